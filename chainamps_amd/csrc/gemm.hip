// Grouped persistent complex GEMM kernels (gfx950 MFMA).
//
// Workgroup = 256 threads = 4 wave64s arranged 2x2; a TM x TN output tile
// per workgroup, each wave owning a (TM/2) x (TN/2) sub-tile made of 16x16
// MFMA fragments.  K is consumed in chunks of KC (16 or 8) staged through LDS
// with real and imaginary parts in separate planes (so one MFMA operand is one
// LDS read); the next chunk is prefetched into registers while the current
// one is multiplied.  A complex product is four real MFMAs,
//   Cre += Are*Bre + (-Aim)*Bim,   Cim += Are*Bim + Aim*Bre,
// or three (Gauss): P1 = Are*Bre, P2 = Aim*Bim, P3 = (Are+Aim)*(Bre+Bim),
//   Cre = P1 - P2,   Cim = P3 - P1 - P2.
//
// MFMA 16x16x4 operand layout (wave64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Accumulator element v of lane l is D[row][l%16] with
// row = l/16 + 4v for the f64 instruction and row = 4(l/16) + v for f32
// (measured on gfx950).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "gemm.hpp"

namespace cam {

namespace {

constexpr int KCmax = 16;  // largest K chunk staged in LDS
constexpr int kXcds = 8;   // MI355X accelerator complex dies (one L2 each)
constexpr int PADL = 4;    // LDS row padding (elements)

// global-address-space access (global_load / global_store rather than
// flat) through native vector types: HIP's double2/float2 have reference
// constructors that would cast the address back to the flat space
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
template <typename T> struct NV;
template <> struct NV<double2> { using T = d2v; };
template <> struct NV<float2> { using T = f2v; };
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  using V = typename NV<T>::T;
  const V v = *(const __attribute__((address_space(1))) V*)(p);
  T r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, const T& a) {
  using V = typename NV<T>::T;
  V v;
  v.x = a.x;
  v.y = a.y;
  *(__attribute__((address_space(1))) V*)(p) = v;
}

// The plan (tiles, problems, segments) is read-only for the whole launch:
// reading it through the constant address space lets the compiler use
// scalar loads, whose waits (lgkmcnt) do not also wait for the operand
// prefetch in flight (vmcnt).
template <typename T>
__device__ __forceinline__ T cload(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "plan records are word-sized");
  using W = const __attribute__((address_space(4))) uint32_t*;
  const W q = (W)(p);
  union {
    T t;
    uint32_t w[sizeof(T) / 4];
  } u;
#pragma unroll
  for (int i = 0; i < static_cast<int>(sizeof(T) / 4); ++i) u.w[i] = q[i];
  return u.t;
}

template <typename R> struct CT;
template <> struct CT<double> { using T = double2; };
template <> struct CT<float> { using T = float2; };

template <typename R> struct Mf;
template <> struct Mf<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int fk, int v) { return fk + 4 * v; }
};
template <> struct Mf<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int fk, int v) { return 4 * fk + v; }
};

// G3: complex products with three real MFMAs (P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi); re = P1 - P2, im = P3 - P1 - P2) instead of four;
// the error stays O(eps |A||B|) per product, as for the four-product form.
// KC = 8 halves the LDS footprint and the prefetch registers (more resident
// workgroups where registers allow) at one barrier per 8 instead of 16 k
template <typename R, int TM, int TN, bool G3, int KC>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const GemmProb* __restrict__ probs,
                                                           const GemmSeg* __restrict__ segs,
                                                           const GemmTile* __restrict__ tiles,
                                                           int ntiles, GemmBases bases) {
  using C = typename CT<R>::T;
  using acc_t = typename Mf<R>::acc;
  constexpr int LDA = TM + PADL;
  constexpr int LDB = TN + PADL;
  constexpr int PA = TM * KC / 256;   // A elements per thread per chunk
  constexpr int PB = TN * KC / 256;   // B elements per thread per chunk
  constexpr int FM = TM / 32;         // 16-row fragments per wave
  constexpr int FN = TN / 32;
  // two LDS buffers: chunk c+1 is stored while chunk c is multiplied, so one
  // barrier per chunk suffices
  __shared__ R sAr[2][KC * LDA], sAi[2][KC * LDA], sBr[2][KC * LDB], sBi[2][KC * LDB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int fr = lane & 15, fk = lane >> 4;

  // XCD-aware order: workgroup b runs on XCD b % 8, so consecutive tiles
  // (the M-tiles of one output column block, which read the same B columns)
  // are given to workgroups of the same XCD and share its L2
  const int G = gridDim.x;
  const int v0 = (G % kXcds == 0) ? (blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds : blockIdx.x;
  for (int t = v0; t < ntiles; t += G) {
    const GemmTile tl = cload(tiles + t);
    const GemmProb P = cload(probs + tl.prob);
    const int M = P.M, N = P.N;
    const int m0 = tl.m0, n0 = tl.n0;

    // four-product form: accr = re, acci = im; G3: accr = P1, acci = P2,
    // accs = P3
    acc_t accr[FM][FN], acci[FM][FN], accs[G3 ? FM : 1][G3 ? FN : 1];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        accr[i][j] = acc_t{0, 0, 0, 0};
        acci[i][j] = acc_t{0, 0, 0, 0};
        if (G3) accs[G3 ? i : 0][G3 ? j : 0] = acc_t{0, 0, 0, 0};
      }

    // prefetched operands of the next chunk (registers) and their in-bounds flags
    C ra0[PA], rb0[PB];
    uint32_t am0 = 0, bm0 = 0;
    const int send = P.seg0 + P.nseg;
    // load cursor: segment and k offset of the next chunk to load (plans
    // hold no empty segments)
    int cs = P.seg0;
    int ck = 0;
    GemmSeg CS;
    if (cs < send) CS = cload(segs + cs);

    // Address state of the current segment, set up once per segment: the
    // operand pointers at this tile's batch, first row / column and the
    // segment's first k, every element's offset from them (32 bits: a tile
    // spans at most 128 rows or columns and the plans keep strides below
    // 2^31 / 128), the element stride along k and which elements lie inside
    // the problem's rows / columns.  A chunk's addresses are then one add per
    // element (the chunk's k offset is uniform).
    const C* segA = nullptr;
    const C* segB = nullptr;
    int32_t aoff[PA], boff[PB];
    int64_t sak = 0, sbk = 0;
    uint32_t arow = 0, bcol = 0;
    bool akc = false, bnc = false;
    auto setup = [&](const GemmSeg& G) {
      const int asp = (G.flags >> 8) & 15, bsp = (G.flags >> 12) & 15;
      akc = G.flags & kSegAKContig;
      bnc = G.flags & kSegBNContig;
      const int32_t csA = static_cast<int32_t>(G.csA), ldB = static_cast<int32_t>(G.ldB);
      segA = (asp ? reinterpret_cast<const C*>(bases.p[asp]) + reinterpret_cast<intptr_t>(G.A)
                  : reinterpret_cast<const C*>(G.A)) +
             static_cast<int64_t>(tl.batch) * G.bsA + (akc ? static_cast<int64_t>(m0) * G.csA : m0);
      segB = (bsp ? reinterpret_cast<const C*>(bases.p[bsp]) + reinterpret_cast<intptr_t>(G.B)
                  : reinterpret_cast<const C*>(G.B)) +
             static_cast<int64_t>(tl.batch) * G.bsB + (bnc ? n0 : static_cast<int64_t>(n0) * G.ldB);
      sak = akc ? 1 : G.csA;
      sbk = bnc ? G.ldB : 1;
      arow = 0;
      bcol = 0;
#pragma unroll
      for (int p = 0; p < PA; ++p) {
        const int e = tid + p * 256;
        const int m = akc ? e / KC : e % TM, k = akc ? e % KC : e / TM;
        aoff[p] = akc ? k + m * csA : m + k * csA;
        arow |= (m0 + m < M ? 1u : 0u) << p;
      }
#pragma unroll
      for (int p = 0; p < PB; ++p) {
        const int e = tid + p * 256;
        const int n = bnc ? e % TN : e / KC, k = bnc ? e / TN : e % KC;
        boff[p] = bnc ? k * ldB + n : k + n * ldB;
        bcol |= (n0 + n < N ? 1u : 0u) << p;
      }
    };
    if (cs < send) setup(CS);
    auto advance = [&]() {
      ck += KC;
      if (ck >= CS.K) {
        ck = 0;
        ++cs;
        if (cs < send) {
          CS = cload(segs + cs);
          setup(CS);
        }
      }
      return cs < send;
    };

    // operand loads of the chunk at k offset kk0 of the current segment:
    // global address space (not flat, whose completion the LDS waits of the
    // MFMA loop would also have to wait for) and branch-free: an element
    // outside the problem reads the segment's first element and is zeroed
    // when it is staged to LDS, so nothing waits on the loads before the
    // MFMAs
    auto load = [&](int kk0, int K, C(&ra)[PA], C(&rb)[PB], uint32_t& amask, uint32_t& bmask) {
      const int64_t da = static_cast<int64_t>(kk0) * sak, db = static_cast<int64_t>(kk0) * sbk;
      const int kleft = K - kk0;
#pragma unroll
      for (int p = 0; p < PA; ++p) {
        const int e = tid + p * 256;
        const int k = akc ? e % KC : e / TM;
        const bool ok = ((arow >> p) & 1u) && k < kleft;
        // a padding element reads the tile's first element of the segment,
        // inside the operand's allocation; the empty asm keeps the compiler
        // from turning the select into two loads under divergent branches
        int64_t o = ok ? aoff[p] + da : 0;
        asm("" : "+v"(o));
        ra[p] = gload(segA + o);
        amask = ok ? (amask | (1u << p)) : (amask & ~(1u << p));
      }
#pragma unroll
      for (int p = 0; p < PB; ++p) {
        const int e = tid + p * 256;
        const int k = bnc ? e / TN : e % KC;
        const bool ok = ((bcol >> p) & 1u) && k < kleft;
        int64_t o = ok ? boff[p] + db : 0;
        asm("" : "+v"(o));
        rb[p] = gload(segB + o);
        bmask = ok ? (bmask | (1u << p)) : (bmask & ~(1u << p));
      }
    };
    auto store = [&](const GemmSeg& G, int buf, const C(&ra)[PA], const C(&rb)[PB], uint32_t amask,
                     uint32_t bmask) {
      R* __restrict__ pAr = sAr[buf];
      R* __restrict__ pAi = sAi[buf];
      R* __restrict__ pBr = sBr[buf];
      R* __restrict__ pBi = sBi[buf];
      const bool cA = G.flags & kSegConjA, cB = G.flags & kSegConjB, al = G.flags & kSegAlpha;
      const R are = static_cast<R>(G.alpha_re), aim = static_cast<R>(G.alpha_im);
      const bool akc = G.flags & kSegAKContig;
#pragma unroll
      for (int p = 0; p < PA; ++p) {
        const int e = tid + p * 256;
        const int m = akc ? e / KC : e % TM, k = akc ? e % KC : e / TM;
        const bool ok = (amask >> p) & 1u;
        R xr = ok ? ra[p].x : R(0), xi = ok ? (cA ? -ra[p].y : ra[p].y) : R(0);
        if (al) {
          const R tr = xr * are - xi * aim;
          xi = xr * aim + xi * are;
          xr = tr;
        }
        pAr[k * LDA + m] = xr;
        pAi[k * LDA + m] = xi;
      }
      if (G.flags & kSegBNContig) {
#pragma unroll
        for (int p = 0; p < PB; ++p) {
          const int e = tid + p * 256;
          const int n = e % TN, k = e / TN;
          const bool ok = (bmask >> p) & 1u;
          pBr[k * LDB + n] = ok ? rb[p].x : R(0);
          pBi[k * LDB + n] = ok ? (cB ? -rb[p].y : rb[p].y) : R(0);
        }
      } else {
#pragma unroll
        for (int p = 0; p < PB; ++p) {
          const int e = tid + p * 256;
          const int k = e % KC, n = e / KC;
          const bool ok = (bmask >> p) & 1u;
          pBr[k * LDB + n] = ok ? rb[p].x : R(0);
          pBi[k * LDB + n] = ok ? (cB ? -rb[p].y : rb[p].y) : R(0);
        }
      }
    };

    auto compute = [&](int buf) {
      const R* __restrict__ cAr = sAr[buf];
      const R* __restrict__ cAi = sAi[buf];
      const R* __restrict__ cBr = sBr[buf];
      const R* __restrict__ cBi = sBi[buf];
      // fragments of step ks+1 are read from LDS before the MFMAs of step ks
      // are issued, so their latency hides behind those MFMAs (tiles up to
      // 64x64; the 128-wide tiles have no registers to spare and read each
      // step's fragments just before its MFMAs)
      constexpr bool kPipe = TM * TN <= 4096;
      R far_[2][FM], fai_[2][FM], fbr_[2][FN], fbi_[2][FN];
      auto frag = [&](int ks, int slot) {
        const int kr = ks * 4 + fk;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          far_[slot][i] = cAr[kr * LDA + wm * (TM / 2) + i * 16 + fr];
          fai_[slot][i] = cAi[kr * LDA + wm * (TM / 2) + i * 16 + fr];
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          fbr_[slot][j] = cBr[kr * LDB + wn * (TN / 2) + j * 16 + fr];
          fbi_[slot][j] = cBi[kr * LDB + wn * (TN / 2) + j * 16 + fr];
        }
      };
      if (kPipe) frag(0, 0);
#pragma unroll
      for (int ks = 0; ks < KC / 4; ++ks) {
        const int cur = kPipe ? (ks & 1) : 0;
        if (!kPipe) frag(ks, 0);
        else if (ks + 1 < KC / 4) frag(ks + 1, cur ^ 1);
        R ar[FM], ai[FM], br[FN], bi[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          ar[i] = far_[cur][i];
          ai[i] = fai_[cur][i];
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          br[j] = fbr_[cur][j];
          bi[j] = fbi_[cur][j];
        }
        if constexpr (G3) {
          // three passes; an accumulator is revisited FM*FN MFMAs later
          R sa[FM], sb[FN];
#pragma unroll
          for (int i = 0; i < FM; ++i) sa[i] = ar[i] + ai[i];
#pragma unroll
          for (int j = 0; j < FN; ++j) sb[j] = br[j] + bi[j];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accr[i][j] = Mf<R>::mma(ar[i], br[j], accr[i][j]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acci[i][j] = Mf<R>::mma(ai[i], bi[j], acci[i][j]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accs[i][j] = Mf<R>::mma(sa[i], sb[j], accs[i][j]);
        } else {
          // four passes over the fragments so that consecutive MFMAs never
          // depend on each other (each accumulator is revisited 2*FM*FN
          // instructions later)
          R nai[FM];
#pragma unroll
          for (int i = 0; i < FM; ++i) nai[i] = -ai[i];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accr[i][j] = Mf<R>::mma(ar[i], br[j], accr[i][j]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acci[i][j] = Mf<R>::mma(ar[i], bi[j], acci[i][j]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) accr[i][j] = Mf<R>::mma(nai[i], bi[j], accr[i][j]);
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acci[i][j] = Mf<R>::mma(ai[i], br[j], acci[i][j]);
        }
      }
    };

    int buf = 0;
    if (cs < send) {
      GemmSeg S0 = CS;
      load(ck, S0.K, ra0, rb0, am0, bm0);
      store(S0, 0, ra0, rb0, am0, bm0);
      __syncthreads();
      // chunk c+1 is loaded into registers while chunk c is multiplied from
      // LDS, then staged into the other LDS buffer.  The load is issued
      // unconditionally (past the last chunk it re-reads the last segment and
      // is never staged), so every path has the same loads in flight.
      for (;;) {
        const bool v1 = advance();
        const GemmSeg S1 = CS;
        load(ck, S1.K, ra0, rb0, am0, bm0);
        // keep every load of the next chunk ahead of the MFMAs (the scheduler
        // would otherwise sink some of them next to their LDS stores, where
        // their whole latency is exposed)
        __builtin_amdgcn_sched_barrier(0);
        compute(buf);
        // the other LDS buffer was last read before the previous barrier
        if (v1) store(S1, buf ^ 1, ra0, rb0, am0, bm0);
        __syncthreads();
        buf ^= 1;
        if (!v1) break;
      }
    } else {
      __syncthreads();
    }

    // epilogue
    C* Cp = (P.cspace ? reinterpret_cast<C*>(const_cast<void*>(bases.p[P.cspace])) +
                            reinterpret_cast<intptr_t>(P.C)
                      : reinterpret_cast<C*>(P.C)) +
            static_cast<int64_t>(tl.batch) * P.bsC;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int n = n0 + wn * (TN / 2) + j * 16 + fr;
        if (n >= N) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int m = m0 + wm * (TM / 2) + i * 16 + Mf<R>::row(fk, v);
          if (m < M) {
            C* dst = Cp + m + static_cast<int64_t>(n) * P.csC;
            C val;
            if constexpr (G3) {
              const R p1 = accr[i][j][v], p2 = acci[i][j][v];
              val.x = p1 - p2;
              val.y = accs[i][j][v] - p1 - p2;
            } else {
              val.x = accr[i][j][v];
              val.y = acci[i][j][v];
            }
            if (P.beta) {
              const C old = gload(dst);
              val.x += old.x;
              val.y += old.y;
            }
            gstore(dst, val);
          }
        }
      }
    __syncthreads();
  }
}

// Skinny problems (N <= kSkinnyN, short segments, A M-contiguous): the W
// stages of small QN sectors, where a 16x16 MFMA fragment would be mostly
// padding.  One thread per output row accumulates its N outputs in
// registers.  Per work item the K x N blocks of B of all the problem's
// segments are staged in LDS once (every thread reads the same element: a
// broadcast), and the segments' k steps are flattened into one table of
// (A column address, B row, operand flags), so the loop keeps eight
// coalesced A loads in flight across segment boundaries with no barrier.
// A is read once per problem and C written once: bound by HBM like the
// copy it nearly is.
template <typename R>
__global__ __launch_bounds__(256) void skinny_gemm_kernel(const GemmProb* __restrict__ probs,
                                                          const GemmSeg* __restrict__ segs,
                                                          const GemmTile* __restrict__ items, int nitems,
                                                          GemmBases bases) {
  using C = typename CT<R>::T;
  __shared__ R sBr[kSkinnyB], sBi[kSkinnyB];
  __shared__ const C* sStepA[kSkinnySteps];    // A column of the step, batch and k applied
  __shared__ int sStepB[kSkinnySteps];         // first element of its B row in sBr / sBi
  __shared__ int sStepF[kSkinnySteps];         // segment flags
  __shared__ R sStepAr[kSkinnySteps], sStepAi[kSkinnySteps];
  __shared__ int sSegK[kSkinnySegs], sSegOff[kSkinnySegs], sSegStep[kSkinnySegs];
  const int tid = threadIdx.x;
  // XCD-aware order, as for the tiles: consecutive items (the problems of
  // one row group over the same rows, which read the same A rows) go to
  // workgroups of one XCD, whose L2 then serves the repeated reads
  const int Gd = gridDim.x;
  const int v0 = (Gd % kXcds == 0) ? (blockIdx.x % kXcds) * (Gd / kXcds) + blockIdx.x / kXcds : blockIdx.x;
  for (int it = v0; it < nitems; it += Gd) {
    const GemmTile w = cload(items + it);
    const GemmProb P = cload(probs + w.prob);
    const int M = P.M, N = P.N, ns = P.nseg;
    __syncthreads();   // the previous item's tables are no longer read
    if (tid < ns) sSegK[tid] = cload(segs + P.seg0 + tid).K;
    __syncthreads();
    if (tid == 0) {
      int off = 0, st = 0;
      for (int j = 0; j < ns; ++j) {
        sSegOff[j] = off;
        sSegStep[j] = st;
        off += sSegK[j] * N;
        st += sSegK[j];
      }
    }
    __syncthreads();
    int nsteps = 0;
    for (int j = 0; j < ns; ++j) nsteps += sSegK[j];
    // B blocks and the step table of every segment
    for (int j = 0; j < ns; ++j) {
      const GemmSeg G = cload(segs + P.seg0 + j);
      const int asp = (G.flags >> 8) & 15, bsp = (G.flags >> 12) & 15;
      const C* Ab = (asp ? reinterpret_cast<const C*>(bases.p[asp]) + reinterpret_cast<intptr_t>(G.A)
                         : reinterpret_cast<const C*>(G.A)) +
                    static_cast<int64_t>(w.batch) * G.bsA;
      const C* Bb = (bsp ? reinterpret_cast<const C*>(bases.p[bsp]) + reinterpret_cast<intptr_t>(G.B)
                         : reinterpret_cast<const C*>(G.B)) +
                    static_cast<int64_t>(w.batch) * G.bsB;
      const bool cB = G.flags & kSegConjB, bnc = G.flags & kSegBNContig;
      const int K = G.K, off = sSegOff[j], st = sSegStep[j];
      for (int e = tid; e < K * N; e += 256) {
        const int k = e / N, n = e % N;
        const C b = gload(Bb + (bnc ? static_cast<int64_t>(k) * G.ldB + n : k + static_cast<int64_t>(n) * G.ldB));
        sBr[off + e] = b.x;
        sBi[off + e] = cB ? -b.y : b.y;
      }
      if (tid < K) {
        sStepA[st + tid] = Ab + static_cast<int64_t>(tid) * G.csA;
        sStepB[st + tid] = off + tid * N;
        sStepF[st + tid] = G.flags;
        sStepAr[st + tid] = static_cast<R>(G.alpha_re);
        sStepAi[st + tid] = static_cast<R>(G.alpha_im);
      }
    }
    __syncthreads();
    const int m = w.m0 + tid;
    if (m >= M) continue;
    R accr[kSkinnyN], acci[kSkinnyN];
#pragma unroll
    for (int n = 0; n < kSkinnyN; ++n) accr[n] = acci[n] = R(0);
    for (int j0 = 0; j0 < nsteps; j0 += 8) {
      C a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = gload(sStepA[j0 + u < nsteps ? j0 + u : j0] + m);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (j0 + u >= nsteps) break;
        const int j = j0 + u, fl = sStepF[j], bo = sStepB[j];
        R xr = a[u].x, xi = (fl & kSegConjA) ? -a[u].y : a[u].y;
        if (fl & kSegAlpha) {
          const R are = sStepAr[j], aim = sStepAi[j];
          const R tr = xr * are - xi * aim;
          xi = xr * aim + xi * are;
          xr = tr;
        }
#pragma unroll
        for (int n = 0; n < kSkinnyN; ++n) {
          if (n >= N) break;
          const R br = sBr[bo + n], bi = sBi[bo + n];
          accr[n] += xr * br - xi * bi;
          acci[n] += xr * bi + xi * br;
        }
      }
    }
    C* Cp = (P.cspace ? reinterpret_cast<C*>(const_cast<void*>(bases.p[P.cspace])) + reinterpret_cast<intptr_t>(P.C)
                      : reinterpret_cast<C*>(P.C)) +
            static_cast<int64_t>(w.batch) * P.bsC + m;
#pragma unroll
    for (int n = 0; n < kSkinnyN; ++n) {
      if (n >= N) break;
      C* dst = Cp + static_cast<int64_t>(n) * P.csC;
      C val;
      val.x = accr[n];
      val.y = acci[n];
      if (P.beta) {
        const C old = gload(dst);
        val.x += old.x;
        val.y += old.y;
      }
      gstore(dst, val);
    }
  }
}

template <typename R>
void launch_skinny(const GemmPlan& plan, const LaunchCfg& c, hipStream_t s, int grid_limit, const GemmBases& bases) {
  const int n = static_cast<int>(plan.skinny.size());
  if (!n) return;
  static int cus = device_cu_count();
  int grid = std::min(n, cus * 8);
  if (grid_limit > 0) grid = std::min(grid, grid_limit);
  (void)c;
  hipLaunchKernelGGL((skinny_gemm_kernel<R>), dim3(grid), dim3(256), 0, s, plan.d_probs.get(), plan.d_segs.get(),
                     plan.d_skinny.get(), n, bases);
  HIP_CHECK(hipGetLastError());
}

template <typename R, int TM, int TN>
void launch_shape(const GemmPlan& plan, const LaunchCfg& c, int shape, hipStream_t s, int grid_limit,
                  const GemmBases& bases) {
  const int nt = static_cast<int>(plan.tiles[c.tile_mode][shape].size());
  if (!nt) return;
  static int cus = device_cu_count();
  int grid = std::min(nt, cus * std::max(1, c.wg_per_cu));
  if (grid_limit > 0) grid = std::min(grid, grid_limit);
  const GemmTile* tiles = plan.d_tiles[c.tile_mode][shape].get();
  const GemmProb* P = plan.d_probs.get();
  const GemmSeg* S = plan.d_segs.get();
  if (c.algo == kAlgo3M && c.kc == 8)
    hipLaunchKernelGGL((grouped_gemm_kernel<R, TM, TN, true, 8>), dim3(grid), dim3(256), 0, s, P, S, tiles, nt, bases);
  else if (c.algo == kAlgo3M)
    hipLaunchKernelGGL((grouped_gemm_kernel<R, TM, TN, true, 16>), dim3(grid), dim3(256), 0, s, P, S, tiles, nt, bases);
  else if (c.kc == 8)
    hipLaunchKernelGGL((grouped_gemm_kernel<R, TM, TN, false, 8>), dim3(grid), dim3(256), 0, s, P, S, tiles, nt, bases);
  else
    hipLaunchKernelGGL((grouped_gemm_kernel<R, TM, TN, false, 16>), dim3(grid), dim3(256), 0, s, P, S, tiles, nt, bases);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

int device_cu_count() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  return prop.multiProcessorCount;
}

int GemmPlan::add_problem(void* C, int64_t csC, int M, int N, int batch, int64_t bsC, int beta,
                          int cspace) {
  GemmProb p{};
  p.C = C;
  p.csC = csC;
  p.bsC = bsC;
  p.M = M;
  p.N = N;
  p.seg0 = static_cast<int32_t>(segs.size());
  p.nseg = 0;
  p.batch = batch;
  p.beta = beta;
  p.tile = 0;
  p.cspace = cspace;
  probs.push_back(p);
  group.push_back(-1);
  uploaded = false;
  for (int m = 0; m < kTileModes; ++m) built[m] = tiles_uploaded[m] = false;
  return static_cast<int>(probs.size()) - 1;
}

void GemmPlan::set_group(int prob, int64_t g) {
  CAM_REQUIRE(prob >= 0 && prob < static_cast<int>(probs.size()), "set_group: no such problem");
  group[prob] = g;
  for (int m = 0; m < kTileModes; ++m) built[m] = tiles_uploaded[m] = false;
}

void GemmPlan::add_segment(int prob, const void* A, int64_t csA, int64_t bsA, const void* B,
                           int64_t ldB, int64_t bsB, int K, int flags, double are, double aim) {
  if (K <= 0) return;   // contributes nothing; the kernels assume K >= 1
  GemmProb& p = probs[prob];
  CAM_REQUIRE(p.seg0 + p.nseg == static_cast<int32_t>(segs.size()),
              "segments of a problem must be added contiguously");
  // the kernels address an element of a tile by a 32-bit offset from the
  // tile's first element: at most 128 rows / columns (and KC <= 16 k) times
  // the stride must stay below 2^31
  constexpr int64_t kOffLimit = (int64_t(1) << 31) - 256;
  CAM_REQUIRE(csA >= 0 && ldB >= 0 && csA * 128 < kOffLimit && ldB * 128 < kOffLimit,
              "add_segment: operand stride too large for 32-bit tile offsets");
  GemmSeg g{};
  g.A = A;
  g.B = B;
  g.csA = csA;
  g.ldB = ldB;
  g.bsA = bsA;
  g.bsB = bsB;
  g.K = K;
  g.flags = flags;
  g.alpha_re = are;
  g.alpha_im = aim;
  if (are != 1.0 || aim != 0.0) g.flags |= kSegAlpha;
  segs.push_back(g);
  p.nseg++;
  flops += 8.0 * double(p.M) * double(p.N) * double(K) * double(p.batch);
  uploaded = false;
}

void GemmPlan::choose_tiles(int mode) {
  CAM_REQUIRE(mode >= 0 && mode < kTileModes, "unknown tile mode");
  cfg[0].tile_mode = cfg[1].tile_mode = mode;
}

void GemmPlan::build_tiles() {
  build_tiles_mode(cfg[0].tile_mode);
  if (cfg[1].tile_mode != cfg[0].tile_mode) build_tiles_mode(cfg[1].tile_mode);
}

static bool skinny_eligible(const GemmPlan& plan, const GemmProb& p) {
  static const bool off = [] {
    const char* e = std::getenv("CAM_SKINNY");
    return e && *e == '0';
  }();
  if (off || p.N > kSkinnyN || p.nseg == 0 || p.nseg > kSkinnySegs) return false;
  int64_t steps = 0;
  for (int s = p.seg0; s < p.seg0 + p.nseg; ++s) {
    const GemmSeg& g = plan.segs[s];
    if (g.K > kSkinnyK || (g.flags & kSegAKContig)) return false;
    steps += g.K;
  }
  // the step table and every B block fit the kernel's LDS tables
  return steps <= kSkinnySteps && steps * p.N <= kSkinnyB;
}

void GemmPlan::build_tiles_mode(int mode) {
  auto& T = tiles[mode];
  for (int t = 0; t < kNumTileShapes; ++t) T[t].clear();
  std::vector<double> cost[kNumTileShapes];
  skinny.clear();
  skinny_uploaded = false;
  for (size_t pi = 0; pi < probs.size(); ++pi) {
    const GemmProb& p = probs[pi];
    if (skinny_eligible(*this, p)) {
      for (int b = 0; b < p.batch; ++b)
        for (int m0 = 0; m0 < p.M; m0 += kSkinnyRows)
          skinny.push_back(GemmTile{static_cast<int32_t>(pi), b, m0, 0});
      continue;
    }
    int shape = 0;
    if (mode != 0) {
      double best = 1e300;
      const int t_end = mode == 3 ? kNumTileShapes : kNumSmallShapes;
      for (int t = (mode == 2 ? 1 : 0); t < t_end; ++t) {
        const double tm = kTileShapes[t].tm, tn = kTileShapes[t].tn;
        const double pm = std::ceil(p.M / tm) * tm, pn = std::ceil(p.N / tn) * tn;
        // padded area, with a small per-tile overhead so bigger tiles win ties
        const double c = pm * pn * (1.0 + 0.02 * (4096.0 / (tm * tn)));
        if (c < best) { best = c; shape = t; }
      }
    }
    int64_t ktot = 0;
    for (int s = p.seg0; s < p.seg0 + p.nseg; ++s) ktot += (segs[s].K + KCmax - 1) / KCmax * KCmax;
    if (ktot == 0 && p.beta) continue;   // nothing to add
    const int tm = kTileShapes[shape].tm, tn = kTileShapes[shape].tn;
    for (int b = 0; b < p.batch; ++b)
      for (int n0 = 0; n0 < p.N; n0 += tn)
        for (int m0 = 0; m0 < p.M; m0 += tm) {
          T[shape].push_back(GemmTile{static_cast<int32_t>(pi), b, m0, n0});
          cost[shape].push_back(double(ktot) + 8.0);
        }
  }
  // Longest first (LPT) for the persistent round-robin, by cost class
  // (within 25 %), and inside a class the tiles that read the same B
  // columns (problems whose first segment has the same B operand, e.g. the
  // three MPO channels applied to one psi block) next to each other, M-tiles
  // innermost: consecutive tiles run at the same time on one XCD, so B is
  // fetched from HBM once and re-read from that XCD's L2.  A plan with row
  // groups orders by (group, batch, m0) first instead: the tiles that read
  // the same A rows run together (such plans have far more tiles than the
  // grid, so the cost order matters less than the shared fetch).
  const bool grouped = std::any_of(group.begin(), group.end(), [](int64_t g) { return g >= 0; });
  auto bkey = [&](const GemmTile& g) -> uint64_t {
    const GemmProb& p = probs[g.prob];
    if (!p.nseg) return 0;
    const GemmSeg& q = segs[p.seg0];
    return (static_cast<uint64_t>((q.flags >> 12) & 15) << 56) ^ reinterpret_cast<uint64_t>(q.B) ^
           (static_cast<uint64_t>(g.batch) << 40);
  };
  for (int t = 0; t < kNumTileShapes; ++t) {
    std::vector<size_t> ord(T[t].size());
    std::iota(ord.begin(), ord.end(), 0);
    std::vector<int> cls(T[t].size());
    std::vector<uint64_t> bk(T[t].size());
    for (size_t i = 0; i < ord.size(); ++i) {
      cls[i] = static_cast<int>(std::floor(std::log(cost[t][i]) / std::log(1.25)));
      bk[i] = bkey(T[t][i]);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
      if (grouped) {
        const GemmTile &x = T[t][a], &y = T[t][b];
        const int64_t gx = group[x.prob], gy = group[y.prob];
        if (gx != gy) return gx < gy;
        if (x.batch != y.batch) return x.batch < y.batch;
        if (x.m0 != y.m0) return x.m0 < y.m0;
        if (x.prob != y.prob) return x.prob < y.prob;
        return x.n0 < y.n0;
      }
      if (cls[a] != cls[b]) return cls[a] > cls[b];
      if (bk[a] != bk[b]) return bk[a] < bk[b];
      if (T[t][a].n0 != T[t][b].n0) return T[t][a].n0 < T[t][b].n0;
      if (T[t][a].prob != T[t][b].prob) return T[t][a].prob < T[t][b].prob;
      return T[t][a].m0 < T[t][b].m0;
    });
    std::vector<GemmTile> sorted;
    sorted.reserve(ord.size());
    for (auto o : ord) sorted.push_back(T[t][o]);
    T[t].swap(sorted);
  }
  // skinny items: the problems of one row group (same A rows) next to each
  // other at equal (batch, m0), so that they run together on one XCD
  std::stable_sort(skinny.begin(), skinny.end(), [&](const GemmTile& x, const GemmTile& y) {
    const int64_t gx = group[x.prob] >= 0 ? group[x.prob] : -1 - x.prob;
    const int64_t gy = group[y.prob] >= 0 ? group[y.prob] : -1 - y.prob;
    if (gx != gy) return gx < gy;
    if (x.batch != y.batch) return x.batch < y.batch;
    if (x.m0 != y.m0) return x.m0 < y.m0;
    return x.prob < y.prob;
  });
  built[mode] = true;
  tiles_uploaded[mode] = false;
}

void GemmPlan::upload(hipStream_t s) {
  if (!uploaded) {
    d_probs.ensure(std::max<size_t>(probs.size(), 1));
    d_probs.upload(probs.data(), probs.size(), s);
    d_segs.ensure(std::max<size_t>(segs.size(), 1));
    d_segs.upload(segs.data(), segs.size(), s);
    uploaded = true;
  }
  if (!skinny_uploaded) {
    d_skinny.ensure(std::max<size_t>(skinny.size(), 1));
    d_skinny.upload(skinny.data(), skinny.size(), s);
    skinny_uploaded = true;
  }
  for (int m = 0; m < kTileModes; ++m) {
    if (!built[m] || tiles_uploaded[m]) continue;
    for (int t = 0; t < kNumTileShapes; ++t) {
      d_tiles[m][t].ensure(std::max<size_t>(tiles[m][t].size(), 1));
      d_tiles[m][t].upload(tiles[m][t].data(), tiles[m][t].size(), s);
    }
    tiles_uploaded[m] = true;
  }
}

void GemmPlan::clear() {
  probs.clear();
  segs.clear();
  group.clear();
  skinny.clear();
  skinny_uploaded = false;
  for (int m = 0; m < kTileModes; ++m) {
    for (auto& t : tiles[m]) t.clear();
    built[m] = false;
    tiles_uploaded[m] = false;
  }
  flops = 0.0;
  uploaded = false;
  tuned[0] = tuned[1] = false;
}

uint64_t GemmPlan::signature(int precision) const {
  auto lg = [](int64_t v) { int b = 0; while ((int64_t(1) << b) < v) ++b; return b; };
  std::map<std::array<int, 4>, double> hist;
  double tot = 0.0;
  for (const GemmProb& p : probs) {
    int64_t k = 0;
    for (int s = p.seg0; s < p.seg0 + p.nseg; ++s) k += segs[s].K;
    const double f = double(p.M) * double(p.N) * double(k) * double(p.batch);
    hist[{lg(p.M), lg(p.N), lg(std::max<int64_t>(k, 1)), p.batch > 1 ? 1 : 0}] += f;
    tot += f;
  }
  uint64_t h = 1469598103934665603ull ^ static_cast<uint64_t>(precision);
  auto mix = [&h](uint64_t v) { h ^= v; h *= 1099511628211ull; };
  for (auto& kv : hist) {
    const int q = tot > 0 ? static_cast<int>(std::lround(20.0 * kv.second / tot)) : 0;
    if (!q) continue;
    for (int v : kv.first) mix(static_cast<uint64_t>(v));
    mix(static_cast<uint64_t>(q));
  }
  return h;
}

namespace {

std::mutex g_tune_mu;
std::unordered_map<uint64_t, LaunchCfg> g_tune_cache;

void apply_choice(GemmPlan& plan, int precision, const LaunchCfg& c, hipStream_t s) {
  plan.cfg[precision] = c;
  if (!plan.built[c.tile_mode]) plan.build_tiles_mode(c.tile_mode);
  plan.upload(s);
}

}  // namespace

bool autotune(GemmPlan& plan, int precision, hipStream_t s, const GemmBases* bases, double min_flops) {
  if (plan.tuned[precision]) return true;
  const char* e = std::getenv("CAM_AUTOTUNE");
  if ((e && *e == '0') || plan.flops < min_flops) return false;
  for (const GemmProb& p : plan.probs)
    CAM_REQUIRE(!p.beta, "autotune: a plan that accumulates into its output cannot be re-run");
  const uint64_t key = plan.signature(precision);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      apply_choice(plan, precision, it->second, s);
      plan.tuned[precision] = true;
      return true;
    }
  }
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  auto time_it = [&](const LaunchCfg& c) {
    apply_choice(plan, precision, c, s);
    launch_grouped_gemm(plan, precision, s, bases);   // warm
    float best = 1e30f;
    for (int r = 0; r < 2; ++r) {
      HIP_CHECK(hipEventRecord(ev0, s));
      launch_grouped_gemm(plan, precision, s, bases);
      HIP_CHECK(hipEventRecord(ev1, s));
      HIP_CHECK(hipEventSynchronize(ev1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
      best = std::min(best, ms);
    }
    return best;
  };
  // coordinate descent: product form, then tile policy, then K chunk, then
  // grid size
  LaunchCfg cur = plan.cfg[precision];
  float tcur = time_it(cur);
  auto try_axis = [&](int LaunchCfg::*field, std::initializer_list<int> vals) {
    for (int v : vals) {
      if (cur.*field == v) continue;
      LaunchCfg c = cur;
      c.*field = v;
      const float t = time_it(c);
      if (t < tcur) { tcur = t; cur = c; }
    }
  };
  try_axis(&LaunchCfg::algo, {kAlgo3M, kAlgo4M});
  try_axis(&LaunchCfg::tile_mode, {1, 0, 2, 3});
  try_axis(&LaunchCfg::kc, {16, 8});
  try_axis(&LaunchCfg::wg_per_cu, {4, 2, 8});
  apply_choice(plan, precision, cur, s);
  // leave the output of the chosen configuration (every candidate computes
  // the same values; this keeps the last write consistent with the choice)
  launch_grouped_gemm(plan, precision, s, bases);
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = cur;
  }
  plan.tuned[precision] = true;
  return true;
}

void launch_grouped_gemm(GemmPlan& plan, int precision, hipStream_t s, const GemmBases* bases,
                         int grid_limit) {
  const LaunchCfg& c = plan.cfg[precision];
  if (!plan.built[c.tile_mode]) plan.build_tiles_mode(c.tile_mode);
  if (!plan.uploaded || !plan.tiles_uploaded[c.tile_mode]) plan.upload(s);
  GemmBases b{};
  if (bases) b = *bases;
  if (precision == 0) {
    launch_shape<double, 64, 64>(plan, c, 0, s, grid_limit, b);
    launch_shape<double, 32, 64>(plan, c, 1, s, grid_limit, b);
    launch_shape<double, 64, 32>(plan, c, 2, s, grid_limit, b);
    launch_shape<double, 128, 64>(plan, c, 3, s, grid_limit, b);
    launch_shape<double, 64, 128>(plan, c, 4, s, grid_limit, b);
    launch_shape<double, 128, 32>(plan, c, 5, s, grid_limit, b);
    launch_skinny<double>(plan, c, s, grid_limit, b);
  } else {
    launch_shape<float, 64, 64>(plan, c, 0, s, grid_limit, b);
    launch_shape<float, 32, 64>(plan, c, 1, s, grid_limit, b);
    launch_shape<float, 64, 32>(plan, c, 2, s, grid_limit, b);
    launch_shape<float, 128, 64>(plan, c, 3, s, grid_limit, b);
    launch_shape<float, 64, 128>(plan, c, 4, s, grid_limit, b);
    launch_shape<float, 128, 32>(plan, c, 5, s, grid_limit, b);
    launch_skinny<float>(plan, c, s, grid_limit, b);
  }
}

}  // namespace cam
