#include <algorithm>

#include "kernels.hpp"

namespace cam {

namespace {

constexpr int kChunk = 4096;   // elements per permutation chunk

__global__ __launch_bounds__(256) void permute_kernel(const PermJob* __restrict__ jobs,
                                                      const int64_t* __restrict__ chunks,
                                                      int nchunks) {
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int64_t code = chunks[c];
    const int j = static_cast<int>(code >> 40);
    const int64_t start = (code & ((int64_t(1) << 40) - 1)) * kChunk;
    const PermJob& J = jobs[j];
    const int64_t end = min(J.size, start + kChunk);
    for (int64_t e = start + threadIdx.x; e < end; e += blockDim.x) {
      int64_t rem = e, off = 0;
      for (int d = 0; d < J.ndim; ++d) {
        const int64_t q = rem / J.odims[d];
        off += (rem - q * J.odims[d]) * J.istr[d];
        rem = q;
      }
      double2 v = reinterpret_cast<const double2*>(J.in)[off];
      if (J.conj) v.y = -v.y;
      reinterpret_cast<double2*>(J.out)[e] = v;
    }
  }
}

__global__ void scatter_map_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                   const int64_t* __restrict__ perm, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    out[perm[i]] = in[i];
}

__global__ void gather_map_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                  const int64_t* __restrict__ perm, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = in[perm[i]];
}

__global__ void to_fp32_kernel(const double2* __restrict__ in, float2* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double2 v = in[i];
    out[i] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
  }
}

__global__ void to_fp64_kernel(const float2* __restrict__ in, double2* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float2 v = in[i];
    out[i] = make_double2(v.x, v.y);
  }
}

__global__ void zero_kernel(uint4* p, int64_t n16) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n16;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void zero_bytes_kernel(unsigned char* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

__global__ void axpy_kernel(double ar, double ai, const double2* __restrict__ x,
                            double2* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double2 a = x[i];
    double2 b = y[i];
    b.x += ar * a.x - ai * a.y;
    b.y += ar * a.y + ai * a.x;
    y[i] = b;
  }
}

__global__ void scale_kernel(double ar, double ai, double2* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double2 a = x[i];
    x[i] = make_double2(ar * a.x - ai * a.y, ar * a.y + ai * a.x);
  }
}

__global__ void copy_kernel(const double2* __restrict__ x, double2* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = x[i];
}

constexpr int kMaxK = 64;
// 8 workgroups of 256 per CU on 256 CUs: the sweeps keep the most loads in
// flight at full occupancy (at 1024 the FP32 dots streamed 4.0 TB/s)
constexpr int kDotBlocks = 2048;

__device__ __forceinline__ double2 ld2(const double2* p) { return *p; }
__device__ __forceinline__ double2 ld2(const float2* p) {
  const float2 v = *p;
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2(double2* p, double2 v) { *p = v; }
__device__ __forceinline__ void st2(float2* p, double2 v) { *p = make_float2(static_cast<float>(v.x), static_cast<float>(v.y)); }

// Each workgroup reduces a contiguous slice of rows for all k columns (FP64
// accumulation whatever the storage of V and w).
template <typename TV, typename TW>
__global__ __launch_bounds__(256) void multi_dot_kernel(const TV* __restrict__ V, int64_t ldv,
                                                        int k, const TW* __restrict__ w,
                                                        int64_t n, double2* __restrict__ partial) {
  __shared__ double red_r[kMaxK][8];
  __shared__ double red_i[kMaxK][8];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = blockIdx.x * per;
  const int64_t hi = min(n, lo + per);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = 0; j0 < k; j0 += 8) {
    double sr[8], si[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) { sr[q] = 0; si[q] = 0; }
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      const double2 b = ld2(w + i);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (j0 + q < k) {
          const double2 a = ld2(V + i + (j0 + q) * ldv);
          sr[q] += a.x * b.x + a.y * b.y;
          si[q] += a.x * b.y - a.y * b.x;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double r = sr[q], im = si[q];
      for (int o = 32; o > 0; o >>= 1) {
        r += __shfl_down(r, o, 64);
        im += __shfl_down(im, o, 64);
      }
      if (lane == 0 && j0 + q < k) {
        red_r[j0 + q][wave] = r;
        red_i[j0 + q][wave] = im;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double r = 0, im = 0;
    for (int q = 0; q < static_cast<int>(blockDim.x >> 6); ++q) {
      r += red_r[threadIdx.x][q];
      im += red_i[threadIdx.x][q];
    }
    partial[blockIdx.x * int64_t(k) + threadIdx.x] = make_double2(r, im);
  }
}

__global__ void multi_dot_final(const double2* __restrict__ partial, int nb, int k,
                                double2* __restrict__ c) {
  const int j = blockIdx.x;
  double r = 0, im = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const double2 p = partial[b * int64_t(k) + j];
    r += p.x;
    im += p.y;
  }
  __shared__ double sr[256], si[256];
  sr[threadIdx.x] = r;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) c[j] = make_double2(sr[0], si[0]);
}

// mode 1: w -= V c; mode 0: w = V c; mode 2: w += V c (FP64 arithmetic)
template <typename TV, typename TW>
__global__ void multi_axpy_kernel(const TV* __restrict__ V, int64_t ldv, int k,
                                  const double2* __restrict__ c, TW* __restrict__ w, int64_t n,
                                  int mode) {
  __shared__ double2 cc[kMaxK];
  if (threadIdx.x < k) cc[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double r = 0, im = 0;
    // columns in chunks of eight: the eight loads of a chunk are issued
    // before any of them is used (the guard is uniform across the wave)
    for (int j0 = 0; j0 < k; j0 += 8) {
      double2 a[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (j0 + q < k) a[q] = ld2(V + i + (j0 + q) * ldv);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (j0 + q < k) {
          const double2 b = cc[j0 + q];
          r += a[q].x * b.x - a[q].y * b.y;
          im += a[q].x * b.y + a[q].y * b.x;
        }
      }
    }
    if (mode) {
      double2 o = ld2(w + i);
      const double sgn = mode == 1 ? -1.0 : 1.0;
      o.x += sgn * r;
      o.y += sgn * im;
      st2(w + i, o);
    } else {
      st2(w + i, make_double2(r, im));
    }
  }
}

inline int grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g, 8192)));
}

__global__ void random_fill_kernel(double2* x, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double a = (z >> 11) * (1.0 / 9007199254740992.0);
    z = z * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z ^= z >> 29;
    const double b = (z >> 11) * (1.0 / 9007199254740992.0);
    x[i] = make_double2(2.0 * a - 1.0, 2.0 * b - 1.0);
  }
}

}  // namespace

void launch_random_fill(cplx* x, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(random_fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, reinterpret_cast<double2*>(x), n,
                     seed);
  HIP_CHECK(hipGetLastError());
}

void launch_permute(const std::vector<PermJob>& jobs, DevBuf<PermJob>& djobs, DevBuf<int64_t>& dchunks,
                    hipStream_t s) {
  if (jobs.empty()) return;
  std::vector<int64_t> chunks;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const int64_t nc = (jobs[j].size + kChunk - 1) / kChunk;
    for (int64_t c = 0; c < nc; ++c) chunks.push_back((int64_t(j) << 40) | c);
  }
  if (chunks.empty()) return;
  djobs.ensure(jobs.size());
  djobs.upload(jobs.data(), jobs.size(), s);
  dchunks.ensure(chunks.size());
  dchunks.upload(chunks.data(), chunks.size(), s);
  const int grid = static_cast<int>(std::min<size_t>(chunks.size(), 4096));
  hipLaunchKernelGGL(permute_kernel, dim3(grid), dim3(256), 0, s, djobs.get(), dchunks.get(),
                     static_cast<int>(chunks.size()));
  HIP_CHECK(hipGetLastError());
  // jobs/chunks buffers are reused by later calls: finish before returning
  HIP_CHECK(hipStreamSynchronize(s));
}

void launch_scatter_map(const cplx* in, cplx* out, const int64_t* perm, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scatter_map_kernel, dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out), perm, n);
  HIP_CHECK(hipGetLastError());
}

void launch_gather_map(const cplx* in, cplx* out, const int64_t* perm, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gather_map_kernel, dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out), perm, n);
  HIP_CHECK(hipGetLastError());
}

void launch_to_fp32(const cplx* in, float2* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_fp32_kernel, dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(in), out, n);
  HIP_CHECK(hipGetLastError());
}

void launch_to_fp64(const float2* in, cplx* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(to_fp64_kernel, dim3(grid_for(n)), dim3(256), 0, s, in,
                     reinterpret_cast<double2*>(out), n);
  HIP_CHECK(hipGetLastError());
}

void launch_zero(void* p, int64_t bytes, hipStream_t s) {
  if (bytes <= 0) return;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0 && (bytes & 15) == 0) {
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(bytes / 16)), dim3(256), 0, s,
                       reinterpret_cast<uint4*>(p), bytes / 16);
  } else {
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(grid_for(bytes)), dim3(256), 0, s,
                       reinterpret_cast<unsigned char*>(p), bytes);
  }
  HIP_CHECK(hipGetLastError());
}

void launch_axpy(cplx a, const cplx* x, cplx* y, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, s, a.re, a.im,
                     reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), n);
  HIP_CHECK(hipGetLastError());
}

void launch_scale(cplx a, cplx* x, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, s, a.re, a.im,
                     reinterpret_cast<double2*>(x), n);
  HIP_CHECK(hipGetLastError());
}

void launch_copy(const cplx* x, cplx* y, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), n);
  HIP_CHECK(hipGetLastError());
}

int multi_dot_blocks(int64_t n) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(kDotBlocks, (n + 1023) / 1024)));
}

void launch_multi_dot(const cplx* V, int64_t ldv, int k, const cplx* w, int64_t n, cplx* partial,
                      cplx* c, hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "multi_dot: k out of range");
  const int nb = multi_dot_blocks(n);
  hipLaunchKernelGGL((multi_dot_kernel<double2, double2>), dim3(nb), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), ldv, k,
                     reinterpret_cast<const double2*>(w), n, reinterpret_cast<double2*>(partial));
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(multi_dot_final, dim3(k), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(partial), nb, k, reinterpret_cast<double2*>(c));
  HIP_CHECK(hipGetLastError());
}

void launch_multi_axpy(const cplx* V, int64_t ldv, int k, const cplx* c, cplx* w, int64_t n,
                       hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "multi_axpy: k out of range");
  hipLaunchKernelGGL((multi_axpy_kernel<double2, double2>), dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), ldv, k,
                     reinterpret_cast<const double2*>(c), reinterpret_cast<double2*>(w), n, 1);
  HIP_CHECK(hipGetLastError());
}

void launch_combine(const cplx* V, int64_t ldv, int k, const cplx* c, cplx* w, int64_t n,
                    hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "combine: k out of range");
  hipLaunchKernelGGL((multi_axpy_kernel<double2, double2>), dim3(grid_for(n)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(V), ldv, k,
                     reinterpret_cast<const double2*>(c), reinterpret_cast<double2*>(w), n, 0);
  HIP_CHECK(hipGetLastError());
}

namespace {
__global__ void coef_scale_kernel(double2* __restrict__ c, const double2* __restrict__ w, int k) {
  const int i = threadIdx.x;
  if (i < k) {
    const double2 a = c[i], b = w[i];
    c[i] = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
}
}  // namespace

void launch_coef_scale(cplx* c, const cplx* w, int k, hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "coef_scale: k out of range");
  hipLaunchKernelGGL(coef_scale_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<double2*>(c),
                     reinterpret_cast<const double2*>(w), k);
  HIP_CHECK(hipGetLastError());
}

void launch_multi_dot32(const float2* V, int64_t ldv, int k, const float2* w, int64_t n, cplx* partial, cplx* c,
                        hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "multi_dot: k out of range");
  const int nb = multi_dot_blocks(n);
  hipLaunchKernelGGL((multi_dot_kernel<float2, float2>), dim3(nb), dim3(256), 0, s, V, ldv, k, w, n,
                     reinterpret_cast<double2*>(partial));
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(multi_dot_final, dim3(k), dim3(256), 0, s, reinterpret_cast<const double2*>(partial), nb, k,
                     reinterpret_cast<double2*>(c));
  HIP_CHECK(hipGetLastError());
}

void launch_multi_axpy32(const float2* V, int64_t ldv, int k, const cplx* c, float2* w, int64_t n, hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "multi_axpy: k out of range");
  hipLaunchKernelGGL((multi_axpy_kernel<float2, float2>), dim3(grid_for(n)), dim3(256), 0, s, V, ldv, k,
                     reinterpret_cast<const double2*>(c), w, n, 1);
  HIP_CHECK(hipGetLastError());
}

void launch_accumulate32(const float2* V, int64_t ldv, int k, const cplx* c, cplx* w, int64_t n, hipStream_t s) {
  CAM_REQUIRE(k >= 1 && k <= kMaxK, "accumulate: k out of range");
  hipLaunchKernelGGL((multi_axpy_kernel<float2, double2>), dim3(grid_for(n)), dim3(256), 0, s, V, ldv, k,
                     reinterpret_cast<const double2*>(c), reinterpret_cast<double2*>(w), n, 2);
  HIP_CHECK(hipGetLastError());
}

}  // namespace cam
