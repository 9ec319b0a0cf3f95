// Two-vertex effective Hamiltonian H_eff = L . W . W . R on QN blocks.
//
// Vector layout ("block layout"): psi[sigma1, a0, sigma2, a2] is stored as
// blocks keyed (s0, s3, s1, s2) -- sectors of a0, a2, sigma1, sigma2 with
// q(s1)+q(s0)+q(s2) == q(s3) -- each block dense over (a0, a2, sigma1,
// sigma2), a0 fastest.  The matvec is four grouped MFMA launches:
//   A: T1[bL](a0',a2,s1,s2)  = sum_a0       L_bL(a0',a0)         psi
//   B: T2[b1](a0',a2,s1',s2) = sum_{bL,s1}  W(s1',bL,s1,b1)      T1
//   C: T3[bR](a0',a2,s1',s2')= sum_{b1,s2}  W(s2',b1,s2,bR)      T2
//   D: out(a0',a2',s1',s2')  = sum_{bR,a2}  R_bR(a2',a2)         T3
// (the optimal order; the reference's LeftPart/RightPart order costs a
// factor ~d more, DMRG_routines.cpp:779-833, 898-974).
#pragma once
#include <array>
#include <map>
#include <vector>

#include "ctx.hpp"

namespace cam {

// operand spaces of the stage plans
enum HeffSpace { kSpL = 1, kSpW, kSpR, kSpT1, kSpT2, kSpT3, kSpIn, kSpOut };

struct EnvMat {            // strided view of L_b or R_b for one sector pair
  int sout;                // output sector (a0' or a2')
  int64_t off;             // element offset in the tensor data
  int64_t ld;              // m_out * nb   (column stride over the inner index)
  int64_t mout, min;
};

struct WMat {              // W(s', bl, s, br) for fixed (bl, br) and sectors
  int64_t off, ld;         // W(s', s) at off + s' + s*ld
  int64_t nout, nin;
};

struct Heff {
  Ctx* ctx = nullptr;
  TensorPtr L, W, R;       // keep the inputs alive (shared)
  Rules rules;
  QIndex sig, a0, a2;      // sigma (W idx 2), a0 (L idx 5), a2 (R idx 4)
  int64_t d = 0, chiL = 0, chiR = 0, w = 0;

  struct PB { int s0, s3, s1, s2; int64_t m0, m3, n1, n2, off, size; };
  std::vector<PB> pblocks;
  std::map<std::array<int, 4>, int> pindex;   // (s0,s3,s1,s2) -> block
  int64_t n = 0;

  // operand views: L (b, s0) -> a0' view; R (b, s3) -> a2' view;
  // W (bl, br, s', s) -> view
  std::map<std::pair<int64_t, int>, EnvMat> Lv, Rv;
  std::map<std::array<int64_t, 4>, WMat> Wv;
  // intermediates of the whole matvec: key (b, s0', s3, sx, sy) -> offset;
  // sizes (the layout the adjoint works on)
  struct TB { int64_t off, m0, m3, n1, n2; };
  std::map<std::array<int, 5>, TB> t1, t2, t3;
  int64_t t1n = 0, t2n = 0, t3n = 0;
  // the block products of the four stages (independent of slicing)
  struct SegA { int pb; EnvMat L; };
  struct SegW { std::array<int, 5> src; WMat W; };
  struct SegR { std::array<int, 5> src; EnvMat R; };
  std::map<std::array<int, 5>, SegA> segA;
  std::map<std::array<int, 5>, std::vector<SegW>> segB, segC;
  std::map<int, std::vector<SegR>> segD;

  // Slicing: the a0' rows of every sector are cut into `rslices` ranges and
  // the a2 columns of every sector into `cslices`; slice (r, c) runs all four
  // stages on intermediates restricted to those rows and columns (the
  // stages act on a0' and a2 independently until stage D contracts a2, which
  // accumulates the column slices into the output).  Intermediates then
  // take 1/(rslices cslices) of the memory, at the cost of reading psi
  // rslices times and the output cslices times.  1 x 1 is the plain matvec.
  int rslices = 1, cslices = 1;
  int64_t ws1n = 0, ws2n = 0;                  // workspace elements (max over slices)
  std::vector<std::array<GemmPlan, 4>> plans;  // per slice (r-major), stages A..D

  // Two intermediate buffers: stage C writes T3 over T1 (stage D reads only
  // T3), and an FP32 matvec uses the same bytes as float2 (half the size),
  // so a matvec needs 2 x max(T) x 16 bytes whatever the precision.
  DevView<cplx> T1, T2;                        // context workspaces; T1 also holds T3
  DevBuf<cplx> vin, vout;
  DevBuf<float2> L32, W32, R32, vinf, voutf;   // FP32 operands
  bool fp32_ready = false;
  std::vector<int> untouched;   // psi blocks no stage-D problem writes
  double flops = 0.0;
  double stage_flops[4] = {0, 0, 0, 0};

  // reference compressed order map (built on demand)
  std::vector<int64_t> allowed;   // flat positions
  DevBuf<int64_t> perm;           // perm[i] = block-layout offset of allowed[i]
  bool perm_ready = false;

  // workspace_limit: bytes the two intermediate buffers may take (< 0:
  // the device's free memory plus what the context workspaces already
  // hold, less a reserve); the fewest slices that fit are used
  void build(Ctx& c, TensorPtr L_, TensorPtr W_, TensorPtr R_, double workspace_limit = -1.0);
  // (re)plans the stages for rs x cs slices and binds the workspaces
  void plan_slices(int rs, int cs);
  // bytes of the two intermediate buffers for rs x cs slices
  double slice_workspace_bytes(int rs, int cs) const;
  // the fewest slices (r x c, powers of two) whose workspace fits `bytes`
  void choose_slices(double bytes);
  int nslices() const { return rslices * cslices; }
  void build_perm();
  void prepare_fp32();
  // operand bases of a launch (re-binds T1/T2 to the context workspaces)
  GemmBases bases(int precision, const void* in, void* out);
  void bind_workspace(int precision = 0);
  // device block-layout vectors
  void apply_blocked(const cplx* in, cplx* out);
  // the four stage launches (autotuning them on first use outside capture)
  void run_stages(int precision, const GemmBases& b);
  bool tuned(int precision) const;
  // fixes the launch configuration of every stage (no autotuning)
  void set_launch(int precision, const LaunchCfg& c);
  // FP64 stages A..(nstages-1) only: leaves T1 (and T2) of `in` for the adjoint
  void forward_stages(const cplx* in, int nstages);
  void apply_blocked32(const float2* in, float2* out);
  // precision: 0 fp64, 1 fp32 (input/output FP64 block layout)
  void apply(const cplx* in, cplx* out, int precision);
  // psi as the 4-index wavefunction [s1(in), a0(in), s2(in), a2(out)] with
  // two row indices (TwoVertexWavefunction, DMRG_routines.cpp:765-766)
  TensorPtr tensor_from_vector(const cplx* x);
  // block-layout vector from such a tensor; blocks it lacks are zero
  void vector_from_tensor(const BSTensor& T, cplx* x);
  // average ms per stage over `reps` matvecs (HIP events on the stream),
  // on the given device block-layout vectors (FP64, or FP32 split complex
  // for precision 1) or, when null, on internal ones
  void profile(int precision, int reps, double* stage_ms, const void* in_dev = nullptr, void* out_dev = nullptr);

  // Excited-state projector terms of TwoVertexMPOMPSMultiply
  // (DMRG_routines.cpp:947-973): out += weight_p conj(P_p) (P_p . in), the
  // product without conjugation.  Stored as Q_p = conj(P_p) in the block
  // layout (n x nproj), so that c = Q^H in and out -= Q (-weight c).
  int nproj = 0;
  DevBuf<cplx> projQ, projW, projC, projPart;
  DevBuf<float2> projQ32;
  void set_projectors(int np, const cplx* const* P_ref, const double* weights);
  void add_projectors(const cplx* in, cplx* out);
  void add_projectors32(const float2* in, float2* out);
};

}  // namespace cam
