// C ABI (include/chainamps_amd.h).  Every entry point converts exceptions
// into a non-zero status plus cam_last_error() text.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/chainamps_amd.h"
#include "ctx.hpp"
#include "heff.hpp"
#include "solvers.hpp"
#include "idmrg.hpp"
#include "itebd.hpp"
#include "ftebd.hpp"
#include "storage.hpp"
#include "ortho.hpp"
#include "noalloc.hpp"
#include <memory>

using namespace cam;

struct cam_ctx_s {
  Ctx ctx;
  explicit cam_ctx_s(int d) : ctx(d) {}
};
struct cam_tensor_s {
  TensorPtr t;
  cam_ctx* owner;
};
struct cam_heff_s {
  Heff h;
  cam_ctx* owner;
  DevBuf<cplx> bin, bout, rin, rout;
};

namespace cam {
static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }
const char* last_error() { return g_last_error.c_str(); }
}  // namespace cam

#define CAM_API_BEGIN try {
#define CAM_API_END                                   \
  return 0;                                           \
  }                                                   \
  catch (const std::exception& e) {                   \
    cam::set_last_error(e.what());                    \
    return 1;                                         \
  }                                                   \
  catch (...) {                                       \
    cam::set_last_error("unknown error");             \
    return 1;                                         \
  }

static void set_device(cam_ctx* c) { HIP_CHECK(hipSetDevice(c->ctx.device)); }

extern "C" {

const char* cam_last_error(void) { return cam::last_error(); }
const char* cam_version(void) { return "chainamps_amd 0.1 (gfx950)"; }

int cam_ctx_create(int device, cam_ctx** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(out, "null output");
  int n = 0;
  HIP_CHECK(hipGetDeviceCount(&n));
  CAM_REQUIRE(device >= 0 && device < n, "no such HIP device");
  auto* c = new cam_ctx(device);
  try {
    // the no-allocation matvec's scratch array on this device (the
    // context constructor made the device current)
    noalloc_setup(device);
  } catch (...) {
    delete c;
    throw;
  }
  *out = c;
  CAM_API_END
}

int cam_ctx_destroy(cam_ctx* c) {
  CAM_API_BEGIN
  delete c;
  CAM_API_END
}

int cam_ctx_sync(cam_ctx* c) {
  CAM_API_BEGIN
  HIP_CHECK(hipStreamSynchronize(c->ctx.stream));
  CAM_API_END
}

int cam_ctx_release_workspaces(cam_ctx* c) {
  CAM_API_BEGIN
  CAM_REQUIRE(c, "null argument");
  set_device(c);
  HIP_CHECK(hipStreamSynchronize(c->ctx.stream));
  c->ctx.ws_krylov.reset();
  c->ctx.ws_t1.reset();
  c->ctx.ws_t2.reset();
  CAM_API_END
}

int cam_ctx_stream(cam_ctx* c, void** stream) {
  CAM_API_BEGIN
  *stream = reinterpret_cast<void*>(c->ctx.stream);
  CAM_API_END
}

int cam_tensor_from_csc(cam_ctx* c, int nrules, const int16_t* rules, int nidx,
                        const cam_index_desc* idx, int num_row_indices, int64_t rows, int64_t cols,
                        const int64_t* p, const int64_t* i, const double* x, cam_tensor** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(c && out && p, "null argument");
  CAM_REQUIRE(nrules >= 1 && nrules <= kMaxRules, "unsupported number of charge rules");
  CAM_REQUIRE(nidx >= 0 && nidx <= 16, "unsupported number of indices");
  set_device(c);
  Rules R;
  R.N.assign(rules, rules + nrules);
  std::vector<QIndex> qi(nidx);
  for (int k = 0; k < nidx; ++k) {
    qi[k].inward = idx[k].inward != 0;
    qi[k].physical = idx[k].physical != 0;
    qi[k].size = idx[k].size;
    CAM_REQUIRE(idx[k].size >= 0, "negative index size");
    qi[k].charges.resize(idx[k].size);
    for (int64_t e = 0; e < idx[k].size; ++e) qi[k].charges[e] = R.normalise(idx[k].charges + e * nrules);
  }
  auto t = tensor_from_csc(R, std::move(qi), num_row_indices, rows, cols, p, i, x, c->ctx.stream);
  *out = new cam_tensor{t, c};
  CAM_API_END
}

int cam_tensor_shape(const cam_tensor* t, int* nidx, int* nrow, int* nrules) {
  CAM_API_BEGIN
  if (nidx) *nidx = t->t->nidx();
  if (nrow) *nrow = t->t->numrows;
  if (nrules) *nrules = t->t->rules.n();
  CAM_API_END
}

int cam_tensor_index(const cam_tensor* t, int k, int32_t* inward, int32_t* physical, int64_t* size,
                     int16_t* charges) {
  CAM_API_BEGIN
  CAM_REQUIRE(k >= 0 && k < t->t->nidx(), "index out of range");
  const QIndex& q = t->t->idx[k];
  if (inward) *inward = q.inward;
  if (physical) *physical = q.physical;
  if (size) *size = q.size;
  if (charges) {
    const int nr = t->t->rules.n();
    for (int64_t e = 0; e < q.size; ++e)
      for (int r = 0; r < nr; ++r) charges[e * nr + r] = q.charges[e].q[r];
  }
  CAM_API_END
}

int cam_tensor_nnz(const cam_tensor* t, int64_t* nnz) {
  CAM_API_BEGIN
  *nnz = t->t->total;
  CAM_API_END
}

int cam_tensor_to_csc(const cam_tensor* t, int64_t cap, int64_t* p, int64_t* i, double* x, int64_t* nnz) {
  CAM_API_BEGIN
  set_device(t->owner);
  const int64_t n = tensor_to_csc(*t->t, cap, p, i, x, t->owner->ctx.stream);
  if (nnz) *nnz = n;
  CAM_REQUIRE(!p || n <= cap, "cam_tensor_to_csc: capacity too small");
  CAM_API_END
}

int cam_tensor_destroy(cam_tensor* t) {
  CAM_API_BEGIN
  delete t;
  CAM_API_END
}

int cam_tensor_shift_num_row_indices(cam_tensor* t, int nrow) {
  CAM_API_BEGIN
  CAM_REQUIRE(nrow >= 1 && nrow <= t->t->nidx(), "Incorrect MPX reshape params for change rows");
  t->t->numrows = nrow;
  CAM_API_END
}

int cam_tensor_remove_dummy_indices(cam_tensor* tt, int n, const int64_t* which) {
  CAM_API_BEGIN
  tt->t->remove_dummy_indices(std::vector<int>(which, which + n));
  CAM_API_END
}

int cam_tensor_move_dummy_indices(cam_tensor* tt, int n, const int64_t* order) {
  CAM_API_BEGIN
  tt->t->move_dummy_indices(std::vector<int>(order, order + n));
  CAM_API_END
}

int cam_tensor_rescale(cam_tensor* t, double re, double im) {
  CAM_API_BEGIN
  set_device(t->owner);
  launch_scale(cplx{re, im}, t->t->data.get(), t->t->total, t->owner->ctx.stream);
  HIP_CHECK(hipStreamSynchronize(t->owner->ctx.stream));
  CAM_API_END
}

int cam_contract(cam_ctx* c, const cam_tensor* A, int conjA, const cam_tensor* B, int conjB, int npairs,
                 const int64_t* pairs, cam_tensor** out) {
  CAM_API_BEGIN
  set_device(c);
  std::vector<std::pair<int, int>> pr;
  for (int k = 0; k < npairs; ++k)
    pr.emplace_back(static_cast<int>(pairs[2 * k]), static_cast<int>(pairs[2 * k + 1]));
  auto t = contract(c->ctx, *A->t, conjA != 0, *B->t, conjB != 0, pr);
  *out = new cam_tensor{t, c};
  CAM_API_END
}

int cam_reorder(cam_ctx* c, const cam_tensor* A, int conjA, int n, const int64_t* order, int nrow,
                cam_tensor** out) {
  CAM_API_BEGIN
  set_device(c);
  std::vector<int> o(order, order + n);
  auto t = reorder(c->ctx, *A->t, conjA != 0, o, nrow);
  *out = new cam_tensor{t, c};
  CAM_API_END
}

int cam_extract_mpo_block(cam_ctx* c, const cam_tensor* W, int64_t r0, int64_t r1, int64_t c0, int64_t c1,
                          cam_tensor** out) {
  CAM_API_BEGIN
  set_device(c);
  *out = new cam_tensor{extract_mpo_block(c->ctx, *W->t, r0, r1, c0, c1), c};
  CAM_API_END
}

int cam_bond_hamiltonian(cam_ctx* c, const cam_tensor* W, int odd, cam_tensor** out) {
  CAM_API_BEGIN
  set_device(c);
  *out = new cam_tensor{bond_hamiltonian(c->ctx, *W->t, odd != 0), c};
  CAM_API_END
}

int cam_bond_evolution_operator(cam_ctx* c, const cam_tensor* H, double timestep, cam_tensor** out) {
  CAM_API_BEGIN
  set_device(c);
  *out = new cam_tensor{bond_evolution_operator(c->ctx, *H->t, timestep), c};
  CAM_API_END
}

int cam_heff_create(cam_ctx* c, const cam_tensor* L, const cam_tensor* W, const cam_tensor* R,
                    cam_heff** out, int64_t* n) {
  CAM_API_BEGIN
  set_device(c);
  auto* h = new cam_heff;
  h->owner = c;
  try {
    h->h.build(c->ctx, L->t, W->t, R->t);
  } catch (...) {
    delete h;
    throw;
  }
  *out = h;
  if (n) *n = h->h.n;
  CAM_API_END
}

int cam_heff_destroy(cam_heff* h) {
  CAM_API_BEGIN
  delete h;
  CAM_API_END
}

int cam_heff_allowed_map(const cam_heff* hc, int64_t* idx) {
  CAM_API_BEGIN
  auto* h = const_cast<cam_heff*>(hc);
  set_device(h->owner);
  h->h.build_perm();
  std::memcpy(idx, h->h.allowed.data(), h->h.allowed.size() * sizeof(int64_t));
  CAM_API_END
}

int cam_heff_set_projectors(cam_heff* h, int nproj, const double* const* P, const double* weights) {
  CAM_API_BEGIN
  CAM_REQUIRE(h && nproj >= 0 && (nproj == 0 || (P && weights)), "null argument");
  set_device(h->owner);
  std::vector<const cplx*> p(static_cast<size_t>(nproj));
  for (int k = 0; k < nproj; ++k) {
    CAM_REQUIRE(P[k], "null projector");
    p[k] = reinterpret_cast<const cplx*>(P[k]);
  }
  h->h.set_projectors(nproj, p.data(), weights);
  CAM_API_END
}

int cam_heff_flops(const cam_heff* h, double* flops) {
  CAM_API_BEGIN
  *flops = h->h.flops;
  CAM_API_END
}

int cam_heff_apply_blocked(cam_heff* h, const double* in, double* out, int precision) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  h->h.apply(reinterpret_cast<const cplx*>(in), reinterpret_cast<cplx*>(out), precision);
  CAM_API_END
}

int cam_heff_apply_blocked_noalloc(cam_heff* h, const double* in, double* out) {
  CAM_API_BEGIN
  CAM_REQUIRE(h && in && out, "null argument");
  set_device(h->owner);
  heff_apply_noalloc(h->h, reinterpret_cast<const cplx*>(in), reinterpret_cast<cplx*>(out));
  CAM_API_END
}

int cam_heff_apply_blocked32(cam_heff* h, const float* in, float* out) {
  CAM_API_BEGIN
  CAM_REQUIRE(h && in && out, "null argument");
  set_device(h->owner);
  h->h.apply_blocked32(reinterpret_cast<const float2*>(in), reinterpret_cast<float2*>(out));
  CAM_API_END
}

int cam_heff_profile(cam_heff* h, int precision, int reps, double* stage_ms, double* stage_flops) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  h->h.profile(precision, reps, stage_ms);
  if (stage_flops)
    for (int i = 0; i < 4; ++i) stage_flops[i] = h->h.stage_flops[i];
  CAM_API_END
}

int cam_heff_profile_device(cam_heff* h, int precision, int reps, const void* in, void* out, double* stage_ms,
                            double* stage_flops) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  CAM_REQUIRE(in && out, "cam_heff_profile_device: null vector");
  h->h.profile(precision, reps, stage_ms, in, out);
  if (stage_flops)
    for (int i = 0; i < 4; ++i) stage_flops[i] = h->h.stage_flops[i];
  CAM_API_END
}

int cam_heff_set_slices(cam_heff* h, int row_slices, int col_slices) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(row_slices >= 1 && row_slices <= 4096 && col_slices >= 1 && col_slices <= 4096,
              "slice counts must be 1..4096");
  h->h.plan_slices(row_slices, col_slices);
  HIP_CHECK(hipStreamSynchronize(h->owner->ctx.stream));
  CAM_API_END
}

int cam_heff_set_workspace_limit(cam_heff* h, double bytes, int precision) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(bytes > 0.0, "workspace limit must be positive");
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  // FP32 intermediates take 8 bytes per element, the FP64 count's half
  h->h.choose_slices(precision == CAM_FP32 ? 2.0 * bytes : bytes);
  h->h.plan_slices(h->h.rslices, h->h.cslices);
  HIP_CHECK(hipStreamSynchronize(h->owner->ctx.stream));
  CAM_API_END
}

int cam_heff_get_slices(const cam_heff* h, int* row_slices, int* col_slices, double* workspace_bytes) {
  CAM_API_BEGIN
  if (row_slices) *row_slices = h->h.rslices;
  if (col_slices) *col_slices = h->h.cslices;
  if (workspace_bytes) *workspace_bytes = h->h.slice_workspace_bytes(h->h.rslices, h->h.cslices);
  CAM_API_END
}

int cam_heff_stage_bytes(const cam_heff* h, double* b) {
  CAM_API_BEGIN
  const cam::Heff& H = h->h;
  const double e = sizeof(cam::cplx);
  b[0] = e * double(H.n + H.L->total + H.t1n);
  b[1] = e * double(H.t1n + H.W->total + H.t2n);
  b[2] = e * double(H.t2n + H.W->total + H.t3n);
  b[3] = e * double(H.t3n + H.R->total + H.n);
  CAM_API_END
}

int cam_heff_set_launch(cam_heff* h, int precision, int algo, int tile_mode, int wg_per_cu, int kchunk) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  CAM_REQUIRE(algo == kAlgo4M || algo == kAlgo3M, "algo must be 0 (four real products) or 1 (three)");
  CAM_REQUIRE(tile_mode >= 0 && tile_mode < kTileModes, "tile_mode out of range");
  CAM_REQUIRE(wg_per_cu >= 1 && wg_per_cu <= 64, "wg_per_cu out of range");
  CAM_REQUIRE(kchunk == 8 || kchunk == 16, "kchunk must be 8 or 16");
  LaunchCfg c;
  c.algo = algo;
  c.tile_mode = tile_mode;
  c.wg_per_cu = wg_per_cu;
  c.kc = kchunk;
  h->h.set_launch(precision, c);
  CAM_API_END
}

int cam_heff_get_launch(const cam_heff* h, int precision, int* cfg) {
  CAM_API_BEGIN
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  for (int i = 0; i < 4; ++i) {
    const LaunchCfg& c = h->h.plans[0][i].cfg[precision];
    cfg[4 * i] = c.algo;
    cfg[4 * i + 1] = c.tile_mode;
    cfg[4 * i + 2] = c.wg_per_cu;
    cfg[4 * i + 3] = c.kc;
  }
  CAM_API_END
}

int cam_heff_to_blocked(cam_heff* h, const double* in_dev, double* out_dev) {
  CAM_API_BEGIN
  set_device(h->owner);
  h->h.build_perm();
  launch_scatter_map(reinterpret_cast<const cplx*>(in_dev), reinterpret_cast<cplx*>(out_dev),
                     h->h.perm.get(), h->h.n, h->owner->ctx.stream);
  CAM_API_END
}

int cam_heff_from_blocked(cam_heff* h, const double* in_dev, double* out_dev) {
  CAM_API_BEGIN
  set_device(h->owner);
  h->h.build_perm();
  launch_gather_map(reinterpret_cast<const cplx*>(in_dev), reinterpret_cast<cplx*>(out_dev),
                    h->h.perm.get(), h->h.n, h->owner->ctx.stream);
  CAM_API_END
}

int cam_heff_apply(cam_heff* h, const double* in, double* out, int mem_kind, int precision) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(precision == CAM_FP64 || precision == CAM_FP32, "bad precision");
  CAM_REQUIRE(mem_kind == CAM_MEM_HOST || mem_kind == CAM_MEM_DEVICE, "bad mem_kind");
  Heff& H = h->h;
  hipStream_t s = h->owner->ctx.stream;
  H.build_perm();
  const int64_t n = H.n;
  h->bin.ensure(std::max<int64_t>(n, 1));
  h->bout.ensure(std::max<int64_t>(n, 1));
  const cplx* rin = reinterpret_cast<const cplx*>(in);
  cplx* rout = reinterpret_cast<cplx*>(out);
  if (mem_kind == CAM_MEM_HOST) {
    h->rin.ensure(std::max<int64_t>(n, 1));
    h->rout.ensure(std::max<int64_t>(n, 1));
    h->rin.upload(rin, n, s);
    rin = h->rin.get();
    rout = h->rout.get();
  }
  launch_scatter_map(rin, h->bin.get(), H.perm.get(), n, s);
  H.apply(h->bin.get(), h->bout.get(), precision);
  launch_gather_map(h->bout.get(), rout, H.perm.get(), n, s);
  if (mem_kind == CAM_MEM_HOST) {
    HIP_CHECK(hipMemcpyAsync(out, rout, n * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  CAM_API_END
}

int cam_heff_lowest(cam_heff* h, const double* guess, int converge, double tol, int ncv, int mixed,
                    double* eval, double* evec, double* info) {
  CAM_API_BEGIN
  set_device(h->owner);
  heff_lowest(h->h, reinterpret_cast<const cplx*>(guess), converge != 0, tol, ncv, mixed != 0, eval,
              reinterpret_cast<cplx*>(evec), info);
  CAM_API_END
}

int cam_heff_lowest_blocked(cam_heff* h, double* x, int have_start, int converge, double tol, int ncv, int mixed,
                            double* eval, double* info) {
  CAM_API_BEGIN
  set_device(h->owner);
  CAM_REQUIRE(x != nullptr, "cam_heff_lowest_blocked: null vector");
  *eval = eigs_lowest(h->h, reinterpret_cast<cplx*>(x), have_start != 0, converge != 0, tol, ncv, mixed != 0, info);
  CAM_API_END
}

int cam_svd_truncate(cam_ctx* c, const cam_tensor* theta, int64_t D, double max_trunc, cam_tensor** U,
                     cam_tensor** Vd, double* values, int64_t cap, int64_t* nvals, double* discarded) {
  CAM_API_BEGIN
  set_device(c);
  TensorPtr u, vd;
  std::vector<double> vals;
  double disc = 0.0;
  svd_truncate(c->ctx, *theta->t, D, max_trunc, u, vd, vals, disc);
  CAM_REQUIRE(static_cast<int64_t>(vals.size()) <= cap, "cam_svd_truncate: values capacity too small");
  std::copy(vals.begin(), vals.end(), values);
  if (nvals) *nvals = static_cast<int64_t>(vals.size());
  if (discarded) *discarded = disc;
  *U = new cam_tensor{u, c};
  *Vd = new cam_tensor{vd, c};
  CAM_API_END
}

int cam_tebd_gate(cam_ctx* c, const cam_tensor* G, const cam_tensor* A0, const cam_tensor* A1,
                  const double* lambda0, int64_t n0, int64_t D, double minS, cam_tensor** A0n,
                  cam_tensor** A1n, double* lambda1, int64_t cap, int64_t* n1, double* truncation) {
  CAM_API_BEGIN
  set_device(c);
  TensorPtr a0, a1;
  std::vector<double> l1;
  double tr = 0.0;
  std::vector<double> l0(lambda0, lambda0 + n0);
  tebd_gate(c->ctx, *G->t, *A0->t, *A1->t, l0, D, minS, a0, a1, l1, tr);
  CAM_REQUIRE(static_cast<int64_t>(l1.size()) <= cap, "cam_tebd_gate: lambda capacity too small");
  std::copy(l1.begin(), l1.end(), lambda1);
  if (n1) *n1 = static_cast<int64_t>(l1.size());
  if (truncation) *truncation = tr;
  *A0n = new cam_tensor{a0, c};
  *A1n = new cam_tensor{a1, c};
  CAM_API_END
}

int cam_heff_vjp(cam_heff* h, const double* psi, double* dpsi, double* dW, int64_t nnzW) {
  CAM_API_BEGIN
  set_device(h->owner);
  heff_vjp(h->h, reinterpret_cast<const cplx*>(psi), reinterpret_cast<cplx*>(dpsi),
           reinterpret_cast<cplx*>(dW), nnzW);
  CAM_API_END
}

int cam_heff_coupling_grad(cam_heff* h, const double* psi, int64_t nnzW, int ncoup, const int64_t* gen_ptr,
                           const int64_t* gen_entry, const double* gen_coef, double* dE, double* dpsi) {
  CAM_API_BEGIN
  CAM_REQUIRE(ncoup >= 0 && nnzW >= 0, "coupling_grad: negative count");
  CAM_REQUIRE(ncoup == 0 || (gen_ptr && dE), "coupling_grad: null generator table");
  CAM_REQUIRE(ncoup == 0 || gen_ptr[0] == 0, "coupling_grad: gen_ptr[0] must be 0");
  for (int c = 0; c < ncoup; ++c) {
    CAM_REQUIRE(gen_ptr[c + 1] >= gen_ptr[c], "coupling_grad: gen_ptr must be non-decreasing");
    for (int64_t k = gen_ptr[c]; k < gen_ptr[c + 1]; ++k)
      CAM_REQUIRE(gen_entry[k] >= 0 && gen_entry[k] < nnzW, "coupling_grad: generator entry outside W");
  }
  set_device(h->owner);
  std::vector<cplx> own;
  cplx* dp = reinterpret_cast<cplx*>(dpsi);
  if (!dp) {
    own.resize(static_cast<size_t>(h->h.n));
    dp = own.data();
  }
  std::vector<cplx> dW(static_cast<size_t>(std::max<int64_t>(nnzW, 1)));
  heff_vjp(h->h, reinterpret_cast<const cplx*>(psi), dp, dW.data(), nnzW);
  // chain rule through W = sum_c theta_c dW/dtheta_c (W linear in the couplings)
  for (int c = 0; c < ncoup; ++c) {
    double re = 0.0, im = 0.0;
    for (int64_t k = gen_ptr[c]; k < gen_ptr[c + 1]; ++k) {
      const cplx g = dW[gen_entry[k]];
      const double cr = gen_coef[2 * k], ci = gen_coef[2 * k + 1];
      re += g.re * cr - g.im * ci;
      im += g.re * ci + g.im * cr;
    }
    dE[2 * c] = re;
    dE[2 * c + 1] = im;
  }
  CAM_API_END
}

}  // extern "C"

struct cam_idmrg_s {
  std::unique_ptr<cam::IDMRG> d;
  cam_ctx* owner;
};

struct cam_itebd_s {
  std::unique_ptr<cam::ITEBD> d;
  cam_ctx* owner;
};
struct cam_unitcell_s {
  cam::StoredBasis basis;
  std::vector<TensorPtr> matrices;
  std::vector<std::vector<double>> lambdas;
  cam_ctx* owner;
};
struct cam_ftebd_s {
  std::unique_ptr<cam::FTEBD> d;
  cam_ctx* owner;
};

extern "C" {

int cam_idmrg_create(cam_ctx* c, const cam_tensor* W, const int16_t* target, int converge, int mixed, int ncv,
                     cam_idmrg** out) {
  CAM_API_BEGIN
  set_device(c);
  Charge t = W->t->rules.normalise(target);
  auto* r = new cam_idmrg;
  r->owner = c;
  r->d.reset(new cam::IDMRG(c->ctx, W->t, t, converge != 0, mixed != 0, ncv > 0 ? ncv : 20));
  *out = r;
  CAM_API_END
}

static void write_record(const cam::IDMRGRecord& r, double* record) {
  const double v[CAM_IDMRG_RECORD] = {double(r.N),    r.energy,      r.energy_per_vertex, r.entropy,
                                      r.truncation,   double(r.chi), double(r.n),         r.matvecs_fp64,
                                      r.matvecs_fp32, r.residual,    r.seconds,           r.heff_flops,
                                      r.eig_seconds,  r.svd_seconds, r.block_seconds,     r.fidelity,
                                      r.guess_dropped, double(r.left), double(r.right)};
  std::copy(v, v + CAM_IDMRG_RECORD, record);
}

int cam_idmrg_grow(cam_idmrg* d, int64_t chi, double truncation, double* record) {
  CAM_API_BEGIN
  CAM_REQUIRE(d && record, "null argument");
  set_device(d->owner);
  write_record(d->d->grow(chi, truncation), record);
  CAM_API_END
}

int cam_idmrg_move(cam_idmrg* d, int to_right, int64_t chi, double truncation, double* record) {
  CAM_API_BEGIN
  CAM_REQUIRE(d && record, "null argument");
  set_device(d->owner);
  write_record(d->d->move(to_right != 0, chi, truncation), record);
  CAM_API_END
}

int cam_fdmrg_sweeps(cam_idmrg* d, int64_t num_sweeps, int64_t chi, double truncation, double* records,
                     int32_t* midpoint, int64_t cap, int64_t* nrec) {
  CAM_API_BEGIN
  CAM_REQUIRE(d && nrec, "null argument");
  CAM_REQUIRE(num_sweeps >= 0, "number of sweeps must be >= 0");
  set_device(d->owner);
  // a sweep moves right to the end (nright steps), left to the other end
  // (N - 2) and right to the middle ((N - 2) / 2); every sweep after the
  // first starts in the middle
  const int64_t span = d->d->size > 2 ? d->d->size - 2 : 0;
  *nrec = num_sweeps > 0 && span > 0 ? d->d->nright + span + span / 2 + (num_sweeps - 1) * 2 * span : 0;
  CAM_REQUIRE(!records || cap >= *nrec, "capacity too small");
  if (!records) return 0;
  std::vector<int> mid;
  const std::vector<cam::IDMRGRecord> rs = d->d->finite_sweeps(num_sweeps, chi, truncation, &mid);
  CAM_REQUIRE(static_cast<int64_t>(rs.size()) == *nrec, "finite sweeps: unexpected number of steps");
  for (size_t i = 0; i < rs.size(); ++i) {
    write_record(rs[i], records + i * CAM_IDMRG_RECORD);
    if (midpoint) midpoint[i] = mid[i];
  }
  CAM_API_END
}

int cam_idmrg_blocks(const cam_idmrg* d, cam_tensor** LB, cam_tensor** RB) {
  CAM_API_BEGIN
  CAM_REQUIRE(d && LB && RB, "null argument");
  CAM_REQUIRE(d->d->LB && d->d->RB, "iDMRG: no blocks yet (grow first)");
  auto* l = new cam_tensor{d->d->LB, d->owner};
  auto* r = new cam_tensor{d->d->RB, d->owner};
  *LB = l;
  *RB = r;
  CAM_API_END
}

int cam_idmrg_values(const cam_idmrg* d, double* values, int64_t cap, int64_t* n) {
  CAM_API_BEGIN
  const auto& v = d->d->values;
  if (n) *n = static_cast<int64_t>(v.size());
  CAM_REQUIRE(static_cast<int64_t>(v.size()) <= cap || !values, "capacity too small");
  if (values) std::copy(v.begin(), v.end(), values);
  CAM_API_END
}

int cam_idmrg_destroy(cam_idmrg* d) {
  CAM_API_BEGIN
  delete d;
  CAM_API_END
}

int cam_itebd_create(cam_ctx* c, const cam_tensor* W, const cam_tensor* A0, const cam_tensor* A1,
                     const double* lambda0, int64_t n0, const double* lambda1, int64_t n1, double dt, int order,
                     cam_itebd** out) {
  CAM_API_BEGIN
  set_device(c);
  CAM_REQUIRE(lambda0 && lambda1 && n0 > 0 && n1 > 0, "iTEBD: bond weights missing");
  auto* r = new cam_itebd;
  r->owner = c;
  try {
    r->d.reset(new cam::ITEBD(c->ctx, *W->t, A0->t, A1->t, std::vector<double>(lambda0, lambda0 + n0),
                              std::vector<double>(lambda1, lambda1 + n1), dt, order));
  } catch (...) {
    delete r;
    throw;
  }
  *out = r;
  CAM_API_END
}

int cam_itebd_evolve(cam_itebd* it, int64_t nsteps, int64_t D, double minS, double* record) {
  CAM_API_BEGIN
  set_device(it->owner);
  const cam::ITEBDRecord r = it->d->evolve(nsteps, D, minS);
  const double v[7] = {double(r.steps), r.time, r.truncation, r.entropy, double(r.chi), r.seconds, r.gate_flops};
  if (record) std::copy(v, v + 7, record);
  CAM_API_END
}

int cam_itebd_evolve_measured(cam_itebd* it, int64_t nsteps, int64_t D, double minS, int64_t interval, int nops,
                              const cam_tensor* const* ops, double* records, int64_t cap, int64_t* nrec) {
  CAM_API_BEGIN
  CAM_REQUIRE(it && nrec && nops >= 0 && (nops == 0 || ops), "null argument");
  set_device(it->owner);
  std::vector<TensorPtr> m;
  for (int k = 0; k < nops; ++k) {
    CAM_REQUIRE(ops[k], "null measured operator");
    m.push_back(ops[k]->t);
  }
  const std::vector<cam::ITEBDMeasure> r = it->d->evolve_measured(nsteps, D, minS, interval, m);
  *nrec = static_cast<int64_t>(r.size());
  const int64_t per = 7 + 2 * nops;
  if (records)
    for (int64_t k = 0; k < std::min<int64_t>(cap, *nrec); ++k) {
      double* o = records + k * per;
      const cam::ITEBDMeasure& e = r[k];
      o[0] = static_cast<double>(e.step);
      o[1] = e.time;
      o[2] = e.truncation;
      o[3] = e.entropy;
      o[4] = e.abs_overlap;
      o[5] = e.overlap.re;
      o[6] = e.overlap.im;
      for (int j = 0; j < nops; ++j) {
        o[7 + 2 * j] = e.values[j].re;
        o[8 + 2 * j] = e.values[j].im;
      }
    }
  CAM_API_END
}

int cam_itebd_lambda(const cam_itebd* it, int k, double* values, int64_t cap, int64_t* n) {
  CAM_API_BEGIN
  CAM_REQUIRE(k == 0 || k == 1, "iTEBD: bond index must be 0 or 1");
  const auto& v = it->d->lam[k];
  if (n) *n = static_cast<int64_t>(v.size());
  CAM_REQUIRE(static_cast<int64_t>(v.size()) <= cap || !values, "capacity too small");
  if (values) std::copy(v.begin(), v.end(), values);
  CAM_API_END
}

int cam_itebd_tensor(const cam_itebd* it, int k, cam_tensor** out) {
  CAM_API_BEGIN
  set_device(it->owner);
  CAM_REQUIRE(k == 0 || k == 1, "iTEBD: tensor index must be 0 or 1");
  const cam::BSTensor& A = *it->d->A[k];
  std::vector<int> ident(A.nidx());
  for (int i = 0; i < A.nidx(); ++i) ident[i] = i;
  *out = new cam_tensor{reorder(it->owner->ctx, A, false, ident, A.numrows), it->owner};
  CAM_API_END
}

int cam_itebd_destroy(cam_itebd* it) {
  CAM_API_BEGIN
  delete it;
  CAM_API_END
}

int cam_ftebd_create(cam_ctx* c, const cam_tensor* W, const cam_tensor* const* sites, int64_t nsites, double dt,
                     int order, cam_ftebd** out) {
  CAM_API_BEGIN
  set_device(c);
  CAM_REQUIRE(sites && nsites >= 2, "finite TEBD: at least two site tensors");
  std::vector<TensorPtr> v;
  for (int64_t i = 0; i < nsites; ++i) {
    CAM_REQUIRE(sites[i], "finite TEBD: null site tensor");
    v.push_back(sites[i]->t);
  }
  auto* r = new cam_ftebd;
  r->owner = c;
  try {
    r->d.reset(new cam::FTEBD(c->ctx, *W->t, v, dt, order));
  } catch (...) {
    delete r;
    throw;
  }
  *out = r;
  CAM_API_END
}

int cam_ftebd_evolve(cam_ftebd* t, int64_t nsteps, int64_t D, double minS, int64_t interval) {
  CAM_API_BEGIN
  set_device(t->owner);
  t->d->evolve(nsteps, D, minS, interval);
  CAM_API_END
}

int cam_ftebd_records(const cam_ftebd* t, double* records, int64_t cap, int64_t* n) {
  CAM_API_BEGIN
  const auto& R = t->d->records;
  if (n) *n = static_cast<int64_t>(R.size());
  CAM_REQUIRE(static_cast<int64_t>(R.size()) <= cap || !records, "capacity too small");
  if (records)
    for (size_t i = 0; i < R.size(); ++i) {
      const double v[7] = {double(R[i].step), R[i].time,       R[i].truncation, R[i].entropy,
                           R[i].overlap_re,   R[i].overlap_im, R[i].abs_overlap};
      std::copy(v, v + 7, records + 7 * i);
    }
  CAM_API_END
}

int cam_ftebd_site(const cam_ftebd* t, int64_t v, cam_tensor** out) {
  CAM_API_BEGIN
  set_device(t->owner);
  CAM_REQUIRE(v >= 1 && v <= t->d->N, "finite TEBD: site index out of range (1..N)");
  const cam::BSTensor& A = *t->d->left[v];
  std::vector<int> ident(A.nidx());
  for (int i = 0; i < A.nidx(); ++i) ident[i] = i;
  *out = new cam_tensor{reorder(t->owner->ctx, A, false, ident, A.numrows), t->owner};
  CAM_API_END
}

int cam_ftebd_destroy(cam_ftebd* t) {
  CAM_API_BEGIN
  delete t;
  CAM_API_END
}

namespace {
cam::StoredBasis make_basis(int nrules, const int16_t* rules, int64_t d, const int16_t* charges,
                            const double* energies) {
  CAM_REQUIRE(nrules >= 1 && nrules <= cam::kMaxRules, "unsupported number of charge rules");
  CAM_REQUIRE(d >= 1 && charges, "basis: no physical states");
  cam::StoredBasis b;
  b.rules.N.assign(rules, rules + nrules);
  for (int64_t e = 0; e < d; ++e) b.charges.push_back(b.rules.normalise(charges + e * nrules));
  b.energies.assign(static_cast<size_t>(d), 0.0);
  if (energies) std::copy(energies, energies + d, b.energies.begin());
  return b;
}
}  // namespace

int cam_tensor_store(const cam_tensor* t, const char* path) {
  CAM_API_BEGIN
  CAM_REQUIRE(t && path, "null argument");
  set_device(t->owner);
  cam::store_tensor(path, *t->t, t->owner->ctx.stream);
  CAM_API_END
}

int cam_tensor_load(cam_ctx* c, const char* path, int nrules, const int16_t* rules, int64_t d,
                    const int16_t* phys_charges, cam_tensor** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(c && path && rules && out, "null argument");
  set_device(c);
  const cam::StoredBasis b = make_basis(nrules, rules, d, phys_charges, nullptr);
  *out = new cam_tensor{cam::load_tensor(path, b.rules, b.charges, c->ctx.stream), c};
  CAM_API_END
}

int cam_unitcell_store(const char* path, int nrules, const int16_t* rules, int64_t d, const int16_t* phys_charges,
                       const double* energies, int n, const cam_tensor* const* matrices, const double* const* lambdas,
                       const int64_t* lambda_sizes) {
  CAM_API_BEGIN
  CAM_REQUIRE(path && rules && (n == 0 || (matrices && lambdas && lambda_sizes)), "null argument");
  CAM_REQUIRE(n >= 0, "negative unit cell size");
  const cam::StoredBasis b = make_basis(nrules, rules, d, phys_charges, energies);
  std::vector<TensorPtr> m;
  std::vector<std::vector<double>> l;
  hipStream_t s = nullptr;
  for (int k = 0; k < n; ++k) {
    CAM_REQUIRE(matrices[k], "null unit cell matrix");
    if (k == 0) {
      set_device(matrices[k]->owner);
      s = matrices[k]->owner->ctx.stream;
    }
    m.push_back(matrices[k]->t);
    l.emplace_back(lambdas[k], lambdas[k] + lambda_sizes[k]);
  }
  cam::store_unitcell(path, b, m, l, s);
  CAM_API_END
}

int cam_unitcell_load(cam_ctx* c, const char* path, cam_unitcell** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(c && path && out, "null argument");
  set_device(c);
  auto* u = new cam_unitcell;
  u->owner = c;
  try {
    cam::load_unitcell(path, u->basis, u->matrices, u->lambdas, c->ctx.stream);
  } catch (...) {
    delete u;
    throw;
  }
  *out = u;
  CAM_API_END
}

int cam_unitcell_basis(const cam_unitcell* u, int* nrules, int16_t* rules, int64_t* d, int16_t* charges,
                       double* energies, int64_t cap) {
  CAM_API_BEGIN
  const cam::StoredBasis& b = u->basis;
  const int nr = b.rules.n();
  if (nrules) *nrules = nr;
  if (d) *d = static_cast<int64_t>(b.charges.size());
  if (rules) std::copy(b.rules.N.begin(), b.rules.N.end(), rules);
  CAM_REQUIRE((!charges && !energies) || static_cast<int64_t>(b.charges.size()) <= cap, "capacity too small");
  if (charges)
    for (size_t e = 0; e < b.charges.size(); ++e) std::copy(b.charges[e].q.begin(), b.charges[e].q.begin() + nr,
                                                            charges + e * nr);
  if (energies) std::copy(b.energies.begin(), b.energies.end(), energies);
  CAM_API_END
}

int cam_unitcell_size(const cam_unitcell* u, int* n) {
  CAM_API_BEGIN
  *n = static_cast<int>(u->matrices.size());
  CAM_API_END
}

int cam_unitcell_matrix(const cam_unitcell* u, int k, cam_tensor** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(k >= 0 && k < static_cast<int>(u->matrices.size()), "unit cell index out of range");
  *out = new cam_tensor{u->matrices[k], u->owner};
  CAM_API_END
}

int cam_unitcell_lambda(const cam_unitcell* u, int k, double* values, int64_t cap, int64_t* n) {
  CAM_API_BEGIN
  CAM_REQUIRE(k >= 0 && k < static_cast<int>(u->lambdas.size()), "unit cell index out of range");
  const auto& v = u->lambdas[k];
  if (n) *n = static_cast<int64_t>(v.size());
  CAM_REQUIRE(static_cast<int64_t>(v.size()) <= cap || !values, "capacity too small");
  if (values) std::copy(v.begin(), v.end(), values);
  CAM_API_END
}

int cam_unitcell_destroy(cam_unitcell* u) {
  CAM_API_BEGIN
  delete u;
  CAM_API_END
}

// ---------------------------------------------------------------------------
// Unit cells on the device, orthogonalisation and the energy per vertex
// ---------------------------------------------------------------------------

static cam_unitcell* new_cell(cam_ctx* c, std::vector<TensorPtr> m, std::vector<std::vector<double>> l) {
  CAM_REQUIRE(!m.empty() && m.size() == l.size(), "Invalid Unit Cell: matrices and lambdas differ in number");
  auto* u = new cam_unitcell;
  u->owner = c;
  u->basis.rules = m.front()->rules;
  CAM_REQUIRE(m.front()->nidx() >= 1 && m.front()->idx[0].physical, "unit cell matrix without a physical index");
  u->basis.charges = m.front()->idx[0].charges;
  u->basis.energies.assign(u->basis.charges.size(), 0.0);
  u->matrices = std::move(m);
  u->lambdas = std::move(l);
  return u;
}

int cam_unitcell_create(cam_ctx* c, int n, const cam_tensor* const* matrices, const double* const* lambdas,
                        const int64_t* lambda_sizes, cam_unitcell** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(c && out && n >= 1 && matrices && lambdas && lambda_sizes, "null argument");
  std::vector<TensorPtr> m;
  std::vector<std::vector<double>> l;
  for (int k = 0; k < n; ++k) {
    CAM_REQUIRE(matrices[k] && (lambdas[k] || lambda_sizes[k] == 0), "null unit cell entry");
    m.push_back(matrices[k]->t);
    l.emplace_back(lambdas[k], lambdas[k] + lambda_sizes[k]);
  }
  *out = new_cell(c, std::move(m), std::move(l));
  CAM_API_END
}

int cam_unitcell_from_decomposition(cam_ctx* c, const cam_tensor* U, const double* values, int64_t nvalues,
                                    const cam_tensor* Vd, const double* previous, int64_t nprevious,
                                    cam_unitcell** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(c && U && Vd && values && previous && out, "null argument");
  set_device(c);
  std::vector<double> v(values, values + nvalues), p(previous, previous + nprevious);
  TensorPtr A1 = cam::unit_cell_second(c->ctx, *Vd->t, v, p);
  *out = new_cell(c, {U->t, A1}, {p, v});
  CAM_API_END
}

int cam_idmrg_unitcell(cam_idmrg* g, cam_unitcell** out) {
  CAM_API_BEGIN
  CAM_REQUIRE(g && out, "null argument");
  CAM_REQUIRE(g->d->U && g->d->Vd, "iDMRG: no decomposition yet (grow first)");
  set_device(g->owner);
  cam::IDMRG& d = *g->d;
  TensorPtr A1 = cam::unit_cell_second(g->owner->ctx, *d.Vd, d.values, d.prev_lambda);
  *out = new_cell(g->owner, {d.U, A1}, {d.prev_lambda, d.values});
  CAM_API_END
}

int cam_transfer_apply(const cam_unitcell* u, int side, const double* in, double* out) {
  CAM_API_BEGIN
  CAM_REQUIRE(u && in && out, "null argument");
  set_device(u->owner);
  cam::transfer_apply(u->owner->ctx, u->matrices, side, reinterpret_cast<const cplx*>(in),
                      reinterpret_cast<cplx*>(out));
  CAM_API_END
}

int cam_transfer_leading(const cam_unitcell* u, int side, double tol, int ncv, double* value, double* info) {
  CAM_API_BEGIN
  CAM_REQUIRE(u && value, "null argument");
  CAM_REQUIRE(ncv <= 0 || (ncv >= 2 && ncv <= 63), "transfer matrix: ncv must be 2..63 (or <= 0 for 20)");
  set_device(u->owner);
  const cam::LeadingEig e =
      cam::transfer_leading(u->owner->ctx, u->matrices, side, tol > 0.0 ? tol : 1e-14, nullptr, ncv);
  value[0] = e.value.re;
  value[1] = e.value.im;
  if (info) {
    info[0] = e.matvecs;
    info[1] = e.restarts;
    info[2] = e.residual;
    info[3] = e.converged ? 1.0 : 0.0;
  }
  CAM_API_END
}

static int orthogonalise_cell(const cam_unitcell* u, cam_unitcell** out, double* info, bool inversion_symmetric) {
  CAM_API_BEGIN
  CAM_REQUIRE(u && out, "null argument");
  CAM_REQUIRE(u->matrices.size() == 2, "Orthogonalise: two-vertex unit cells only");
  set_device(u->owner);
  TensorPtr A0n, A1n;
  std::vector<double> l0;
  cam::OrthoInfo oi;
  if (inversion_symmetric)
    cam::orthogonalise_inversion_symmetric(u->owner->ctx, u->matrices[0], u->matrices[1], u->lambdas[0], A0n, A1n,
                                           l0, oi);
  else
    cam::orthogonalise(u->owner->ctx, u->matrices[0], u->matrices[1], A0n, A1n, l0, oi);
  cam_unitcell* r = new_cell(u->owner, {A0n, A1n}, {l0, u->lambdas[1]});
  r->basis = u->basis;
  *out = r;
  if (info) {
    const double v[12] = {oi.left.value.re, oi.left.value.im, oi.right.value.re, oi.right.value.im,
                          oi.entropy, static_cast<double>(l0.size()),
                          static_cast<double>(oi.left.matvecs), static_cast<double>(oi.right.matvecs),
                          oi.left.residual, oi.right.residual, static_cast<double>(oi.kept_left + oi.kept_right),
                          oi.seconds};
    std::copy(v, v + 12, info);
  }
  CAM_API_END
}

int cam_orthogonalise(const cam_unitcell* u, cam_unitcell** out, double* info) {
  return orthogonalise_cell(u, out, info, false);
}

int cam_orthogonalise_inversion_symmetric(const cam_unitcell* u, cam_unitcell** out, double* info) {
  return orthogonalise_cell(u, out, info, true);
}

int cam_itwovertex_energy(const cam_tensor* W, const cam_unitcell* u, double* energy, double* terms) {
  CAM_API_BEGIN
  CAM_REQUIRE(W && u && energy, "null argument");
  CAM_REQUIRE(u->matrices.size() == 2 && u->lambdas.size() == 2, "Only two vertex basis accepted for now");
  set_device(u->owner);
  cplx t[3];
  const cplx e = cam::i_two_vertex_energy(u->owner->ctx, *W->t, u->matrices[0], u->matrices[1], u->lambdas[0], t);
  energy[0] = e.re;
  energy[1] = e.im;
  if (terms)
    for (int k = 0; k < 3; ++k) {
      terms[2 * k] = t[k].re;
      terms[2 * k + 1] = t[k].im;
    }
  CAM_API_END
}

}  // extern "C"
