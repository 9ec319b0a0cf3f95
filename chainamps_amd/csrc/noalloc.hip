// The H_eff matvec for callers that cannot allocate device memory.
//
// No counterpart in the reference.  Code under HIP graph capture, tight
// serving loops and harnesses that forbid hipMalloc cannot let the matvec
// grow the context's intermediate workspaces; this variant keeps the two
// intermediate buffers (T1/T3 and T2 of the four stages) in `scratch`, a
// file-scope device array of kNoAllocScratchBytes that every GPU the library
// runs on holds until the process ends.  Its address on each device is taken
// once with hipGetSymbolAddress while that device is current (noalloc_setup,
// called by cam_ctx_create); host code uses only that address.
//
// The kernels of one matvec write every intermediate element before they
// read it (each intermediate block is the output of exactly one problem of
// its stage), so the array needs no clearing between calls.  Every call on
// one GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, two host threads) race.
#include <array>
#include <mutex>

#include "heff.hpp"
#include "noalloc.hpp"

namespace cam {

namespace {

constexpr int64_t kScratchElems = kNoAllocScratchBytes / static_cast<int64_t>(sizeof(double2));
static __device__ __align__(256) double2 scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<void*, kMaxDevices> g_ptr{};

}  // namespace

void noalloc_setup(int device) {
  CAM_REQUIRE(device >= 0 && device < kMaxDevices, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return;
  int cur = -1;
  HIP_CHECK(hipGetDevice(&cur));
  CAM_REQUIRE(cur == device, "noalloc: set-up must run with its device current");
  void* p = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CAM_REQUIRE(p != nullptr, "noalloc: no device address for the scratch array");
  g_ptr[device] = p;
}

void heff_apply_noalloc(Heff& H, const cplx* in, cplx* out) {
  const int device = H.ctx->device;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    CAM_REQUIRE(device >= 0 && device < kMaxDevices && g_ptr[device],
                "noalloc: the scratch array of this device was not set up (create the context first)");
    base = g_ptr[device];
  }
  const int64_t need = H.ws1n + H.ws2n;
  CAM_REQUIRE(need <= kScratchElems,
              "noalloc: the matvec's intermediates (" + std::to_string(need * 16) + " bytes) exceed the " +
                  std::to_string(kNoAllocScratchBytes) +
                  "-byte scratch array; slice the matvec (cam_heff_set_slices) or use cam_heff_apply_blocked");
  // the plans must already be on the device at their current launch
  // configuration: uploading them (or autotuning) would allocate
  for (const auto& st : H.plans)
    for (const GemmPlan& p : st) {
      const LaunchCfg& c = p.cfg[0];
      CAM_REQUIRE(p.uploaded && p.built[c.tile_mode] && p.tiles_uploaded[c.tile_mode],
                  "noalloc: the matvec plans are not on the device at their launch configuration "
                  "(run cam_heff_apply_blocked once, or cam_heff_set_launch, before this call)");
    }
  hipStream_t s = H.ctx->stream;
  cplx* t1 = static_cast<cplx*>(base);
  cplx* t2 = t1 + H.ws1n;
  GemmBases b{};
  b.p[kSpL] = H.L->data.get();
  b.p[kSpW] = H.W->data.get();
  b.p[kSpR] = H.R->data.get();
  b.p[kSpT1] = t1;
  b.p[kSpT2] = t2;
  b.p[kSpT3] = t1;     // stage C writes T3 over T1
  b.p[kSpIn] = in;
  b.p[kSpOut] = out;
  for (auto& st : H.plans)
    for (int i = 0; i < 4; ++i) launch_grouped_gemm(st[i], 0, s, &b);
  for (int pi : H.untouched) launch_zero(out + H.pblocks[pi].off, H.pblocks[pi].size * sizeof(cplx), s);
  if (H.nproj) H.add_projectors(in, out);
}

}  // namespace cam
