// Auxiliary device kernels (HBM-bound): block permutation, index-map
// gather/scatter, precision conversion and Lanczos vector operations.
#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace cam {

constexpr int kMaxDims = 8;

// out block (column-major over odims) <- in block with strides istr per
// output dimension; optional conjugation and complex scale.
struct PermJob {
  const cplx* in;
  cplx* out;
  int64_t size;
  int32_t ndim;
  int32_t conj;
  int64_t odims[kMaxDims];
  int64_t istr[kMaxDims];
};

void launch_permute(const std::vector<PermJob>& jobs, DevBuf<PermJob>& djobs, DevBuf<int64_t>& dchunks,
                    hipStream_t s);

// out[perm[i]] = in[i]   (reference order -> block layout), FP64
void launch_scatter_map(const cplx* in, cplx* out, const int64_t* perm, int64_t n, hipStream_t s);
// out[i] = in[perm[i]]   (block layout -> reference order), FP64
void launch_gather_map(const cplx* in, cplx* out, const int64_t* perm, int64_t n, hipStream_t s);

void launch_to_fp32(const cplx* in, float2* out, int64_t n, hipStream_t s);
void launch_to_fp64(const float2* in, cplx* out, int64_t n, hipStream_t s);
void launch_zero(void* p, int64_t bytes, hipStream_t s);

// y = a*x + y (complex a); x,y FP64 complex
void launch_axpy(cplx a, const cplx* x, cplx* y, int64_t n, hipStream_t s);
// x *= a (complex)
void launch_scale(cplx a, cplx* x, int64_t n, hipStream_t s);
// y = x
void launch_copy(const cplx* x, cplx* y, int64_t n, hipStream_t s);

// Multi dot: c[j] = sum_i conj(V[i + j*ldv]) * w[i], j < k (k <= 64).
// partial: device scratch of size nblocks*k; result written to c (device).
void launch_multi_dot(const cplx* V, int64_t ldv, int k, const cplx* w, int64_t n, cplx* partial,
                      cplx* c, hipStream_t s);
// w -= sum_j V[:,j] * c[j]   (c on device)
void launch_multi_axpy(const cplx* V, int64_t ldv, int k, const cplx* c, cplx* w, int64_t n,
                       hipStream_t s);
// w = sum_j V[:,j] * c[j]   (c on device)
void launch_combine(const cplx* V, int64_t ldv, int k, const cplx* c, cplx* w, int64_t n,
                    hipStream_t s);
int multi_dot_blocks(int64_t n);
// The same on FP32 (split-complex float2) storage, FP64 accumulation:
// c = V^H w; w -= V c; and w (FP64) += V c.
void launch_multi_dot32(const float2* V, int64_t ldv, int k, const float2* w, int64_t n, cplx* partial, cplx* c,
                        hipStream_t s);
void launch_multi_axpy32(const float2* V, int64_t ldv, int k, const cplx* c, float2* w, int64_t n, hipStream_t s);
// c[i] *= w[i], i < k (k <= 64), on the device
void launch_coef_scale(cplx* c, const cplx* w, int k, hipStream_t s);
void launch_accumulate32(const float2* V, int64_t ldv, int k, const cplx* c, cplx* w, int64_t n, hipStream_t s);
// x[i] = (a, b), a and b uniform in [-1, 1) from a counter-based splitmix64
// stream (deterministic for a given seed)
void launch_random_fill(cplx* x, int64_t n, uint64_t seed, hipStream_t s);

}  // namespace cam
