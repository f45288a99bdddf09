// dice_canary.hip -- level-0 compiler canaries for the exactness contract.
//
// The bit-exact path needs every binary64 and binary32 operation rounded
// once, in source order: no contraction of a*b+c into a fused multiply-add,
// correctly rounded division, narrowing f64 -> f32 to nearest, and float32
// denormals kept (docs/cuda-decision-record.md:23-40 of the reference: the
// CUDA backend probes (1+2^-29)(1-2^-29)-1 == 0 and checks f64 mul-add,
// division and narrowing on 100k samples; tests/test_fast_cpu_bitexact.py:
// 347-375 is the CPU analogue).  These kernels evaluate exactly those
// operations on operands read from device memory, so nothing can be folded
// at compile time, and the host compares every result with numpy.
//
// The multiply-add kernels hold no other arithmetic: the only way a fused
// instruction can appear in them is contraction, which the build forbids.
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace dice {

__global__ void k_canary_mul_add_f64(const double* __restrict__ a, const double* __restrict__ b,
                                     const double* __restrict__ c, double* __restrict__ out,
                                     int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * b[i] + c[i];
}

__global__ void k_canary_mul_add_f32(const float* __restrict__ a, const float* __restrict__ b,
                                     const float* __restrict__ c, float* __restrict__ out,
                                     int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * b[i] + c[i];
}

__global__ void k_canary_rest(const double* __restrict__ a, const double* __restrict__ b,
                              const float* __restrict__ fa, const float* __restrict__ fb,
                              double* __restrict__ quotient, float* __restrict__ narrowed,
                              float* __restrict__ f32_quotient, float* __restrict__ f32_product,
                              int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  quotient[i] = a[i] / b[i];
  narrowed[i] = (float)a[i];
  f32_quotient[i] = fa[i] / fb[i];
  f32_product[i] = fa[i] * fb[i];
}

hipError_t launch_fp_canaries(int64_t n, const double* a, const double* b, const double* c,
                              const float* fa, const float* fb, const float* fc,
                              double* mul_add, double* quotient, float* narrowed,
                              float* f32_mul_add, float* f32_quotient, float* f32_product,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_canary_mul_add_f64, dim3(blocks), dim3(256), 0, s, a, b, c, mul_add, n);
  hipLaunchKernelGGL(k_canary_mul_add_f32, dim3(blocks), dim3(256), 0, s, fa, fb, fc,
                     f32_mul_add, n);
  hipLaunchKernelGGL(k_canary_rest, dim3(blocks), dim3(256), 0, s, a, b, fa, fb, quotient,
                     narrowed, f32_quotient, f32_product, n);
  return hipGetLastError();
}

}  // namespace dice
