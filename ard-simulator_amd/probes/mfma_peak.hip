// Matrix-core throughput of the three instructions the transform kernels
// use, on this GPU: v_mfma_f32_32x32x2_f32 (k_gemm_mfma, k_gemm_mfma_fold),
// v_mfma_f32_16x16x4_f32 (k_fold_direct) and v_mfma_f32_32x32x16_bf16
// (k_gemm_bf16x3_fold). Each wave runs four independent accumulation chains
// on register operands (no memory traffic in the loop), 8 waves per SIMD;
// one store per thread at the end keeps the work. Prints TFLOP/s per
// instruction (2 x M x N x K flops each), two timed launches each.
#include <hip/hip_runtime.h>

#include <cstdio>

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void k_f32_32x32x2(float* out, int iters) {
  f32x16 a0 = {}, a1 = {}, a2 = {}, a3 = {};
  const float x = threadIdx.x * 1e-3f, y = blockIdx.x * 1e-3f;
  for (int i = 0; i < iters; ++i) {
    a0 = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, a0, 0, 0, 0);
    a1 = __builtin_amdgcn_mfma_f32_32x32x2f32(y, x, a1, 0, 0, 0);
    a2 = __builtin_amdgcn_mfma_f32_32x32x2f32(x, x, a2, 0, 0, 0);
    a3 = __builtin_amdgcn_mfma_f32_32x32x2f32(y, y, a3, 0, 0, 0);
  }
  float s = 0.f;
  for (int r = 0; r < 16; ++r) s += a0[r] + a1[r] + a2[r] + a3[r];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_f32_16x16x4(float* out, int iters) {
  f32x4 a0 = {}, a1 = {}, a2 = {}, a3 = {};
  const float x = threadIdx.x * 1e-3f, y = blockIdx.x * 1e-3f;
  for (int i = 0; i < iters; ++i) {
    a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, a0, 0, 0, 0);
    a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(y, x, a1, 0, 0, 0);
    a2 = __builtin_amdgcn_mfma_f32_16x16x4f32(x, x, a2, 0, 0, 0);
    a3 = __builtin_amdgcn_mfma_f32_16x16x4f32(y, y, a3, 0, 0, 0);
  }
  float s = 0.f;
  for (int r = 0; r < 4; ++r) s += a0[r] + a1[r] + a2[r] + a3[r];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_bf16_32x32x16(float* out, int iters) {
  f32x16 a0 = {}, a1 = {}, a2 = {}, a3 = {};
  bf16x8 x, y;
  for (int i = 0; i < 8; ++i) {
    x[i] = (__bf16)(threadIdx.x * 1e-3f + i);
    y[i] = (__bf16)(blockIdx.x * 1e-3f - i);
  }
  for (int i = 0; i < iters; ++i) {
    a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, a0, 0, 0, 0);
    a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(y, x, a1, 0, 0, 0);
    a2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, x, a2, 0, 0, 0);
    a3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(y, y, a3, 0, 0, 0);
  }
  float s = 0.f;
  for (int r = 0; r < 16; ++r) s += a0[r] + a1[r] + a2[r] + a3[r];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

int main() {
  const int blocks = 2048, iters = 4000;  // 8192 waves: 8 per SIMD of 256 CUs
  float* d = nullptr;
  if (hipMalloc(&d, size_t(blocks) * 256 * 4) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  struct Case {
    const char* name;
    void (*kern)(float*, int);
    double flops;  // per instruction
  } cases[] = {{"v_mfma_f32_32x32x2_f32", k_f32_32x32x2, 2.0 * 32 * 32 * 2},
               {"v_mfma_f32_16x16x4_f32", k_f32_16x16x4, 2.0 * 16 * 16 * 4},
               {"v_mfma_f32_32x32x16_bf16", k_bf16_32x32x16, 2.0 * 32 * 32 * 16}};
  for (const Case& k : cases) {
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipEventRecord(e0);
      hipLaunchKernelGGL(k.kern, dim3(blocks), dim3(256), 0, 0, d, iters);
      (void)hipEventRecord(e1);
      if (hipEventSynchronize(e1) != hipSuccess) return 1;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double fl = double(blocks) * 4 * iters * 4 * k.flops;
      std::printf("%-26s launch %d: %8.3f ms  %7.1f TFLOP/s\n", k.name, rep, ms, fl / ms / 1e9);
    }
  }
  (void)hipFree(d);
  return 0;
}
