// Streaming floor of the modal pass's state traffic: the access pattern of
// k_modal alone (no spectra, no reductions). Blocks of 4 waves own a 64 x 16
// tile of (kx, ky) and walk 48 kz planes; per plane each wave reads 4 rows x
// 64 floats of two arrays (the current and previous modes) and writes 4 rows
// of one (the new modes over the previous), non-temporal, one plane ahead in
// registers: 12 bytes per cell. Layout 0: rows of pitch W, as the mode arrays
// are stored; layout 1: each (plane, tile) contiguous. LDS per block sets the
// blocks resident per CU. 74.3M cells, grid = nbx * nby * nbz blocks over
// W = 64 nbx, H = 16 nby, D = 48 nbz, every index in bounds.
#include <hip/hip_runtime.h>

#include <cstdio>

template <int LAYOUT, int LDSKB>
__global__ __launch_bounds__(256) void k_stream(const float* __restrict__ a, float* __restrict__ b, int nbx, int nby) {
  __shared__ float pad[LDSKB * 256];
  const int bid = blockIdx.x;
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long W = 64LL * nbx, H = 16LL * nby;
  auto idx = [&](int kz, int r) -> long long {
    const int row = wave * 4 + r;
    if (LAYOUT == 0) return ((long long)kz * H + by * 16 + row) * W + bx * 64 + lane;
    return ((((long long)kz * nby + by) * nbx + bx) * 16 + row) * 64 + lane;
  };
  float acc = 0.f;
  float ca[4], cb[4], na[4], nb[4];
  const int kz0 = bz * 48;
  for (int r = 0; r < 4; ++r) {
    ca[r] = __builtin_nontemporal_load(&a[idx(kz0, r)]);
    cb[r] = __builtin_nontemporal_load(&b[idx(kz0, r)]);
  }
  for (int kz = kz0; kz < kz0 + 48; ++kz) {
    const int kn = kz + 1 < kz0 + 48 ? kz + 1 : kz;
    for (int r = 0; r < 4; ++r) {
      na[r] = __builtin_nontemporal_load(&a[idx(kn, r)]);
      nb[r] = __builtin_nontemporal_load(&b[idx(kn, r)]);
    }
    for (int r = 0; r < 4; ++r) {
      const float v = 1.998f * ca[r] - 0.999f * cb[r];
      __builtin_nontemporal_store(v, &b[idx(kz, r)]);
      acc += v;
    }
    for (int r = 0; r < 4; ++r) {
      ca[r] = na[r];
      cb[r] = nb[r];
    }
  }
  if (acc == 12345.f) pad[threadIdx.x] = acc;  // keeps the LDS allocation
  if (acc == 12345.f) b[0] = pad[(threadIdx.x + 1) & 255];
}

int main() {
  const int nbx = 7, nby = 6, nbz = 36;  // 448 x 96 x 1728 = 74.3M cells
  const size_t n = size_t(64 * nbx) * (16 * nby) * (48 * nbz);
  float *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 4) != hipSuccess || hipMalloc(&b, n * 4) != hipSuccess) return 1;
  if (hipMemset(a, 0, n * 4) != hipSuccess || hipMemset(b, 0, n * 4) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  const int grid = nbx * nby * nbz;
  auto run = [&](void (*kern)(const float*, float*, int, int), const char* name) {
    for (int i = 0; i < 3; ++i) hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, 0, a, b, nbx, nby);
    (void)hipEventRecord(e0);
    for (int i = 0; i < 10; ++i) hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, 0, a, b, nbx, nby);
    (void)hipEventRecord(e1);
    if (hipEventSynchronize(e1) != hipSuccess) return false;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= 10;
    std::printf("%-34s %8.1f us  %6.2f TB/s (12 B/cell)\n", name, ms * 1e3, 12.0 * n / (ms * 1e-3) / 1e12);
    return true;
  };
  if (!run(k_stream<0, 40>, "pitch rows, 40 KB LDS (3 blk/CU)") || !run(k_stream<1, 40>, "tiled, 40 KB LDS (3 blk/CU)") ||
      !run(k_stream<0, 30>, "pitch rows, 30 KB LDS (5 blk/CU)") || !run(k_stream<1, 30>, "tiled, 30 KB LDS (5 blk/CU)") ||
      !run(k_stream<0, 12>, "pitch rows, 12 KB LDS (8 blk/CU)") || !run(k_stream<1, 12>, "tiled, 12 KB LDS (8 blk/CU)"))
    return 1;
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}
