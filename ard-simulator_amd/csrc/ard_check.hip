// Non-finite state check (ard_check_finite): one launch reads every state
// array of the context (DCT force, pressure and modes, PML pressure, phi and
// force; in the batched form the modes and the face-slab force and pressure)
// and reports the lowest partition id holding a NaN or an infinity. The
// reference has no such check: a diverging run fills its traces with NaN and
// a bad recorder dereferences a null partition (ARD/recorder.cpp:50).
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "ard_ctx_impl.h"

namespace ard {

namespace {

struct CheckSeg {
  const float* p;
  long long n;      // floats
  long long blk0;   // first block of the segment in the launch
  int part;
};

constexpr int kPerThread = 8;  // floats per thread, 256 threads per block

__global__ __launch_bounds__(256) void k_check_finite(const CheckSeg* __restrict__ segs, int nseg,
                                                      int* __restrict__ bad) {
  const long long b = blockIdx.x;
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {  // the segment of this block (block-uniform)
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const CheckSeg& s = segs[lo];
  const long long base = (b - s.blk0) * 256LL * kPerThread + threadIdx.x;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const long long i = base + 256LL * j;
    if (i < s.n) ok &= __builtin_isfinite(s.p[i]);
  }
  if (!ok) atomicMin(bad, s.part);
}

__global__ void k_fill_int(int* p, int v) { *p = v; }

void add_seg(std::vector<CheckSeg>& v, const float* p, long long n, int part) {
  if (p && n > 0) v.push_back(CheckSeg{p, n, 0, part});
}

}  // namespace

int check_finite(Ctx& c, int* bad_part) {
  std::vector<CheckSeg> segs;
  for (int i = 0; i < int(c.topo.parts.size()); ++i) {
    const PartGeom& g = c.topo.parts[size_t(i)];
    const long long n = (long long)g.w * g.h * g.d;
    if (g.is_pml) {
      const PmlBuffers& B = c.pml[size_t(i)];
      for (int k = 0; k < 2; ++k) {
        add_seg(segs, B.p[k], n, i);
        for (int a = 0; a < 3; ++a) add_seg(segs, B.phi[k][a], n, i);
      }
      add_seg(segs, B.f, n, i);
    } else if (c.bs) {
      std::vector<std::pair<const float*, long long>> arr;
      batched_state_arrays(c, i, arr);
      for (auto& a : arr) add_seg(segs, a.first, a.second, i);
    } else {
      const DctPlain& D = c.dct[size_t(i)];
      add_seg(segs, D.F, n, i);
      add_seg(segs, D.p, n, i);
      add_seg(segs, D.M[0], n, i);
      add_seg(segs, D.M[1], n, i);
    }
  }
  long long blocks = 0;
  for (CheckSeg& s : segs) {
    s.blk0 = blocks;
    blocks += (s.n + 256LL * kPerThread - 1) / (256LL * kPerThread);
  }
  *bad_part = -1;
  if (segs.empty()) return ARD_OK;
  if (blocks >= (1LL << 31)) return set_error(ARD_ERR_UNSUPPORTED, "state too large for one check launch");
  CheckSeg* d_segs = nullptr;
  int* d_bad = nullptr;
  hipError_t e = hipMalloc(&d_segs, segs.size() * sizeof(CheckSeg));
  if (e == hipSuccess) e = hipMalloc(&d_bad, sizeof(int));
  if (e == hipSuccess) e = hipMemcpyAsync(d_segs, segs.data(), segs.size() * sizeof(CheckSeg), hipMemcpyHostToDevice,
                                          c.stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, c.stream, d_bad, INT_MAX);
    hipLaunchKernelGGL(k_check_finite, dim3(unsigned(blocks)), dim3(256), 0, c.stream, d_segs, int(segs.size()), d_bad);
    e = hipGetLastError();
  }
  int bad = INT_MAX;
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  if (d_segs) (void)hipFree(d_segs);
  if (d_bad) (void)hipFree(d_bad);
  if (e != hipSuccess) return hip_fail(e, "non-finite state check");
  *bad_part = bad == INT_MAX ? -1 : bad;
  return ARD_OK;
}

}  // namespace ard
