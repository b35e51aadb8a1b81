// Standalone scaled 3D DCT-II / DCT-III of a device volume: the reference's
// GPU seam VkFFT_DCT(vkGPU, dctType, w, h, d, in, out).execute()
// (ARD/utils_VkFFT.h:38-60, ARD/utils_VkFFT.cpp:623-691) together with the
// scaling DctVolume applies after it (ARD/dct_volume.cpp:38-68), on device
// pointers and a HIP stream instead of host arrays and Vulkan staging copies.
//
// Three per-axis passes on the matrix cores (the plain form's k_gemm_mfma):
// x: in -> out, y: out -> work volume, z: work volume -> out (scaled). The
// workspace holds the intermediate volume, the three cosine matrices of the
// pass (FFTW's REDFT10 or REDFT01 kernel, ARD/dct_volume.cpp:17,19, built on
// the device in FP64 and stored FP32) and the three pass descriptors. It is
// allocated per call by ard_dct3d and is the file-scope device array
// `scratch` for ard_dct3d_noalloc, which allocates nothing.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ard.h"
#include "ard_ctx_impl.h"
#include "ard_kernels.h"

// ard_dct3d_noalloc's workspace: one copy per GPU (at its own address on
// each), held until the process ends. In host code this name is a stand-in,
// not a device address: host code uses the per-device pointer that
// hipGetSymbolAddress returns (g_scratch below), never the name itself.
static __device__ __align__(256) char scratch[ARD_DCT3D_SCRATCH_BYTES];

namespace ard {
namespace {

constexpr size_t kAlign = 256;
size_t up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

struct Dct3dLayout {
  size_t work, tab[3], desc, total;  // byte offsets into the workspace, total size
};

Dct3dLayout layout(int w, int h, int d) {
  Dct3dLayout L;
  size_t o = 0;
  L.work = o;
  o += up(size_t(w) * h * d * 4);
  const int n[3] = {w, h, d};
  for (int a = 0; a < 3; ++a) {
    L.tab[a] = o;
    o += up(size_t(n[a]) * n[a] * 4);
  }
  L.desc = o;
  o += up(3 * sizeof(GemmDesc));
  L.total = o;
  return L;
}

// Cosine matrix of one axis: A[k][j] = 2 cos(pi k (j + 1/2) / n) (REDFT10) or
// A[k][j] = j ? 2 cos(pi j (k + 1/2) / n) : 1 (REDFT01), in FP64.
__global__ void k_dct_table(float* __restrict__ A, int n, int inverse) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * n) return;
  const int k = int(i / n), j = int(i % n);
  double v;
  if (!inverse) v = 2.0 * cospi(double(k) * (j + 0.5) / n);
  else v = j == 0 ? 1.0 : 2.0 * cospi(double(j) * (k + 0.5) / n);
  A[i] = float(v);
}

struct ThreeDescs {
  GemmDesc g[3];
};

// The pass descriptors go to device memory through a kernel argument (no host
// copy, no allocation).
__global__ void k_put_descs(GemmDesc* __restrict__ dst, ThreeDescs v) {
  if (threadIdx.x < 3) dst[threadIdx.x] = v.g[threadIdx.x];
}

}  // namespace

size_t dct3d_workspace_bytes(int w, int h, int d) { return layout(w, h, d).total; }

// Enqueues the transform on stream s with the workspace at ws (device
// memory of at least dct3d_workspace_bytes). Asynchronous.
int dct3d_enqueue(char* ws, int type, int w, int h, int d, const float* in, float* out, float scale,
                  hipStream_t s) {
  const Dct3dLayout L = layout(w, h, d);
  float* work = reinterpret_cast<float*>(ws + L.work);
  const int n[3] = {w, h, d};
  float* tab[3];
  for (int a = 0; a < 3; ++a) {
    tab[a] = reinterpret_cast<float*>(ws + L.tab[a]);
    const long long cnt = (long long)n[a] * n[a];
    hipLaunchKernelGGL(k_dct_table, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, s, tab[a], n[a], type == 3);
  }
  const long long W = w, H = h, D = d;
  ThreeDescs v{};
  // x: B(k = x, n = row) contiguous in k; C(m = kx, n = row) contiguous in m
  GemmDesc& gx = v.g[0];
  gx.A = tab[0]; gx.B = in; gx.C = out;
  gx.M = w; gx.K = w; gx.N = int(H * D); gx.N1 = gx.N;
  gx.bK = 1; gx.b1 = W; gx.cM = 1; gx.c1 = W;
  // y: per z plane, n = (x, z)
  GemmDesc& gy = v.g[1];
  gy.A = tab[1]; gy.B = out; gy.C = work;
  gy.M = h; gy.K = h; gy.N = int(W * D); gy.N1 = w;
  gy.bK = W; gy.b1 = 1; gy.b2 = W * H; gy.cM = W; gy.c1 = 1; gy.c2 = W * H;
  // z: n = (x, y) flattened
  GemmDesc& gz = v.g[2];
  gz.A = tab[2]; gz.B = work; gz.C = out;
  gz.M = d; gz.K = d; gz.N = int(W * H); gz.N1 = gz.N;
  gz.bK = W * H; gz.b1 = 1; gz.cM = W * H; gz.c1 = 1;
  for (int a = 0; a < 3; ++a) {
    v.g[a].alpha = a == 2 ? scale : 1.f;
    v.g[a].tiles_n = int((v.g[a].N + 63) / 64);
  }
  GemmDesc* dd = reinterpret_cast<GemmDesc*>(ws + L.desc);
  hipLaunchKernelGGL(k_put_descs, dim3(1), dim3(64), 0, s, dd, v);
  if (hipGetLastError() != hipSuccess) return set_error(ARD_ERR_HIP, "DCT table launch");
  int r = 0;
  r |= launch_gemm(dd + 0, 1, int(((w + 63) / 64) * v.g[0].tiles_n), true, true, 0, s, 1);
  r |= launch_gemm(dd + 1, 1, int(((h + 63) / 64) * v.g[1].tiles_n), false, false, 0, s, 1);
  r |= launch_gemm(dd + 2, 1, int(((d + 63) / 64) * v.g[2].tiles_n), false, false, 0, s, 1);
  return r ? set_error(ARD_ERR_HIP, "DCT pass launch") : ARD_OK;
}

int dct3d_check(int type, int w, int h, int d, const float* in, float* out) {
  if (type != 2 && type != 3) return set_error(ARD_ERR_INVALID, "type must be 2 (DCT-II) or 3 (DCT-III)");
  if (w <= 0 || h <= 0 || d <= 0 || !in || !out) return set_error(ARD_ERR_INVALID, "bad arguments");
  const long long n = (long long)w * h * d;
  if (n >= (1LL << 31) || (long long)w * w >= (1LL << 31) || (long long)h * h >= (1LL << 31) ||
      (long long)d * d >= (1LL << 31))
    return set_error(ARD_ERR_INVALID, "volume too large");
  const char* a = reinterpret_cast<const char*>(in);
  const char* b = reinterpret_cast<const char*>(out);
  if (a < b + n * 4 && b < a + n * 4) return set_error(ARD_ERR_INVALID, "in and out overlap");
  return ARD_OK;
}

}  // namespace ard

extern "C" int ard_dct3d(int device, int type, int w, int h, int d, const float* in, float* out, float scale,
                         void* stream) {
  using namespace ard;
  int r;
  if ((r = dct3d_check(type, w, h, d, in, out))) return r;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = nullptr;
  if ((e = hipMalloc((void**)&ws, dct3d_workspace_bytes(w, h, d))) != hipSuccess) return hip_fail(e, "hipMalloc");
  r = dct3d_enqueue(ws, type, w, h, d, in, out, scale, s);
  e = hipStreamSynchronize(s);
  hipFree(ws);
  if (r) return r;
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};  // per device: the address of `scratch` there
}  // namespace

// No counterpart in the reference: ard_dct3d for callers that cannot allocate
// (graph capture, serving loops). The workspace is `scratch`; its device
// address is taken once per device, with that device current, on the first
// call there. Asynchronous on `stream`. Every call on one GPU shares that
// GPU's scratch, so calls that may overlap on one GPU (other streams, other
// host threads) race: serialise them. Everything the workspace holds is
// rewritten by each call before it is read (tables, descriptors, volume).
extern "C" int ard_dct3d_noalloc(int device, int type, int w, int h, int d, const float* in, float* out, float scale,
                                 void* stream) {
  using namespace ard;
  int r;
  if ((r = dct3d_check(type, w, h, d, in, out))) return r;
  if (dct3d_workspace_bytes(w, h, d) > size_t(ARD_DCT3D_SCRATCH_BYTES))
    return set_error(ARD_ERR_RANGE, "volume too large for ard_dct3d_noalloc's workspace (ARD_DCT3D_SCRATCH_BYTES)");
  if (device < 0 || device >= kMaxDevices) return set_error(ARD_ERR_INVALID, "no such device");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  char* ws;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
      void* p = nullptr;
      if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess || !p)
        return hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)");
      g_scratch[device] = static_cast<char*>(p);
    }
    ws = g_scratch[device];
  }
  return dct3d_enqueue(ws, type, w, h, d, in, out, scale, static_cast<hipStream_t>(stream));
}
