// The visualisation step of Simulation::Update (ARD/simulation.cpp:320-420)
// on the device: one pressure plane through the scene, colour-mapped into a
// persistent pixel array.
//
// Per the reference: the plane is z = the first source's z (xy view) or x =
// its x (yz view) unless the caller gives one; partitions are drawn in list
// order (DCT partitions, then the PML slabs), later ones over earlier ones,
// each only if its [start, end] range along the view axis holds the plane;
// a partition reads its plane with the plane's GLOBAL coordinate as its local
// index (Partition::get_xy_plane / get_yz_plane, ARD/partition.cpp:58-78),
// through DctVolume::get_value (no bounds, ARD/dct_volume.cpp:70-73; a cell
// past the array, undefined there, is drawn as pressure 0 here) or the PML's
// GetIndex (out of range reads the zero sentinel, ARD/pml_partition.cpp:6-15).
// Colour (simulation.cpp:346-364): norm = 0.5 clamp(10 p, -1, 1) + 0.5;
// norm >= 0.5: r = g = 255 - round(510 (norm - 0.5)), b = 255; else r = 255,
// g = b = 255 - round(255 (1 - 2 norm)); pixel = SDL_MapRGBA(RGBA8888, 255,
// r, g, b) = 0xFF000000 | r << 16 | g << 8 | b. The xy view writes pixel
// (y - y_start) * size_x + (x - x_start); the yz view (z - z_start) * size_y +
// (y - y_start), into the same array (ARD/simulation.cpp:405-413).
//
// The batched form keeps no dense pressure in DCT partitions, so their plane
// is evaluated from the modes: a contraction of the modes with the REDFT01
// weights at the plane's local coordinate, then the 2D inverse transform of
// the plane (two GEMM launches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "ard_ctx_impl.h"
#include "ard_kernels.h"

namespace ard {

namespace {

// xy plane of a partition's modes at local z: P[ky][kx] = sum_kz wz[kz] M[kz][ky][kx]
__global__ __launch_bounds__(256) void k_plane_z(const float* __restrict__ M, const float* __restrict__ wz, int W,
                                                 int H, int D, int Wp, float* __restrict__ out) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)W * H) return;
  const int ky = int(e / W), kx = int(e % W);
  const float* m = M + (long long)ky * Wp + kx;
  const long long plane = (long long)H * Wp;
  float acc = 0.f;
  for (int kz = 0; kz < D; ++kz) acc += wz[kz] * m[kz * plane];
  out[e] = acc;
}

// yz plane at local x: P[kz][ky] = sum_kx wx[kx] M[kz][ky][kx], one wave per row
__global__ __launch_bounds__(256) void k_plane_x(const float* __restrict__ M, const float* __restrict__ wx, int W,
                                                 int H, int D, int Wp, float* __restrict__ out) {
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= (long long)D * H) return;
  const float* m = M + row * Wp;
  float acc = 0.f;
  for (int kx = lane; kx < W; kx += 64) acc += wx[kx] * m[kx];
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) acc += __shfl_xor(acc, s);
  if (lane == 0) out[row] = acc;
}

struct ViewSrc {
  const float* src;  // null: every cell reads as 0
  long long base, si, sj, limit;
  int rows, cols;
  long long row_off, col_off, stride;
};

// Colour of one cell, in the reference's FP32 operation order (no contraction).
__device__ unsigned int colour(float p) {
  float v = __fmul_rn(p, 10.f);
  v = (v < 1.f) ? v : 1.f;     // std::min(1.0f, v)
  v = (-1.f < v) ? v : -1.f;   // std::max(-1.0f, v)
  const float norm = __fadd_rn(__fmul_rn(0.5f, v), 0.5f);
  unsigned int r, g, b;
  if (norm >= 0.5f) {
    r = g = unsigned(int(255.f - roundf(__fmul_rn(510.f, __fsub_rn(norm, 0.5f)))));
    b = 255u;
  } else {
    r = 255u;
    g = b = unsigned(int(255.f - roundf(__fmul_rn(255.f, __fsub_rn(1.f, __fmul_rn(2.f, norm))))));
  }
  return 0xFF000000u | (r << 16) | (g << 8) | b;
}

__global__ __launch_bounds__(256) void k_render(ViewSrc s, unsigned int* __restrict__ pixels) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)s.rows * s.cols) return;
  const int i = int(e / s.cols), j = int(e % s.cols);
  const long long idx = s.base + i * s.si + j * s.sj;
  const float p = (s.src && idx >= 0 && idx < s.limit) ? s.src[idx] : 0.f;
  pixels[(s.row_off + i) * s.stride + s.col_off + j] = colour(p);
}

std::vector<float> redft01_row(int n, int x) {
  const double pi = 3.14159265358979323846;
  std::vector<float> w(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) w[size_t(k)] = float((k == 0 ? 1.0 : 2.0) * std::cos(pi * k * (x + 0.5) / n));
  return w;
}

long long cdiv(long long a, long long b) { return (a + b - 1) / b; }
}  // namespace

struct ViewState {
  unsigned int* d_pixels = nullptr;
  long long npix = 0;
  int box[6] = {0, 0, 0, 0, 0, 0};  // x_start, y_start, z_start, x_end, y_end, z_end
  float* d_scratch = nullptr;       // plane evaluation of the batched form
  size_t scratch_n = 0;
};

void destroy_view(Ctx& c) {
  delete c.view;
  c.view = nullptr;
}

int render_view(Ctx& c, int view, int coord, unsigned int* pixels, int* dims) {
  if (!c.view) {
    c.view = new ViewState();
    ViewState& V = *c.view;
    // bounding box of all partitions, PML slabs included (ARD/simulation.cpp:272-288)
    int b[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    for (const PartGeom& p : c.topo.parts) {
      b[0] = std::min(b[0], p.x0); b[1] = std::min(b[1], p.y0); b[2] = std::min(b[2], p.z0);
      b[3] = std::max(b[3], p.x1()); b[4] = std::max(b[4], p.y1()); b[5] = std::max(b[5], p.z1());
    }
    std::copy(b, b + 6, V.box);
    const long long sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
    // one array for both views, as the reference's pixels_ (sized for the
    // larger of them: the reference's own is size_x * size_y)
    V.npix = std::max(sx, sz) * sy;
    int r;
    if ((r = c.dalloc((void**)&V.d_pixels, size_t(V.npix) * 4)) || (r = c.dzero(V.d_pixels, size_t(V.npix) * 4)))
      return r;
  }
  ViewState& V = *c.view;
  const int* b = V.box;
  const long long sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
  if (view == ARD_VIEW_XY) {
    dims[0] = int(sx); dims[1] = int(sy);
  } else {
    dims[0] = int(sy); dims[1] = int(sz);
  }
  if (!pixels) return ARD_OK;  // a query of the array's shape
  if (coord < 0) {
    if (c.topo.sources.empty()) return set_error(ARD_ERR_STATE, "no source to place the plane at (give coord)");
    coord = view == ARD_VIEW_XY ? c.topo.sources[0].z : c.topo.sources[0].x;
  }
  const int par = c.time_step & 1;
  const bool batched = c.prm.form == ARD_FORM_BATCHED;
  int r;
  // batched DCT partitions: evaluate each drawn plane into its own scratch slot
  struct Job {
    int part;
    ViewSrc s;
    long long eval_off = -1;  // scratch offset of the evaluated plane, or -1
    int local = 0;            // local coordinate of the evaluated plane
  };
  std::vector<Job> jobs;
  size_t scratch = 0;
  for (int i = 0; i < int(c.topo.parts.size()); ++i) {
    const PartGeom& p = c.topo.parts[size_t(i)];
    const int lo = view == ARD_VIEW_XY ? p.z0 : p.x0, hi = view == ARD_VIEW_XY ? p.z1() : p.x1();
    if (lo > coord || hi < coord) continue;
    Job j{};
    j.part = i;
    ViewSrc& s = j.s;
    const long long HW = (long long)p.h * p.w, N = HW * p.d;
    if (view == ARD_VIEW_XY) {
      s.rows = p.h; s.cols = p.w; s.row_off = p.y0 - b[1]; s.col_off = p.x0 - b[0]; s.stride = sx;
    } else {
      s.rows = p.d; s.cols = p.h; s.row_off = p.z0 - b[2]; s.col_off = p.y0 - b[1]; s.stride = sy;
    }
    if (p.is_pml) {  // GetIndex: a coordinate outside the slab reads the zero sentinel
      const int dim = view == ARD_VIEW_XY ? p.d : p.w;
      s.src = (coord >= 0 && coord < dim) ? c.pml[size_t(i)].p[par] : nullptr;
      s.base = view == ARD_VIEW_XY ? coord * HW : coord;
      s.si = view == ARD_VIEW_XY ? p.w : HW;
      s.sj = view == ARD_VIEW_XY ? 1 : p.w;
      s.limit = N;
    } else if (!batched) {  // the dense field, addressed as get_value does
      s.src = c.dct[size_t(i)].p;
      s.base = view == ARD_VIEW_XY ? coord * HW : coord;
      s.si = view == ARD_VIEW_XY ? p.w : HW;
      s.sj = view == ARD_VIEW_XY ? 1 : p.w;
      s.limit = N;
    } else if (view == ARD_VIEW_XY) {  // plane at local z = coord, if inside the array
      s.base = 0; s.si = p.w; s.sj = 1; s.limit = HW;
      if (coord < p.d) {
        j.eval_off = (long long)scratch;
        j.local = coord;
        scratch += size_t(2 * HW);
      }
    } else {  // cells i*HW + j*W + coord: the plane at x' = coord % W, shifted by coord / W rows
      s.base = coord / p.w; s.si = p.h; s.sj = 1; s.limit = (long long)p.d * p.h;
      j.eval_off = (long long)scratch;
      j.local = coord % p.w;
      scratch += size_t(2 * (long long)p.d * p.h);
    }
    jobs.push_back(j);
  }
  if (scratch > 0) {
    // per plane: [weights | GEMM descriptors] then two plane buffers
    std::vector<float> wts;
    std::vector<GemmDesc> descs;
    std::vector<long long> woff;
    for (const Job& j : jobs)
      if (j.eval_off >= 0) {
        const PartGeom& p = c.topo.parts[size_t(j.part)];
        woff.push_back((long long)wts.size());
        std::vector<float> w = redft01_row(view == ARD_VIEW_XY ? p.d : p.w, j.local);
        wts.insert(wts.end(), w.begin(), w.end());
      }
    const size_t nw = (wts.size() + 63) & ~size_t(63);
    const size_t need = scratch + nw + (woff.size() * 2 * sizeof(GemmDesc) + 3) / 4;
    if (V.scratch_n < need) {
      if (V.d_scratch) {
        hipError_t e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        hipFree(V.d_scratch);
        c.allocs.erase(std::find(c.allocs.begin(), c.allocs.end(), (void*)V.d_scratch));
        V.d_scratch = nullptr;
        V.scratch_n = 0;
      }
      if ((r = c.dalloc((void**)&V.d_scratch, need * 4))) return r;
      V.scratch_n = need;
    }
    float* d_w = V.d_scratch + scratch;
    GemmDesc* d_g = reinterpret_cast<GemmDesc*>(d_w + nw);
    size_t k = 0;
    for (Job& j : jobs) {
      if (j.eval_off < 0) continue;
      const PartGeom& p = c.topo.parts[size_t(j.part)];
      float* P0 = V.d_scratch + j.eval_off;
      const long long plane = view == ARD_VIEW_XY ? (long long)p.h * p.w : (long long)p.d * p.h;
      float* T = P0 + plane;
      const float *fa, *ia, *fb, *ib;  // inner axis (x or y), outer axis (y or z)
      if (view == ARD_VIEW_XY) {
        if ((r = c.cos_tables(p.w, &fa, &ia)) || (r = c.cos_tables(p.h, &fb, &ib))) return r;
      } else {
        if ((r = c.cos_tables(p.h, &fa, &ia)) || (r = c.cos_tables(p.d, &fb, &ib))) return r;
      }
      const int n_in = view == ARD_VIEW_XY ? p.w : p.h, n_out = view == ARD_VIEW_XY ? p.h : p.d;
      // along the inner axis: T[row][x] = sum_k ia[x][k] P0[row][k]
      GemmDesc g1{};
      g1.A = ia; g1.B = P0; g1.C = T;
      g1.M = n_in; g1.K = n_in; g1.N = n_out; g1.N1 = n_out;
      g1.bK = 1; g1.b1 = n_in; g1.cM = 1; g1.c1 = n_in; g1.alpha = 1.f;
      g1.tiles_n = int(cdiv(g1.N, 64));
      // along the outer axis, back into P0: P0[y][x] = sum_k ib[y][k] T[k][x]
      GemmDesc g2{};
      g2.A = ib; g2.B = T; g2.C = P0;
      g2.M = n_out; g2.K = n_out; g2.N = n_in; g2.N1 = n_in;
      g2.bK = n_in; g2.b1 = 1; g2.cM = n_in; g2.c1 = 1; g2.alpha = 1.f;
      g2.tiles_n = int(cdiv(g2.N, 64));
      descs.push_back(g1);
      descs.push_back(g2);
      j.s.src = P0;
    }
    if ((r = c.upload(d_w, wts.data(), wts.size() * 4)) ||
        (r = c.upload(d_g, descs.data(), descs.size() * sizeof(GemmDesc))))
      return r;
    for (Job& j : jobs) {
      if (j.eval_off < 0) continue;
      const PartGeom& p = c.topo.parts[size_t(j.part)];
      const float* Mp[2];
      int wp;
      if ((r = batched_modes(c, j.part, Mp, &wp))) return r;
      float* P0 = V.d_scratch + j.eval_off;
      if (view == ARD_VIEW_XY)
        hipLaunchKernelGGL(k_plane_z, dim3(unsigned(cdiv((long long)p.w * p.h, 256))), dim3(256), 0, c.stream,
                           Mp[par], d_w + woff[k], p.w, p.h, p.d, wp, P0);
      else
        hipLaunchKernelGGL(k_plane_x, dim3(unsigned(cdiv((long long)p.d * p.h, 4))), dim3(256), 0, c.stream,
                           Mp[par], d_w + woff[k], p.w, p.h, p.d, wp, P0);
      const GemmDesc& g1 = descs[2 * k];
      const GemmDesc& g2 = descs[2 * k + 1];
      if (launch_gemm(d_g + 2 * k, 1, int(cdiv(g1.M, 64)) * g1.tiles_n, true, true, 0, c.stream) ||
          launch_gemm(d_g + 2 * k + 1, 1, int(cdiv(g2.M, 64)) * g2.tiles_n, false, false, 0, c.stream))
        return set_error(ARD_ERR_HIP, "plane transform launch");
      ++k;
    }
  }
  for (const Job& j : jobs)
    hipLaunchKernelGGL(k_render, dim3(unsigned(cdiv((long long)j.s.rows * j.s.cols, 256))), dim3(256), 0, c.stream,
                       j.s, V.d_pixels);
  if (hipGetLastError() != hipSuccess) return set_error(ARD_ERR_HIP, "render launch");
  hipError_t e = hipMemcpyAsync(pixels, V.d_pixels, size_t(dims[0]) * dims[1] * 4, hipMemcpyDeviceToHost, c.stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
  e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

}  // namespace ard
