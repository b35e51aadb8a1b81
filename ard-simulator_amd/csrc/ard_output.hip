// Recorder cubes on the device: the 10 x 10 x 10 cells around each recorder,
// sampled every step (the out_<id>.txt row that Recorder::RecordField writes,
// ARD/recorder.cpp:44-54). The cube of recorder r at step t lands in
// d_cube[t][r][1000], element e = 100 (i + 5) + 10 (j + 5) + (k + 5) for the
// cell (x + k, y + j, z + i), i outermost as in the reference's loops.
//
// Addressing follows the point sample: the recorder's resolved flat index f0
// into its owner's array (global coordinates into the array with
// ARD_REC_REFERENCE, ARD/dct_volume.cpp:70-73; the local cell with
// ARD_REC_LOCAL), and cell (i, j, k) is f0 + (i H + j) W + k.
//
// Plain form: a gather from the owner's dense pressure. Batched form: the
// pressure is not materialised, so the cube is evaluated from the owner's
// modes, p = sum M[kz][ky][kx] Rz(kz, z) Ry(ky, y) Rx(kx, x) with the REDFT01
// weights R(k, x) = (k ? 2 : 1) cos(pi k (x + 1/2) / n), separably: the x and
// y contractions of one kz plane per block on matrix cores
// (v_mfma_f32_16x16x4_f32, up to 16 distinct x and y values per cube), then
// the z contraction per cube cell in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ard_ctx_impl.h"

namespace ard {

namespace {
constexpr int kCube = 1000;
constexpr int kSet = 16;  // distinct coordinates per axis the batched evaluation takes

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct CubeRec {
  const float* field;  // plain: the owner's dense pressure
  const float* M[2];   // batched: the owner's mode buffers, rows padded to Wp
  long long f0;        // flat index of the recorder's cell in the owner's array
  int W, H, D, Wp;
  const float* cx;     // batched: [16][Wp] x weights at the distinct x values (zero rows and pads)
  const float* cy;     // [16][Hp] y weights, Hp = H rounded up to 16
  const float* cz;     // [16][D] z weights
  float* t2;           // [D][16 y][16 x] partial sums after the x and y contractions
  const int* sel;      // [1000] ix | iy << 4 | iz << 8 of each cube cell
  long long block0;    // first block of this cube in the contraction grid
};

__global__ __launch_bounds__(256) void k_cube_gather(const CubeRec* __restrict__ r, int nrec, float* __restrict__ out,
                                                     int cap, const int* __restrict__ d_step) {
  const int t = *d_step;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (t >= cap || e >= kCube) return;
  const CubeRec& R = r[blockIdx.y];
  const int i = e / 100 - 5, j = (e / 10) % 10 - 5, k = e % 10 - 5;
  const long long idx = R.f0 + ((long long)i * R.H + j) * R.W + k;
  out[((long long)t * nrec + blockIdx.y) * kCube + e] = R.field[idx];
}

// One block per (cube, kz plane): T1[ky][ix] = sum_kx M[kz][ky][kx] cx[ix][kx]
// on 16-row tiles, then t2[kz][iy][ix] = sum_ky cy[iy][ky] T1[ky][ix]. Operand
// layouts of v_mfma_f32_16x16x4_f32 (lane l, g = l / 16, q = l % 16): A[q][g],
// B[g][q], D[4 g + v][q]. Each lane loads four consecutive kx (or ky) and
// feeds them to four products, so product v covers index 4 g + v of a
// 16-wide chunk.
__global__ __launch_bounds__(256) void k_cube_contract(const CubeRec* __restrict__ r, int n, int par) {
  __shared__ float s[4][256];
  int ri = 0;
  for (int q = 1; q < n; ++q)
    if ((long long)blockIdx.x >= r[q].block0) ri = q;
  const CubeRec& R = r[ri];
  const int kz = int((long long)blockIdx.x - R.block0);
  const int H = R.H, Wp = R.Wp, Hp = (H + 15) & ~15;
  const float* Mp = R.M[par ^ 1] + (long long)kz * H * Wp;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, q = lane & 15;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int ky0 = wave * 16; ky0 < H; ky0 += 64) {
    floatx4 t1 = {0.f, 0.f, 0.f, 0.f};
    const int row = ky0 + q;
    const float* mrow = Mp + (long long)row * Wp + 4 * g;
    const float* crow = R.cx + (long long)q * Wp + 4 * g;
    for (int kx0 = 0; kx0 < Wp; kx0 += 16) {
      floatx4 a = {0.f, 0.f, 0.f, 0.f};
      if (row < H) a = *reinterpret_cast<const floatx4*>(mrow + kx0);
      const floatx4 b = *reinterpret_cast<const floatx4*>(crow + kx0);
      t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], t1, 0, 0, 0);
      t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], t1, 0, 0, 0);
      t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], t1, 0, 0, 0);
      t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], t1, 0, 0, 0);
    }
    // lane holds T1[ky0 + 4 g + v][q]: the B operand of the y contraction
    const floatx4 cy = *reinterpret_cast<const floatx4*>(R.cy + (long long)q * Hp + ky0 + 4 * g);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(cy[0], t1[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(cy[1], t1[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(cy[2], t1[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(cy[3], t1[3], acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) s[wave][(4 * g + v) * 16 + q] = acc[v];
  __syncthreads();
  const int e = threadIdx.x;
  R.t2[(long long)kz * 256 + e] = s[0][e] + s[1][e] + s[2][e] + s[3][e];
}

// z contraction per cube cell, into the cube trace row of this step.
__global__ __launch_bounds__(256) void k_cube_eval(const CubeRec* __restrict__ r, int nrec, float* __restrict__ out,
                                                   int cap, const int* __restrict__ d_step) {
  const int t = *d_step;
  const int e = blockIdx.y * 256 + threadIdx.x;
  if (t >= cap || e >= kCube) return;
  const CubeRec& R = r[blockIdx.x];
  const int s = R.sel[e];
  const float* t2 = R.t2 + ((s >> 4) & 15) * 16 + (s & 15);
  const float* cz = R.cz + (long long)(s >> 8) * R.D;
  double acc = 0.0;
  for (int kz = 0; kz < R.D; ++kz) acc += double(cz[kz]) * double(t2[(long long)kz * 256]);
  out[((long long)t * nrec + blockIdx.x) * kCube + e] = float(acc);
}

float redft01_weight(int n, int k, int x) {
  const double pi = 3.14159265358979323846;
  return float((k == 0 ? 1.0 : 2.0) * std::cos(pi * k * (x + 0.5) / n));
}

// weights [16][stride] at the given coordinates (rows past the set and columns past n zero)
std::vector<float> weight_rows(int n, int stride, const std::vector<int>& xs) {
  std::vector<float> w(size_t(kSet) * stride, 0.f);
  for (size_t i = 0; i < xs.size(); ++i)
    for (int k = 0; k < n; ++k) w[i * stride + k] = redft01_weight(n, k, xs[i]);
  return w;
}
}  // namespace

struct OutputState {
  CubeRec* d_cube = nullptr;
  int ncube = 0;
  long long contract_blocks = 0;
};

void destroy_output(Ctx& c) {
  delete c.out;
  c.out = nullptr;
}

// Every cube cell inside its owner's array (out of it the reference reads past
// the array). Host-side: runs before the device set-up.
int check_cubes(const Ctx& c) {
  if (!c.cubes) return ARD_OK;
  for (size_t i = 0; i < c.topo.recorders.size(); ++i) {
    const RecorderDef& rd = c.topo.recorders[i];
    const PartGeom& p = c.topo.parts[size_t(rd.part)];
    const long long f0 = ((long long)rd.lz * p.h + rd.ly) * p.w + rd.lx, plane = (long long)p.h * p.w;
    if (f0 - (5 * plane + 5LL * p.w + 5) < 0 || f0 + 4 * plane + 4LL * p.w + 4 >= p.cells())
      return set_error(ARD_ERR_RANGE, "recorder " + std::to_string(i) +
                                          "'s cube addresses cells outside its partition's array");
  }
  return ARD_OK;
}

int setup_cubes(Ctx& c) {
  const int nr = int(c.topo.recorders.size());
  if (!c.cubes || nr == 0) return ARD_OK;
  if (!c.out) c.out = new OutputState();
  OutputState& O = *c.out;
  const bool batched = c.prm.form == ARD_FORM_BATCHED;
  std::vector<CubeRec> h(static_cast<size_t>(nr));
  long long blocks = 0;
  int r;
  for (int i = 0; i < nr; ++i) {
    const RecorderDef& rd = c.topo.recorders[size_t(i)];
    const PartGeom& p = c.topo.parts[size_t(rd.part)];
    CubeRec& q = h[size_t(i)];
    q = CubeRec{};
    q.W = p.w; q.H = p.h; q.D = p.d;
    q.f0 = ((long long)rd.lz * p.h + rd.ly) * p.w + rd.lx;
    const long long plane = (long long)p.h * p.w;
    if (!batched) {
      q.field = c.dct[size_t(rd.part)].p;
      continue;
    }
    if ((r = batched_modes(c, rd.part, q.M, &q.Wp))) return r;
    // distinct local coordinates of the cube's cells along each axis
    std::vector<int> X, Y, Z, sel(kCube);
    std::vector<long long> flat(kCube);
    for (int e = 0; e < kCube; ++e) {
      const int a = e / 100 - 5, b = (e / 10) % 10 - 5, k = e % 10 - 5;
      flat[size_t(e)] = q.f0 + (a * (long long)p.h + b) * p.w + k;
      X.push_back(int(flat[size_t(e)] % p.w));
      Y.push_back(int(flat[size_t(e)] / p.w % p.h));
      Z.push_back(int(flat[size_t(e)] / plane));
    }
    for (auto* v : {&X, &Y, &Z}) {
      std::sort(v->begin(), v->end());
      v->erase(std::unique(v->begin(), v->end()), v->end());
      if (int(v->size()) > kSet)
        return set_error(ARD_ERR_UNSUPPORTED, "recorder " + std::to_string(i) +
                                                  "'s cube spans more than 16 rows of its partition (batched form)");
    }
    for (int e = 0; e < kCube; ++e) {
      const long long f = flat[size_t(e)];
      const int ix = int(std::lower_bound(X.begin(), X.end(), int(f % p.w)) - X.begin());
      const int iy = int(std::lower_bound(Y.begin(), Y.end(), int(f / p.w % p.h)) - Y.begin());
      const int iz = int(std::lower_bound(Z.begin(), Z.end(), int(f / plane)) - Z.begin());
      sel[size_t(e)] = ix | iy << 4 | iz << 8;
    }
    const int Hp = (p.h + 15) & ~15;
    const std::vector<float> cx = weight_rows(p.w, q.Wp, X), cy = weight_rows(p.h, Hp, Y),
                             cz = weight_rows(p.d, p.d, Z);
    float *dx, *dy, *dz, *t2;
    int* ds;
    if ((r = c.dalloc((void**)&dx, cx.size() * 4)) || (r = c.upload(dx, cx.data(), cx.size() * 4)) ||
        (r = c.dalloc((void**)&dy, cy.size() * 4)) || (r = c.upload(dy, cy.data(), cy.size() * 4)) ||
        (r = c.dalloc((void**)&dz, cz.size() * 4)) || (r = c.upload(dz, cz.data(), cz.size() * 4)) ||
        (r = c.dalloc((void**)&ds, sel.size() * 4)) || (r = c.upload(ds, sel.data(), sel.size() * 4)) ||
        (r = c.dalloc((void**)&t2, size_t(p.d) * 256 * 4)))
      return r;
    q.cx = dx; q.cy = dy; q.cz = dz; q.sel = ds; q.t2 = t2;
    q.block0 = blocks;
    blocks += p.d;
  }
  if ((r = c.dalloc((void**)&O.d_cube, h.size() * sizeof(CubeRec))) ||
      (r = c.upload(O.d_cube, h.data(), h.size() * sizeof(CubeRec))))
    return r;
  O.ncube = nr;
  O.contract_blocks = blocks;
  return ARD_OK;
}

// Queues this step's cube samples; before the recorder kernel, which advances
// the device step counter. Batched: reads the modes the step's modal pass wrote.
int enqueue_cubes(Ctx& c, int par, hipStream_t st) {
  if (!c.cubes || !c.out || c.out->ncube == 0) return ARD_OK;
  const OutputState& O = *c.out;
  if (c.prm.form == ARD_FORM_BATCHED) {
    hipLaunchKernelGGL(k_cube_contract, dim3(unsigned(O.contract_blocks)), dim3(256), 0, st, O.d_cube, O.ncube, par);
    hipLaunchKernelGGL(k_cube_eval, dim3(O.ncube, (kCube + 255) / 256), dim3(256), 0, st, O.d_cube, O.ncube,
                       c.d_cube, c.trace_cap, c.d_step);
  } else {
    hipLaunchKernelGGL(k_cube_gather, dim3((kCube + 255) / 256, O.ncube), dim3(256), 0, st, O.d_cube, O.ncube,
                       c.d_cube, c.trace_cap, c.d_step);
  }
  return hipGetLastError() == hipSuccess ? ARD_OK : set_error(ARD_ERR_HIP, "recorder cube launch");
}

}  // namespace ard
