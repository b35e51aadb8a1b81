// The C-ABI context of libard.so: scene registration, device state, stepping.
//
// Replaces the reference's per-step driver Simulation::Update
// (ARD/simulation.cpp:300-317) and the device glue of utils_VkFFT
// (ARD/utils_VkFFT.cpp): all state is device-resident, nothing crosses the
// host boundary inside a step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ard.h"
#include "ard_ctx_impl.h"
#include "ard_kernels.h"

using namespace ard;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace

namespace ard {

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return ARD_ERR_HIP;
}

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int PhaseTimer::begin(hipStream_t s, int phase) {
  if (!on) return -1;
  if (used + 2 > pool.size()) {
    for (int k = 0; k < 64; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      pool.push_back(e);
    }
  }
  const int pair = int(used / 2);
  hipEventRecord(pool[used], s);
  used += 2;
  phase_of.resize(pair + 1);
  phase_of[pair] = phase;
  return pair;
}

void PhaseTimer::end(hipStream_t s, int pair) {
  if (pair >= 0) hipEventRecord(pool[2 * size_t(pair) + 1], s);
}

int PhaseTimer::collect() {
  for (size_t i = 0; i + 1 < used; i += 2) {
    hipError_t e = hipEventSynchronize(pool[i + 1]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    float ms = 0.f;
    hipEventElapsedTime(&ms, pool[i], pool[i + 1]);
    const int ph = phase_of[i / 2];
    total_ms[ph] += ms;
    count[ph] += 1;
  }
  used = 0;
  return ARD_OK;
}

void PhaseTimer::release() {
  for (hipEvent_t e : pool) hipEventDestroy(e);
  pool.clear();
  used = 0;
}

Ctx::~Ctx() {
  if (device >= 0) hipSetDevice(device);
  timer.release();
  if (graph_exec[0]) hipGraphExecDestroy(graph_exec[0]);
  if (graph_exec[1]) hipGraphExecDestroy(graph_exec[1]);
  destroy_batched(*this);
  destroy_adjoint(*this);
  delete twin;
  if (device >= 0) hipSetDevice(device);
  destroy_output(*this);
  destroy_view(*this);
  for (void* p : allocs) hipFree(p);
  if (stream) hipStreamDestroy(stream);
}

int Ctx::dalloc(void** out, size_t bytes) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 4));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  allocs.push_back(p);
  *out = p;
  return ARD_OK;
}

int Ctx::dzero(void* p, size_t bytes) {
  hipError_t e = hipMemsetAsync(p, 0, bytes, stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipMemsetAsync");
}

int Ctx::upload(void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync H2D");
  e = hipStreamSynchronize(stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

// Cosine matrices of one axis length, FFTW's REDFT10 (forward) and REDFT01
// (inverse) unnormalised kernels (ARD/dct_volume.cpp:17,19), built in FP64.
int Ctx::cos_tables(int n, const float** fwd, const float** inv) {
  auto it = tables.find(n);
  if (it == tables.end()) {
    std::vector<float> hf(size_t(n) * n), hi(size_t(n) * n);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        hf[size_t(k) * n + j] = float(2.0 * std::cos(pi * k * (j + 0.5) / n));
        hi[size_t(k) * n + j] = j == 0 ? 1.0f : float(2.0 * std::cos(pi * j * (k + 0.5) / n));
      }
    float *df, *di;
    int r;
    if ((r = dalloc((void**)&df, hf.size() * 4)) || (r = dalloc((void**)&di, hi.size() * 4))) return r;
    if ((r = upload(df, hf.data(), hf.size() * 4)) || (r = upload(di, hi.data(), hi.size() * 4))) return r;
    it = tables.emplace(n, std::make_pair(df, di)).first;
  }
  *fwd = it->second.first;
  *inv = it->second.second;
  return ARD_OK;
}

// Modal coefficients of one DCT partition (ARD/dct_partition.cpp:20-36,66-68),
// computed in FP64 and stored in FP32:
//   a = .999 * 2 cos(w dt),  g = .999 * 2 (1 - cos(w dt)) / w^2 * s_f * s_i
// with 1-based wavenumbers w = c0 pi sqrt((k/Lx)^2 + (j/Ly)^2 + (i/Lz)^2) and the
// transform scales s_f = 1/(2 sqrt(2N)), s_i = 1/sqrt(2N) folded in (the device
// stores modes multiplied by s_i, so its inverse transform needs no scale).
void modal_tables(int W, int H, int D, double h, double dt, double c0, std::vector<float>& a,
                  std::vector<float>& g) {
  const double pi = 3.14159265358979323846;
  const double damp = double(0.999f);
  const double N = double(W) * H * D;
  const double sfsi = 1.0 / (2.0 * std::sqrt(2.0 * N)) / std::sqrt(2.0 * N);
  const double lx = W * h, ly = H * h, lz = D * h;
  a.resize(size_t(N));
  g.resize(size_t(N));
  for (int i = 1; i <= D; ++i)
    for (int j = 1; j <= H; ++j)
      for (int k = 1; k <= W; ++k) {
        const size_t idx = (size_t(i - 1) * H + (j - 1)) * W + (k - 1);
        const double w = c0 * pi * std::sqrt((i / lz) * (i / lz) + (j / ly) * (j / ly) + (k / lx) * (k / lx));
        const double th = w * dt;
        const double s = std::sin(0.5 * th);
        a[idx] = float(damp * 2.0 * std::cos(th));
        g[idx] = float(damp * 4.0 * s * s / (w * w) * sfsi);  // 1 - cos = 2 sin^2(th/2)
      }
}

}  // namespace ard

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* ard_last_error(void) { return g_err.c_str(); }

int ard_ctx_create(int device, const ard_params* params, ard_ctx** out) {
  if (!params || !out) return fail(ARD_ERR_INVALID, "null argument");
  if (!(params->h > 0) || !(params->dt > 0) || !(params->c0 > 0))
    return fail(ARD_ERR_INVALID, "h, dt and c0 must be positive");
  if (params->pml_layers < 3) return fail(ARD_ERR_INVALID, "pml_layers must be at least 3");
  if (params->form != ARD_FORM_PLAIN && params->form != ARD_FORM_BATCHED)
    return fail(ARD_ERR_INVALID, "unknown form");
  if (params->recorder_mode != ARD_REC_REFERENCE && params->recorder_mode != ARD_REC_LOCAL)
    return fail(ARD_ERR_INVALID, "unknown recorder mode");
  // device -1: a host-only context (topology queries only; ard_finalize fails)
  if (device != -1) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= ndev) return fail(ARD_ERR_INVALID, "no such device");
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  }
  auto* c = new Ctx();
  c->device = device;
  c->prm = *params;
  c->topo.pml_layers = params->pml_layers;
  c->topo.absorption = params->absorption;
  if (device != -1) {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return hip_fail(e, "hipStreamCreate");
    }
  }
  *out = reinterpret_cast<ard_ctx*>(c);
  return ARD_OK;
}

int ard_ctx_destroy(ard_ctx* ctx) {
  if (!ctx) return fail(ARD_ERR_INVALID, "null context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (c->device >= 0) {
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
  }
  delete c;
  return ARD_OK;
}

int ard_set_pml_scale(ard_ctx* ctx, const float scale[6]) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !scale) return fail(ARD_ERR_INVALID, "null argument");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  for (int i = 0; i < 6; ++i) c->pml_scale[i] = scale[i];
  return ARD_OK;
}

int ard_add_dct_partition(ard_ctx* ctx, int x0, int y0, int z0, int w, int h, int d) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized || c->topology_built) return fail(ARD_ERR_STATE, "topology is fixed");
  if (w < 6 || h < 6 || d < 6) return fail(ARD_ERR_INVALID, "partition sides must be at least 6 cells");
  return c->topo.add_dct(x0, y0, z0, w, h, d);
}

int ard_add_pml_partition(ard_ctx* ctx, int neighbor, int type, int x0, int y0, int z0, int w, int h, int d) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  if (type < 0 || type > 5 || w <= 0 || h <= 0 || d <= 0) return fail(ARD_ERR_INVALID, "bad PML slab");
  if (neighbor < 0 || neighbor >= int(c->topo.parts.size()) || c->topo.parts[neighbor].is_pml)
    return fail(ARD_ERR_INVALID, "PML neighbour must be a DCT partition");
  c->topology_built = true;
  return c->topo.add_pml(neighbor, type, x0, y0, z0, w, h, d);
}

int ard_add_boundary(ard_ctx* ctx, int type, int a, int b, int xs, int xe, int ys, int ye, int zs, int ze,
                     float absorption) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  const int np = int(c->topo.parts.size());
  if (type < 0 || type > 2 || a < 0 || a >= np || b < 0 || b >= np || a == b)
    return fail(ARD_ERR_INVALID, "bad boundary");
  BoundaryGeom g;
  g.type = type; g.a = a; g.b = b;
  g.xs = xs; g.xe = xe; g.ys = ys; g.ye = ye; g.zs = zs; g.ze = ze;
  g.absorption = absorption;
  c->topology_built = true;
  return c->topo.add_boundary(g);
}

int ard_build_topology(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized || c->topology_built) return fail(ARD_ERR_STATE, "topology already built");
  c->topo.build();
  c->topology_built = true;
  return ARD_OK;
}

int ard_add_source(ard_ctx* ctx, int x, int y, int z, const float* wave, int n) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  if (wave && n <= 0) return fail(ARD_ERR_INVALID, "waveform length must be positive");
  SourceDef s;
  s.x = x; s.y = y; s.z = z;
  if (wave) s.wave.assign(wave, wave + n);
  c->topo.sources.push_back(s);
  return int(c->topo.sources.size()) - 1;
}

int ard_add_recorder(ard_ctx* ctx, int x, int y, int z) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  RecorderDef r;
  r.x = x; r.y = y; r.z = z;
  c->topo.recorders.push_back(r);
  return int(c->topo.recorders.size()) - 1;
}

int ard_finalize(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  std::string err = c->topo.resolve(c->prm.recorder_mode == ARD_REC_LOCAL);
  if (!err.empty()) return fail(ARD_ERR_RANGE, err);
  if (int r = check_cubes(*c)) return r;
  if (c->device < 0) return fail(ARD_ERR_STATE, "host-only context (device -1) cannot be finalized");
  hipSetDevice(c->device);
  int r = c->prm.form == ARD_FORM_PLAIN ? finalize_plain(*c) : finalize_batched(*c);
  if (r == ARD_OK) r = setup_cubes(*c);
  if (r != ARD_OK) return r;
  c->finalized = true;
  return ARD_OK;
}

int ard_num_partitions(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  return c ? int(c->topo.parts.size()) : fail(ARD_ERR_INVALID, "null context");
}
int ard_num_boundaries(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  return c ? int(c->topo.bnds.size()) : fail(ARD_ERR_INVALID, "null context");
}
int ard_num_recorders(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  return c ? int(c->topo.recorders.size()) : fail(ARD_ERR_INVALID, "null context");
}

int ard_partition_info(ard_ctx* ctx, int part, int* info9) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !info9) return fail(ARD_ERR_INVALID, "null argument");
  if (part < 0 || part >= int(c->topo.parts.size())) return fail(ARD_ERR_RANGE, "no such partition");
  const PartGeom& p = c->topo.parts[part];
  const int v[9] = {p.is_pml ? 1 : 0, p.pml_type, p.parent, p.x0, p.y0, p.z0, p.w, p.h, p.d};
  std::memcpy(info9, v, sizeof v);
  return ARD_OK;
}

int ard_boundary_info(ard_ctx* ctx, int bnd, int* info9) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !info9) return fail(ARD_ERR_INVALID, "null argument");
  if (bnd < 0 || bnd >= int(c->topo.bnds.size())) return fail(ARD_ERR_RANGE, "no such boundary");
  const BoundaryGeom& b = c->topo.bnds[bnd];
  const int v[9] = {b.type, b.a, b.b, b.xs, b.xe, b.ys, b.ye, b.zs, b.ze};
  std::memcpy(info9, v, sizeof v);
  return ARD_OK;
}

int ard_check_finite(ard_ctx* ctx, int* bad_partition) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !bad_partition) return fail(ARD_ERR_INVALID, "null argument");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  hipSetDevice(c->device);
  return check_finite(*c, bad_partition);
}

int ard_find_boundary(const int* box_a6, const int* box_b6, int* bnd_info9) {
  if (!box_a6 || !box_b6 || !bnd_info9) return fail(ARD_ERR_INVALID, "null argument");
  for (int k = 3; k < 6; ++k)
    if (box_a6[k] <= 0 || box_b6[k] <= 0) return fail(ARD_ERR_INVALID, "box sides must be positive");
  Topology t;
  t.add_dct(box_a6[0], box_a6[1], box_a6[2], box_a6[3], box_a6[4], box_a6[5]);
  t.add_dct(box_b6[0], box_b6[1], box_b6[2], box_b6[3], box_b6[4], box_b6[5]);
  BoundaryGeom b;
  if (!t.find_boundary(0, 1, 1.0f, &b)) return 0;
  const int v[9] = {b.type, -1, -1, b.xs, b.xe, b.ys, b.ye, b.zs, b.ze};
  std::memcpy(bnd_info9, v, sizeof v);
  return 1;
}

int ard_recorder_partition(ard_ctx* ctx, int rec) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (rec < 0 || rec >= int(c->topo.recorders.size())) return fail(ARD_ERR_RANGE, "no such recorder");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  return c->topo.recorders[rec].part;
}

int ard_step(ard_ctx* ctx, int n) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  if (n < 0) return fail(ARD_ERR_INVALID, "negative step count");
  if (n == 0) return ARD_OK;
  hipSetDevice(c->device);
  int r = ensure_trace(*c, c->time_step + n);
  if (r != ARD_OK) return r;
  r = c->prm.form == ARD_FORM_PLAIN ? step_plain(*c, n) : step_batched(*c, n);
  if (r != ARD_OK) return r;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return ARD_OK;
}

int ard_set_kernel_timing(ard_ctx* ctx, int on) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  hipSetDevice(c->device);
  int r = c->timer.collect();
  if (r) return r;
  c->timer.on = on != 0;
  for (int i = 0; i < ARD_NUM_PHASES; ++i) {
    c->timer.total_ms[i] = 0;
    c->timer.count[i] = 0;
  }
  return ARD_OK;
}

int ard_kernel_times(ard_ctx* ctx, float* avg_ms, int* launches, int nphase) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !avg_ms || !launches) return fail(ARD_ERR_INVALID, "null argument");
  if (nphase < 0 || nphase > ARD_NUM_PHASES) return fail(ARD_ERR_INVALID, "bad phase count");
  if (c->device >= 0) hipSetDevice(c->device);
  int r = c->timer.collect();
  if (r) return r;
  for (int i = 0; i < nphase; ++i) {
    launches[i] = int(c->timer.count[i]);
    avg_ms[i] = c->timer.count[i] ? float(c->timer.total_ms[i] / c->timer.count[i]) : 0.f;
  }
  return ARD_OK;
}

int ard_time_step(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  return c ? c->time_step : fail(ARD_ERR_INVALID, "null context");
}

int ard_synchronize(ard_ctx* ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  hipSetDevice(c->device);
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int ard_fetch_recorders(ard_ctx* ctx, float* out, int t0, int nt) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || (!out && nt > 0)) return fail(ARD_ERR_INVALID, "null argument");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  const int nr = int(c->topo.recorders.size());
  if (t0 < 0 || nt < 0 || t0 + nt > c->time_step) return fail(ARD_ERR_RANGE, "steps not taken yet");
  if (nt == 0 || nr == 0) return ARD_OK;
  hipSetDevice(c->device);
  hipError_t e = hipMemcpyAsync(out, c->d_trace + size_t(t0) * nr, size_t(nt) * nr * 4, hipMemcpyDeviceToHost,
                                c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int ard_set_recorder_cubes(ard_ctx* ctx, int on) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (c->finalized) return fail(ARD_ERR_STATE, "already finalized");
  c->cubes = on != 0;
  return ARD_OK;
}

int ard_fetch_recorder_cubes(ard_ctx* ctx, float* out, int t0, int nt) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || (!out && nt > 0)) return fail(ARD_ERR_INVALID, "null argument");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  if (!c->cubes) return fail(ARD_ERR_STATE, "recorder cubes were not switched on before ard_finalize");
  const size_t row = c->topo.recorders.size() * 1000;
  if (t0 < 0 || nt < 0 || t0 + nt > c->time_step) return fail(ARD_ERR_RANGE, "steps not taken yet");
  if (nt == 0 || row == 0) return ARD_OK;
  hipSetDevice(c->device);
  hipError_t e = hipMemcpyAsync(out, c->d_cube + size_t(t0) * row, size_t(nt) * row * 4, hipMemcpyDeviceToHost,
                                c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int ard_render_view(ard_ctx* ctx, int view, int coord, unsigned int* pixels, int* dims) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !dims) return fail(ARD_ERR_INVALID, "null argument");
  if (view != ARD_VIEW_XY && view != ARD_VIEW_YZ) return fail(ARD_ERR_INVALID, "view must be ARD_VIEW_XY or ARD_VIEW_YZ");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  hipSetDevice(c->device);
  return render_view(*c, view, coord, pixels, dims);
}

int ard_copy_pressure_field(ard_ctx* ctx, int part, float* out) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !out) return fail(ARD_ERR_INVALID, "null argument");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  if (part < 0 || part >= int(c->topo.parts.size())) return fail(ARD_ERR_RANGE, "no such partition");
  hipSetDevice(c->device);
  const float* src = nullptr;
  int r = pressure_field_ptr(*c, part, &src);
  if (r != ARD_OK) return r;
  const size_t n = size_t(c->topo.parts[part].cells());
  hipError_t e = hipMemcpyAsync(out, src, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int ard_get_pressure(ard_ctx* ctx, int part, int x, int y, int z, float* out) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !out) return fail(ARD_ERR_INVALID, "null argument");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  if (part < 0 || part >= int(c->topo.parts.size())) return fail(ARD_ERR_RANGE, "no such partition");
  const PartGeom& p = c->topo.parts[part];
  if (x < 0 || x >= p.w || y < 0 || y >= p.h || z < 0 || z >= p.d) return fail(ARD_ERR_RANGE, "cell outside partition");
  hipSetDevice(c->device);
  const float* src = nullptr;
  int r = pressure_field_ptr(*c, part, &src);
  if (r != ARD_OK) return r;
  const size_t idx = (size_t(z) * p.h + y) * p.w + x;
  hipError_t e = hipMemcpyAsync(out, src + idx, 4, hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2H");
  e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int ard_set_force(ard_ctx* ctx, int part, int x, int y, int z, float f) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  if (!c->finalized) return fail(ARD_ERR_STATE, "not finalized");
  if (part < 0 || part >= int(c->topo.parts.size())) return fail(ARD_ERR_RANGE, "no such partition");
  const PartGeom& p = c->topo.parts[part];
  if (x < 0 || x >= p.w || y < 0 || y >= p.h || z < 0 || z >= p.d) return fail(ARD_ERR_RANGE, "cell outside partition");
  hipSetDevice(c->device);
  return set_force_cell(*c, part, x, y, z, f);
}

int ard_batched_config(ard_ctx* ctx, int* out4) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !out4) return fail(ARD_ERR_INVALID, "null argument");
  if (c->finalized && c->prm.form == ARD_FORM_PLAIN) {  // the plain form: its transform kernel only
    out4[0] = out4[1] = out4[3] = 0;
    out4[2] = c->plain_variant;
    return ARD_OK;
  }
  return batched_config(*c, out4);
}

int ard_set_batched_tiling(ard_ctx* ctx, int rows_per_wave, int kz_per_block) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  return batched_set_tiling(*c, rows_per_wave, kz_per_block);
}

int ard_set_batched_option(ard_ctx* ctx, int option, int value) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(ARD_ERR_INVALID, "null context");
  return batched_set_option(*c, option, value);
}

int ard_batched_option(ard_ctx* ctx, int option, int* value) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !value) return fail(ARD_ERR_INVALID, "null argument");
  return batched_option(*c, option, value);
}

int ard_set_transform_variant(int variant) {
  if (variant < 0 || variant > 5) return fail(ARD_ERR_INVALID, "unknown transform variant");
  g_gemm_variant = variant;
  return ARD_OK;
}

int ard_selftest_transform(int device, int variant, int b_layout, int M, int N, int K, const float* A, const float* B,
                           float* C) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0) return fail(ARD_ERR_INVALID, "bad arguments");
  if (variant < 0 || variant > 2 || b_layout < 0 || b_layout > 1) return fail(ARD_ERR_INVALID, "bad variant");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  GemmDesc* dD = nullptr;
  const size_t sa = size_t(M) * K * 4, sb = size_t(K) * N * 4, sc = size_t(M) * N * 4;
  int r = ARD_OK;
  if ((e = hipMalloc(&dA, sa)) != hipSuccess || (e = hipMalloc(&dB, sb)) != hipSuccess ||
      (e = hipMalloc(&dC, sc)) != hipSuccess || (e = hipMalloc(&dD, sizeof(GemmDesc))) != hipSuccess) {
    r = hip_fail(e, "hipMalloc");
  } else {
    GemmDesc g{};
    g.A = dA; g.B = dB; g.C = dC;
    g.M = M; g.N = N; g.K = K; g.N1 = N;
    if (b_layout == 0) { g.bK = N; g.b1 = 1; } else { g.bK = 1; g.b1 = K; }
    g.cM = N; g.c1 = 1;
    g.alpha = 1.f;
    g.tiles_n = (N + 63) / 64;
    g.tile0 = 0;
    const int tiles = ((M + 63) / 64) * g.tiles_n;
    const int saved = g_gemm_variant;
    g_gemm_variant = variant;
    if (hipMemcpy(dA, A, sa, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dB, B, sb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dD, &g, sizeof g, hipMemcpyHostToDevice) != hipSuccess)
      r = fail(ARD_ERR_HIP, "hipMemcpy");
    else if (launch_gemm(dD, 1, tiles, b_layout == 1, false, 0, nullptr))
      r = fail(ARD_ERR_HIP, "launch");
    else if ((e = hipDeviceSynchronize()) != hipSuccess)
      r = hip_fail(e, "hipDeviceSynchronize");
    else if ((e = hipMemcpy(C, dC, sc, hipMemcpyDeviceToHost)) != hipSuccess)
      r = hip_fail(e, "hipMemcpy D2H");
    g_gemm_variant = saved;
  }
  hipFree(dA); hipFree(dB); hipFree(dC); hipFree(dD);
  return r;
}

int ard_selftest_fft(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y) {
  if (!X || !Y || L < 2 || N <= 0 || b_layout < 0 || b_layout > 1) return fail(ARD_ERR_INVALID, "bad arguments");
  if ((long long)L * N >= (1LL << 30)) return fail(ARD_ERR_INVALID, "too large");
  return fft_selftest(device, inverse != 0, b_layout, L, N, X, Y);
}

int ard_selftest_fold(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y) {
  // L >= 6: partition sides are at least 6 cells, and the kernel relies on it
  if (!X || !Y || L < 6 || N <= 0 || b_layout < 0 || b_layout > 1) return fail(ARD_ERR_INVALID, "bad arguments");
  if ((long long)L * N + 16 >= (1LL << 30)) return fail(ARD_ERR_INVALID, "too large");
  return fold_selftest(device, inverse != 0, b_layout, L, N, X, Y, 0);
}

int ard_selftest_fold_bf16(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y) {
  if (!X || !Y || L < 6 || N <= 0 || b_layout < 0 || b_layout > 1) return fail(ARD_ERR_INVALID, "bad arguments");
  if ((long long)L * N + 16 >= (1LL << 30)) return fail(ARD_ERR_INVALID, "too large");
  return fold_selftest(device, inverse != 0, b_layout, L, N, X, Y, 1);
}

int ard_selftest_fold_direct(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y) {
  if (!X || !Y || L < 6 || N <= 0 || b_layout < 0 || b_layout > 1) return fail(ARD_ERR_INVALID, "bad arguments");
  if ((long long)L * N + 16 >= (1LL << 30)) return fail(ARD_ERR_INVALID, "too large");
  return fold_selftest(device, inverse != 0, b_layout, L, N, X, Y, 2);
}

}  // extern "C"
