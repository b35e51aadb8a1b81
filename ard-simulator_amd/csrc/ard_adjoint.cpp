// Discrete adjoint of the plain time step: gradients of a recorder loss with
// respect to the source samples and the six per-face PML damping scales
// (ard_adjoint, include/ard.h). No counterpart in the reference, which has no
// gradients; the forward operators it transposes are ARD/simulation.cpp:300-317
// (step order), ARD/dct_partition.cpp:48-73 (modal update and transforms),
// ARD/pml_partition.cpp:113-200 (PML), ARD/boundary.cpp:21-173 (interface),
// ARD/partition.cpp:187-197 (sources) and ARD/recorder.cpp:40-57 (recorders).
//
// Forward state carried between steps: DCT modes M, M- and force F; PML p,
// p-, phi and f. The sweep keeps a checkpoint of that state every K steps
// (K = sqrt(T) by default), replays each K-step segment forward from its
// checkpoint (saving the PML state each step, the only forward values the
// damping-scale derivative needs), then walks the segment backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ard_adjoint.h"
#include "ard_ctx_impl.h"
#include "ard_kernels.h"

namespace ard {

namespace {
long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

struct Region {
  float* p;
  size_t n;  // floats
};

struct AdjDct {
  int part = -1;
  long long n = 0;
  float *lM = nullptr, *lMp = nullptr, *lF = nullptr, *lp = nullptr;
  GemmDesc pass[6];  // IDCT^T x, y, z; DCT^T x, y, z
  int tiles[6] = {};
  GemmDesc* d_pass = nullptr;
};

struct AdjPml {
  int part = -1;
  long long n = 0;
  // two sets of adjoint state: [set] = {p, q, phi x, phi y, phi z, f}
  float* set[2][6] = {};
  const float* prof = nullptr;
};

}  // namespace

struct AdjState {
  std::vector<AdjDct> dct;
  std::vector<AdjPml> pml;
  std::vector<Region> regions;  // forward state saved in a checkpoint
  size_t ck_floats = 0;
  float* tmp[4] = {};
  BndDesc* d_bnd = nullptr;
  long long bnd_total = 0;
  int nbnd = 0;
  PartAccess* d_parts[2] = {};
  RecDesc* d_rec = nullptr;
  AdjSource* d_src = nullptr;
  int nsrc = 0;
  double* d_dtheta = nullptr;
  std::map<int, std::pair<float*, float*>> ttables;  // axis length -> transposed (REDFT10, REDFT01)
};

void destroy_adjoint(Ctx& c) {
  delete c.adj;
  c.adj = nullptr;
}

namespace {

int transposed_tables(Ctx& c, int n, const float** fT, const float** iT) {
  AdjState& S = *c.adj;
  auto it = S.ttables.find(n);
  if (it == S.ttables.end()) {
    // transposes of the matrices of Ctx::cos_tables
    std::vector<float> hf(size_t(n) * n), hi(size_t(n) * n);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        hf[size_t(j) * n + k] = float(2.0 * std::cos(pi * k * (j + 0.5) / n));
        hi[size_t(j) * n + k] = j == 0 ? 1.0f : float(2.0 * std::cos(pi * j * (k + 0.5) / n));
      }
    float *df, *di;
    int r;
    if ((r = c.dalloc((void**)&df, hf.size() * 4)) || (r = c.dalloc((void**)&di, hi.size() * 4))) return r;
    if ((r = c.upload(df, hf.data(), hf.size() * 4)) || (r = c.upload(di, hi.data(), hi.size() * 4))) return r;
    it = S.ttables.emplace(n, std::make_pair(df, di)).first;
  }
  *fT = it->second.first;
  *iT = it->second.second;
  return ARD_OK;
}

// the plain form's per-axis passes (ard_plain.cpp), with any matrix
GemmDesc axis_pass(int axis, const PartGeom& p, const float* A, const float* in, float* out) {
  const long long W = p.w, H = p.h, D = p.d;
  GemmDesc g{};
  g.A = A; g.B = in; g.C = out; g.alpha = 1.f;
  if (axis == 0) {
    g.M = p.w; g.K = p.w; g.N = int(H * D); g.N1 = g.N;
    g.bK = 1; g.b1 = W; g.cM = 1; g.c1 = W;
  } else if (axis == 1) {
    g.M = p.h; g.K = p.h; g.N = int(W * D); g.N1 = p.w;
    g.bK = W; g.b1 = 1; g.b2 = W * H; g.cM = W; g.c1 = 1; g.c2 = W * H;
  } else {
    g.M = p.d; g.K = p.d; g.N = int(W * H); g.N1 = g.N;
    g.bK = W * H; g.b1 = 1; g.cM = W * H; g.c1 = 1;
  }
  g.tiles_n = int(cdiv(g.N, 64));
  g.tile0 = 0;
  return g;
}

int setup_adjoint(Ctx& c) {
  c.adj = new AdjState();
  AdjState& S = *c.adj;
  int r;
  const int np = int(c.topo.parts.size());
  long long maxn = 0;
  for (int i = 0; i < np; ++i)
    if (!c.topo.parts[i].is_pml) maxn = std::max<long long>(maxn, c.topo.parts[i].cells());
  for (int k = 0; k < 4; ++k)
    if ((r = c.dalloc((void**)&S.tmp[k], size_t(std::max<long long>(maxn, 1)) * 4))) return r;
  std::vector<PartAccess> acc[2];
  acc[0].assign(np, PartAccess());
  acc[1].assign(np, PartAccess());
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = c.topo.parts[i];
    const long long n = p.cells();
    for (int s = 0; s < 2; ++s) {
      acc[s][i].W = p.w; acc[s][i].H = p.h; acc[s][i].D = p.d;
    }
    if (p.is_pml) {
      AdjPml A;
      A.part = i;
      A.n = n;
      float* blk;
      if ((r = c.dalloc((void**)&blk, size_t(12 * n) * 4))) return r;
      for (int s = 0; s < 2; ++s)
        for (int q = 0; q < 6; ++q) A.set[s][q] = blk + (6 * s + q) * n;
      std::vector<float> prof = pml_profile(p, c.prm.pml_layers, c.prm.h, c.prm.c0, 1.0);
      float* dp;
      if ((r = c.dalloc((void**)&dp, prof.size() * 4)) || (r = c.upload(dp, prof.data(), prof.size() * 4))) return r;
      A.prof = dp;
      for (int s = 0; s < 2; ++s) {
        acc[s][i].kind = 1;
        acc[s][i].p[0] = acc[s][i].p[1] = A.set[s][0];
        acc[s][i].f = A.set[s][5];
      }
      S.pml.push_back(A);
      const PmlBuffers& B = c.pml[i];
      S.regions.push_back(Region{B.p[0], size_t(9 * n)});  // p[2], phi[2][3], f
    } else {
      AdjDct A;
      A.part = i;
      A.n = n;
      float* blk;
      if ((r = c.dalloc((void**)&blk, size_t(4 * n) * 4))) return r;
      A.lM = blk; A.lMp = blk + n; A.lF = blk + 2 * n; A.lp = blk + 3 * n;
      const float *fx, *ix, *fy, *iy, *fz, *iz;
      if ((r = transposed_tables(c, p.w, &fx, &ix)) || (r = transposed_tables(c, p.h, &fy, &iy)) ||
          (r = transposed_tables(c, p.d, &fz, &iz)))
        return r;
      A.pass[0] = axis_pass(0, p, ix, A.lp, S.tmp[0]);
      A.pass[1] = axis_pass(1, p, iy, S.tmp[0], S.tmp[1]);
      A.pass[2] = axis_pass(2, p, iz, S.tmp[1], S.tmp[2]);
      A.pass[3] = axis_pass(0, p, fx, S.tmp[2], S.tmp[0]);
      A.pass[4] = axis_pass(1, p, fy, S.tmp[0], S.tmp[1]);
      A.pass[5] = axis_pass(2, p, fz, S.tmp[1], S.tmp[3]);
      for (int k = 0; k < 6; ++k) A.tiles[k] = int(cdiv(A.pass[k].M, 64) * A.pass[k].tiles_n);
      if ((r = c.dalloc((void**)&A.d_pass, sizeof A.pass)) || (r = c.upload(A.d_pass, A.pass, sizeof A.pass)))
        return r;
      for (int s = 0; s < 2; ++s) {
        acc[s][i].kind = 0;
        acc[s][i].p[0] = acc[s][i].p[1] = A.lp;
        acc[s][i].f = A.lF;
      }
      S.dct.push_back(A);
      S.regions.push_back(Region{c.dct[i].F, size_t(4 * n)});  // F, p, M[0], M[1]
    }
  }
  for (const Region& g : S.regions) S.ck_floats += g.n;
  for (int s = 0; s < 2; ++s)
    if ((r = c.dalloc((void**)&S.d_parts[s], np * sizeof(PartAccess))) ||
        (r = c.upload(S.d_parts[s], acc[s].data(), np * sizeof(PartAccess))))
      return r;
  // interface: the forward descriptors with last-writer masks, one launch
  std::vector<BndDesc> bd = c.h_bnd;
  std::vector<uint8_t> mask;
  std::vector<long long> mstart;
  final_writer_masks(c.topo, bd, c.bnd_elems, mask, mstart);
  uint8_t* dm = nullptr;
  if (!mask.empty() && ((r = c.dalloc((void**)&dm, mask.size())) || (r = c.upload(dm, mask.data(), mask.size()))))
    return r;
  long long e0 = 0;
  for (size_t b = 0; b < bd.size(); ++b) {
    bd[b].mask = dm + mstart[b];
    bd[b].elem0 = e0;
    e0 += c.bnd_elems[b];
  }
  S.bnd_total = e0;
  S.nbnd = int(bd.size());
  if (!bd.empty() && ((r = c.dalloc((void**)&S.d_bnd, bd.size() * sizeof(BndDesc))) ||
                      (r = c.upload(S.d_bnd, bd.data(), bd.size() * sizeof(BndDesc)))))
    return r;
  // recorders and sources read and write the adjoint arrays
  auto adct = [&](int part) -> AdjDct& {
    for (AdjDct& A : S.dct)
      if (A.part == part) return A;
    return S.dct[0];
  };
  std::vector<RecDesc> rec;
  for (const RecorderDef& rd : c.topo.recorders) {
    const PartGeom& p = c.topo.parts[rd.part];
    RecDesc d{};
    d.p = adct(rd.part).lp;
    d.cell = ((long long)rd.lz * p.h + rd.ly) * p.w + rd.lx;
    rec.push_back(d);
  }
  if (!rec.empty() && ((r = c.dalloc((void**)&S.d_rec, rec.size() * sizeof(RecDesc))) ||
                       (r = c.upload(S.d_rec, rec.data(), rec.size() * sizeof(RecDesc)))))
    return r;
  std::vector<AdjSource> src;
  for (size_t i = 0; i < c.topo.sources.size(); ++i) {
    const SourceDef& s = c.topo.sources[i];
    if (s.part < 0) continue;
    const PartGeom& p = c.topo.parts[s.part];
    AdjSource a{};
    a.lF = adct(s.part).lF;
    a.cell = ((long long)(s.z - p.z0) * p.h + (s.y - p.y0)) * p.w + (s.x - p.x0);
    a.index = int(i);
    src.push_back(a);
  }
  S.nsrc = int(src.size());
  if (!src.empty() && ((r = c.dalloc((void**)&S.d_src, src.size() * sizeof(AdjSource))) ||
                       (r = c.upload(S.d_src, src.data(), src.size() * sizeof(AdjSource)))))
    return r;
  return c.dalloc((void**)&S.d_dtheta, 6 * sizeof(double));
}

int copy_regions(Ctx& c, float* store, bool save) {
  size_t o = 0;
  for (const Region& g : c.adj->regions) {
    hipError_t e = save ? hipMemcpyAsync(store + o, g.p, g.n * 4, hipMemcpyDeviceToDevice, c.stream)
                        : hipMemcpyAsync(g.p, store + o, g.n * 4, hipMemcpyDeviceToDevice, c.stream);
    if (e != hipSuccess) return hip_fail(e, "checkpoint copy");
    o += g.n;
  }
  return ARD_OK;
}

int set_step(Ctx& c, int n) {
  c.time_step = n;
  hipError_t e = hipMemcpyAsync(c.d_step, &n, sizeof(int), hipMemcpyHostToDevice, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);  // n lives on the host stack
  return e == hipSuccess ? ARD_OK : hip_fail(e, "step counter");
}

// A batched context's adjoint runs on a plain-form twin: both forms compute
// the same discrete step (the same DCT-II of the same force field, modal
// update, PML and interface forcing; only the order of floating-point
// additions differs), and the transposed kernels work on the plain form's
// dense fields. The twin is built once, from the context's resolved topology,
// parameters and PML damping scales (all fixed at finalize), and keeps its own
// device state; the batched context's state is not touched.
int plain_twin(Ctx& c, Ctx** out) {
  if (!c.twin) {
    auto* t = new Ctx();
    t->device = c.device;
    t->prm = c.prm;
    t->prm.form = ARD_FORM_PLAIN;
    t->prm.autotune = 0;
    for (int i = 0; i < 6; ++i) t->pml_scale[i] = c.pml_scale[i];
    t->topo = c.topo;
    t->topology_built = true;
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      t->stream = nullptr;
      delete t;
      return hip_fail(e, "hipStreamCreate");
    }
    const int r = finalize_plain(*t);
    if (r != ARD_OK) {
      delete t;
      return r;
    }
    t->finalized = true;
    c.twin = t;
  }
  *out = c.twin;
  return ARD_OK;
}

}  // namespace

}  // namespace ard

using namespace ard;

extern "C" int ard_adjoint(ard_ctx* ctx, int nsteps, const float* dl_drec, float* dl_dwave, float* dl_dscale,
                           int checkpoint_stride) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return set_error(ARD_ERR_INVALID, "null context");
  if (!c->finalized) return set_error(ARD_ERR_STATE, "not finalized");
  if (c->device < 0) return set_error(ARD_ERR_STATE, "host-only context");
  if (nsteps <= 0 || !dl_drec) return set_error(ARD_ERR_INVALID, "nsteps and dl_drec are required");
  hipSetDevice(c->device);
  int r;
  if (c->prm.form != ARD_FORM_PLAIN) {
    Ctx* t = nullptr;
    if ((r = plain_twin(*c, &t))) return r;
    r = ard_adjoint(reinterpret_cast<ard_ctx*>(t), nsteps, dl_drec, dl_dwave, dl_dscale, checkpoint_stride);
    hipSetDevice(c->device);
    return r;
  }
  if (!c->adj && (r = setup_adjoint(*c))) return r;
  AdjState& S = *c->adj;
  const int nrec = int(c->topo.recorders.size());
  const int nsrc_all = int(c->topo.sources.size());
  const int K = checkpoint_stride > 0 ? checkpoint_stride : std::max(1, int(std::ceil(std::sqrt(double(nsteps)))));
  const int nck = int(cdiv(nsteps, K));
  long long pml_seg = 0;  // floats of saved PML state per step
  for (const AdjPml& A : S.pml) pml_seg += 5 * A.n;
  // checkpoints (nck + the final state) and one segment of PML states
  // (all per-call buffers are freed by cleanup(), whatever the exit path)
  float *ck = nullptr, *seg = nullptr, *d_g = nullptr, *d_du = nullptr;
  AdjPmlDesc* d_pd = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(c->stream);
    if (ck) hipFree(ck);
    if (seg) hipFree(seg);
    if (d_g) hipFree(d_g);
    if (d_du) hipFree(d_du);
    if (d_pd) hipFree(d_pd);
  };
  hipError_t e = hipMalloc((void**)&ck, std::max<size_t>(size_t(nck + 1) * S.ck_floats * 4, 4));
  if (e == hipSuccess) e = hipMalloc((void**)&seg, std::max<size_t>(size_t(K) * pml_seg * 4, 4));
  if (e == hipSuccess) e = hipMalloc((void**)&d_g, std::max<size_t>(size_t(nsteps) * nrec * 4, 4));
  if (e == hipSuccess) e = hipMalloc((void**)&d_du, std::max<size_t>(size_t(nsrc_all) * nsteps * 4, 4));
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "adjoint buffers");
  }
  auto fail_with = [&](int code) {
    cleanup();
    return code;
  };
  if ((r = ensure_trace(*c, nsteps))) return fail_with(r);
  // forward from the zero state, checkpoint every K steps
  for (const Region& g : S.regions)
    if ((e = hipMemsetAsync(g.p, 0, g.n * 4, c->stream)) != hipSuccess) return fail_with(hip_fail(e, "reset"));
  if ((r = set_step(*c, 0))) return fail_with(r);
  for (int n = 0; n < nsteps; ++n) {
    if (n % K == 0 && (r = copy_regions(*c, ck + size_t(n / K) * S.ck_floats, true))) return fail_with(r);
    if ((r = step_plain(*c, 1))) return fail_with(r);
  }
  if ((r = copy_regions(*c, ck + size_t(nck) * S.ck_floats, true))) return fail_with(r);
  // adjoint state: zero, loss gradient on the device
  if (nrec && (e = hipMemcpyAsync(d_g, dl_drec, size_t(nsteps) * nrec * 4, hipMemcpyHostToDevice, c->stream)) !=
                  hipSuccess)
    return fail_with(hip_fail(e, "dl_drec upload"));
  for (AdjDct& A : S.dct)
    if ((e = hipMemsetAsync(A.lM, 0, size_t(4 * A.n) * 4, c->stream)) != hipSuccess) return fail_with(hip_fail(e, "zero"));
  for (AdjPml& A : S.pml)
    if ((e = hipMemsetAsync(A.set[0][0], 0, size_t(12 * A.n) * 4, c->stream)) != hipSuccess)
      return fail_with(hip_fail(e, "zero"));
  if ((e = hipMemsetAsync(d_du, 0, size_t(nsrc_all) * nsteps * 4, c->stream)) != hipSuccess ||
      (e = hipMemsetAsync(S.d_dtheta, 0, 6 * sizeof(double), c->stream)) != hipSuccess)
    return fail_with(hip_fail(e, "zero"));
  // PML descriptors of the reverse step for (input set, segment slot)
  std::vector<AdjPmlDesc> pd;
  std::vector<long long> pblocks;
  long long blocks = 0;
  for (const AdjPml& A : S.pml) {
    pblocks.push_back(blocks);
    blocks += cdiv(A.n, 256);
  }
  const size_t npd = S.pml.size();
  if (npd) {
    for (int s = 0; s < 2; ++s)
      for (int k = 0; k < K; ++k) {
        long long o = 0;
        for (size_t i = 0; i < npd; ++i) {
          const AdjPml& A = S.pml[i];
          const PartGeom& p = c->topo.parts[A.part];
          AdjPmlDesc d{};
          d.W = p.w; d.H = p.h; d.D = p.d;
          d.axis = p.pml_type / 2;
          d.face = p.pml_type;
          d.zeta = c->pml[A.part].zprof;
          d.prof = A.prof;
          d.lP = A.set[s][0]; d.lQ = A.set[s][1];
          for (int a = 0; a < 3; ++a) d.lPhi[a] = A.set[s][2 + a];
          d.op = A.set[s ^ 1][0]; d.oq = A.set[s ^ 1][1];
          for (int a = 0; a < 3; ++a) d.ophi[a] = A.set[s ^ 1][2 + a];
          d.of = A.set[s ^ 1][5];
          const float* st = seg + size_t(k) * pml_seg + o;
          d.p = st; d.q = st + A.n;
          for (int a = 0; a < 3; ++a) d.phi[a] = st + (2 + a) * A.n;
          d.block0 = pblocks[i];
          o += 5 * A.n;
          pd.push_back(d);
        }
      }
    if ((e = hipMalloc((void**)&d_pd, pd.size() * sizeof(AdjPmlDesc))) != hipSuccess)
      return fail_with(hip_fail(e, "adjoint PML descriptors"));
    if ((r = c->upload(d_pd, pd.data(), pd.size() * sizeof(AdjPmlDesc)))) return fail_with(r);
  }
  int set = 0;
  const float b = 0.999f;
  for (int sgi = nck - 1; sgi >= 0; --sgi) {
    const int n0 = sgi * K, n1 = std::min(n0 + K, nsteps);
    if ((r = copy_regions(*c, ck + size_t(sgi) * S.ck_floats, false)) || (r = set_step(*c, n0))) return fail_with(r);
    for (int n = n0; n < n1; ++n) {
      // PML state before step n: p (current), q (previous), phi
      const int par = n & 1;
      long long o = 0;
      for (const AdjPml& A : S.pml) {
        const PmlBuffers& B = c->pml[A.part];
        float* st = seg + size_t(n - n0) * pml_seg + o;
        const float* src[5] = {B.p[par], B.p[par ^ 1], B.phi[par][0], B.phi[par][1], B.phi[par][2]};
        for (int q = 0; q < 5; ++q)
          if ((e = hipMemcpyAsync(st + q * A.n, src[q], size_t(A.n) * 4, hipMemcpyDeviceToDevice, c->stream)) !=
              hipSuccess)
            return fail_with(hip_fail(e, "segment copy"));
        o += 5 * A.n;
      }
      if ((r = step_plain(*c, 1))) return fail_with(r);
    }
    for (int n = n1 - 1; n >= n0; --n) {
      int lr = 0;
      for (AdjDct& A : S.dct)
        if (hipMemsetAsync(A.lp, 0, size_t(A.n) * 4, c->stream) != hipSuccess) lr = -2;
      lr |= launch_adj_record(S.d_rec, nrec, d_g + size_t(n) * nrec, c->stream);
      lr |= launch_adj_interface(S.d_bnd, S.nbnd, S.d_parts[set], S.bnd_total, c->stream);
      if (npd)
        lr |= launch_adj_pml(d_pd + (size_t(set) * K + (n - n0)) * npd, int(npd), blocks, c->prm.h, c->prm.dt,
                             c->prm.c0, S.d_dtheta, c->stream);
      for (AdjDct& A : S.dct) {
        const DctPlain& D = c->dct[A.part];
        for (int k = 0; k < 3; ++k)
          lr |= launch_gemm(A.d_pass + k, 1, A.tiles[k], k == 0, k == 0, 0, c->stream);
        lr |= launch_adj_modal(S.tmp[2], A.lM, A.lMp, D.a, D.g, b, A.n, c->stream);
        for (int k = 3; k < 6; ++k)
          lr |= launch_gemm(A.d_pass + k, 1, A.tiles[k], k == 3, k == 3, 0, c->stream);
        lr |= launch_adj_add(A.lF, S.tmp[3], A.n, c->stream);
      }
      lr |= launch_adj_sources(S.d_src, S.nsrc, d_du, n, nsteps, c->stream);
      if (lr) return fail_with(set_error(ARD_ERR_HIP, "adjoint step launch"));
      set ^= 1;
    }
  }
  // leave the context in its forward state after nsteps
  if ((r = copy_regions(*c, ck + size_t(nck) * S.ck_floats, false)) || (r = set_step(*c, nsteps)))
    return fail_with(r);
  if (dl_dwave && nsrc_all &&
      (e = hipMemcpyAsync(dl_dwave, d_du, size_t(nsrc_all) * nsteps * 4, hipMemcpyDeviceToHost, c->stream)) !=
          hipSuccess)
    return fail_with(hip_fail(e, "gradient download"));
  double th[6];
  if ((e = hipMemcpyAsync(th, S.d_dtheta, sizeof th, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
    return fail_with(hip_fail(e, "gradient download"));
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail_with(hip_fail(e, "hipStreamSynchronize"));
  if (dl_dscale)
    for (int f = 0; f < 6; ++f) dl_dscale[f] = float(th[f]);
  cleanup();
  return ARD_OK;
}
