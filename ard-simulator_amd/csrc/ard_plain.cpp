// Plain form of the ARD step: the reference's structure (one set of launches
// per partition and per boundary, dense pressure and force volumes) on the
// device. Also the set-up shared with the batched form (PML slabs, sources,
// recorders, boundary descriptors).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ard_ctx_impl.h"
#include "ard_kernels.h"

namespace ard {

namespace {
long long ceil_div(long long a, long long b) { return (a + b - 1) / b; }
}  // namespace

// zeta along the normal of a PML slab (ARD/pml_partition.cpp:28-92): the
// profile zeta0 (x/L - sin(2 pi x/L)/(2 pi)), x = q h, growing outward, with
// zeta0 = c0/L log10(1/R), R = 1e-5, L = layers*h.
std::vector<float> pml_profile(const PartGeom& p, int layers, double h, double c0, double scale) {
  const int axis = p.pml_type / 2;
  const int n = axis == 0 ? p.w : (axis == 1 ? p.h : p.d);
  const double L = layers * h;
  const double z0 = c0 / L * std::log10(1.0 / double(0.00001f));
  const double pi = 3.14159265358979323846;
  std::vector<float> prof(n);
  const bool grows = (p.pml_type == P_RIGHT || p.pml_type == P_BOTTOM || p.pml_type == P_BACK);
  for (int i = 0; i < n; ++i) {
    const int q = grows ? i + 1 : n - i;
    const double x = q * h;
    prof[i] = float(scale * z0 * (x / L - std::sin(2.0 * pi * x / L) / (2.0 * pi)));
  }
  return prof;
}

int ensure_trace(Ctx& c, int steps_needed) {
  const int nr = int(c.topo.recorders.size());
  if (steps_needed <= c.trace_cap) return ARD_OK;
  const int cap = std::max({steps_needed, 2 * c.trace_cap, 1024});
  float* nt = nullptr;
  int r = c.dalloc((void**)&nt, size_t(cap) * std::max(nr, 1) * 4);
  if (r) return r;
  if (c.d_trace && nr > 0) {
    hipError_t e = hipMemcpyAsync(nt, c.d_trace, size_t(c.trace_cap) * nr * 4, hipMemcpyDeviceToDevice, c.stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2D");
    e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    hipFree(c.d_trace);
    c.allocs.erase(std::find(c.allocs.begin(), c.allocs.end(), (void*)c.d_trace));
  }
  c.d_trace = nt;
  if (c.cubes && nr > 0) {  // the cube rows grow with the point trace
    const size_t row = size_t(nr) * 1000 * 4;
    float* nc = nullptr;
    if ((r = c.dalloc((void**)&nc, size_t(cap) * row))) return r;
    if (c.d_cube) {
      hipError_t e = hipMemcpyAsync(nc, c.d_cube, size_t(c.trace_cap) * row, hipMemcpyDeviceToDevice, c.stream);
      if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync D2D");
      e = hipStreamSynchronize(c.stream);
      if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
      hipFree(c.d_cube);
      c.allocs.erase(std::find(c.allocs.begin(), c.allocs.end(), (void*)c.d_cube));
    }
    c.d_cube = nc;
  }
  c.trace_cap = cap;
  return ARD_OK;
}

// Resolves each boundary's sides the way Boundary::ComputeForcingTerms does
// (ARD/boundary.cpp:30-34, 78-82, 126-130) and lays out its face elements in
// the reference's loop order.
int build_boundary_descs(Ctx& c, std::vector<BndDesc>& out, std::vector<long long>& elems) {
  int code = ARD_OK;
  const std::string err = boundary_descs(c.topo, c.prm.h, c.prm.c0, out, elems, &code);
  return err.empty() ? ARD_OK : set_error(code, err);
}

int setup_common(Ctx& c) {
  int r;
  if ((r = c.dalloc((void**)&c.d_step, sizeof(int)))) return r;
  if ((r = c.dzero(c.d_step, sizeof(int)))) return r;
  c.courant = c.prm.c0 * c.prm.dt / c.prm.h;
  // PML slabs
  const int np = int(c.topo.parts.size());
  c.pml.assign(np, PmlBuffers());
  c.h_pml.clear();
  c.pml_blocks.clear();
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = c.topo.parts[i];
    if (!p.is_pml) continue;
    PmlBuffers& B = c.pml[i];
    const size_t n = size_t(p.cells());
    if (n >= (size_t(1) << 29))  // the PML kernels: 32-bit byte offsets within a slab
      return set_error(ARD_ERR_UNSUPPORTED, "a PML slab of 2^29 or more cells");
    float* blk = nullptr;
    if ((r = c.dalloc((void**)&blk, n * 9 * 4))) return r;
    if ((r = c.dzero(blk, n * 9 * 4))) return r;
    B.p[0] = blk; B.p[1] = blk + n;
    for (int q = 0; q < 3; ++q) { B.phi[0][q] = blk + (2 + q) * n; B.phi[1][q] = blk + (5 + q) * n; }
    B.f = blk + 8 * n;
    std::vector<float> prof = pml_profile(p, c.prm.pml_layers, c.prm.h, c.prm.c0, c.pml_scale[p.pml_type]);
    if ((r = c.dalloc((void**)&B.zprof, prof.size() * 4))) return r;
    if ((r = c.upload(B.zprof, prof.data(), prof.size() * 4))) return r;
    PmlDesc d{};
    d.W = p.w; d.H = p.h; d.D = p.d; d.axis = p.pml_type / 2;
    d.p[0] = B.p[0]; d.p[1] = B.p[1];
    for (int s = 0; s < 2; ++s)
      for (int q = 0; q < 3; ++q) d.phi[s][q] = B.phi[s][q];
    d.f = B.f;
    d.zprof = B.zprof;
    d.block0 = 0;
    const long long nblk = pml_slab_blocks(d, c.prm.pml_layers);
    B.desc = int(c.h_pml.size());
    c.h_pml.push_back(d);
    c.pml_blocks.push_back(nblk);
  }
  if (!c.h_pml.empty()) {
    if ((r = c.dalloc((void**)&c.d_pml, c.h_pml.size() * sizeof(PmlDesc)))) return r;
    if ((r = c.upload(c.d_pml, c.h_pml.data(), c.h_pml.size() * sizeof(PmlDesc)))) return r;
    // and every slab in one launch (the plain form's step)
    std::vector<PmlDesc> g = c.h_pml;
    long long b0 = 0;
    for (size_t k = 0; k < g.size(); ++k) {
      g[k].block0 = b0;
      b0 += c.pml_blocks[k];
    }
    c.pml_grp_blocks = b0;
    if ((r = c.dalloc((void**)&c.d_pml_grp, g.size() * sizeof(PmlDesc))) ||
        (r = c.upload(c.d_pml_grp, g.data(), g.size() * sizeof(PmlDesc))))
      return r;
  }
  return ensure_trace(c, 1024);
}

int finalize_plain(Ctx& c) {
  int r;
  if ((r = setup_common(c))) return r;
  const int np = int(c.topo.parts.size());
  c.dct.assign(np, DctPlain());
  size_t maxn = 0;
  for (int i = 0; i < np; ++i)
    if (!c.topo.parts[i].is_pml) maxn = std::max(maxn, size_t(c.topo.parts[i].cells()));
  // pass intermediates: shared by the partitions, which run one after the
  // other, or (folded passes, variant 3) each partition's own, since pass j
  // of every partition is one launch. 16 floats of slack: the folded passes'
  // 16-byte operand loads may end up to 3 floats past an operand array.
  const bool grouped = g_gemm_variant == 3 || g_gemm_variant == 5;
  if (!grouped &&
      ((r = c.dalloc((void**)&c.tmp[0], (maxn + 16) * 4)) || (r = c.dalloc((void**)&c.tmp[1], (maxn + 16) * 4))))
    return r;
  c.h_gemm.clear();
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = c.topo.parts[i];
    if (p.is_pml) continue;
    DctPlain& D = c.dct[i];
    const size_t n = size_t(p.cells());
    float* blk = nullptr;
    if ((r = c.dalloc((void**)&blk, n * 6 * 4))) return r;
    if ((r = c.dzero(blk, n * 4 * 4))) return r;
    D.F = blk; D.p = blk + n; D.M[0] = blk + 2 * n; D.M[1] = blk + 3 * n; D.a = blk + 4 * n; D.g = blk + 5 * n;
    D.t[0] = c.tmp[0];
    D.t[1] = c.tmp[1];
    if (grouped) {
      float* tb = nullptr;
      if ((r = c.dalloc((void**)&tb, (2 * n + 32) * 4))) return r;
      D.t[0] = tb;
      D.t[1] = tb + n + 16;
    }
    std::vector<float> a, g;
    modal_tables(p.w, p.h, p.d, c.prm.h, c.prm.dt, c.prm.c0, a, g);
    if ((r = c.upload(D.a, a.data(), n * 4)) || (r = c.upload(D.g, g.data(), n * 4))) return r;
    const float *fw, *iw, *fh, *ih, *fd, *id;
    if ((r = c.cos_tables(p.w, &fw, &iw)) || (r = c.cos_tables(p.h, &fh, &ih)) || (r = c.cos_tables(p.d, &fd, &id)))
      return r;
    const long long W = p.w, H = p.h, Dd = p.d;
    D.gemm0 = int(c.h_gemm.size());
    for (int par = 0; par < 2; ++par) {
      float* cur = D.M[par];
      float* nxt = D.M[par ^ 1];
      // pass along x: B(k=x, n=row) contiguous in k; C(m=kx, n=row) contiguous in m
      auto xpass = [&](const float* A, const float* in, float* out) {
        GemmDesc g{};
        g.A = A; g.B = in; g.C = out;
        g.M = p.w; g.K = p.w; g.N = int(H * Dd); g.N1 = g.N;
        g.bK = 1; g.b1 = W; g.b2 = 0;
        g.cM = 1; g.c1 = W; g.c2 = 0;
        g.alpha = 1.f;
        return g;
      };
      // pass along y: per z-plane; n = (x, z)
      auto ypass = [&](const float* A, const float* in, float* out) {
        GemmDesc g{};
        g.A = A; g.B = in; g.C = out;
        g.M = p.h; g.K = p.h; g.N = int(W * Dd); g.N1 = p.w;
        g.bK = W; g.b1 = 1; g.b2 = W * H;
        g.cM = W; g.c1 = 1; g.c2 = W * H;
        g.alpha = 1.f;
        return g;
      };
      // pass along z: n = (x, y) flattened
      auto zpass = [&](const float* A, const float* in, float* out) {
        GemmDesc g{};
        g.A = A; g.B = in; g.C = out;
        g.M = p.d; g.K = p.d; g.N = int(W * H); g.N1 = g.N;
        g.bK = W * H; g.b1 = 1; g.b2 = 0;
        g.cM = W * H; g.c1 = 1; g.c2 = 0;
        g.alpha = 1.f;
        return g;
      };
      GemmDesc f0 = xpass(fw, D.F, D.t[0]);
      GemmDesc f1 = ypass(fh, D.t[0], D.t[1]);
      GemmDesc f2 = zpass(fd, D.t[1], nxt);  // modal update into the previous-modes buffer
      f2.ep_cur = cur; f2.ep_a = D.a; f2.ep_g = D.g; f2.ep_b = 0.999f;
      GemmDesc i0 = xpass(iw, nxt, D.t[0]);
      GemmDesc i1 = ypass(ih, D.t[0], D.t[1]);
      GemmDesc i2 = zpass(id, D.t[1], D.p);
      for (GemmDesc* g : {&f0, &f1, &f2, &i0, &i1, &i2}) {
        g->tiles_n = int(ceil_div(g->N, 64));
        g->tile0 = 0;
        g->b_elems = unsigned(n);  // every operand array holds the partition's n cells
        c.h_gemm.push_back(*g);
      }
      if (par == 0) {
        const GemmDesc* gs[6] = {&f0, &f1, &f2, &i0, &i1, &i2};
        for (int j = 0; j < 6; ++j) D.tiles[j] = int(ceil_div(gs[j]->M, 64) * gs[j]->tiles_n);
      }
    }
  }
  if ((r = c.dalloc((void**)&c.d_gemm, c.h_gemm.size() * sizeof(GemmDesc)))) return r;
  if ((r = c.upload(c.d_gemm, c.h_gemm.data(), c.h_gemm.size() * sizeof(GemmDesc)))) return r;
  // transform kernel of the passes, fixed here: the FFT-style passes
  // (variant 4) when every axis length factors into radices up to 13, else
  // the dense passes; variants 3 and 5 (the batched form's folded face
  // passes) mean the even/odd folded matrix-core passes (k_gemm_mfma_fold,
  // the modal update fused into the last forward pass as in variant 1)
  c.plain_variant = grouped ? 3 : g_gemm_variant;
  if (c.plain_variant == 3 && !c.h_gemm.empty()) {
    for (int par = 0; par < 2; ++par)
      for (int j = 0; j < 6; ++j) {
        std::vector<GemmDesc> f;
        for (const DctPlain& D : c.dct)
          if (D.F) f.push_back(c.h_gemm[size_t(D.gemm0 + par * 6 + j)]);
        Ctx::PassGroup& G = c.pass_group[par][j];
        G.blocks = 0;
        for (GemmDesc& g : f) {
          if ((r = fold_twin(c, g, j >= 3))) return r;
          G.blocks += int(ceil_div(g.M, 64)) * g.tiles_n;
        }
        if ((r = fold_pass_tiles(c, f, &G.tiles)) || (r = c.dalloc((void**)&G.d, f.size() * sizeof(GemmDesc))) ||
            (r = c.upload(G.d, f.data(), f.size() * sizeof(GemmDesc))))
          return r;
      }
  }
  if (c.plain_variant == 4 && !c.h_gemm.empty()) {
    std::vector<GemmDesc> f = c.h_gemm;
    bool ok = true;
    for (size_t j = 0; j < f.size() && ok; ++j) ok = fft_setup(c, f[j], j % 6 >= 3) == ARD_OK;
    if (ok) {
      c.fft_lds = 0;
      for (const GemmDesc& g : f) c.fft_lds = std::max(c.fft_lds, fft_lds_bytes(g));
      for (size_t i = 0; i < c.dct.size(); ++i)
        if (c.dct[i].F)
          for (int j = 0; j < 6; ++j) c.dct[i].fft_tiles[j] = fft_tiles(f[c.dct[i].gemm0 + j]);
      if ((r = c.dalloc((void**)&c.d_fft, f.size() * sizeof(GemmDesc))) ||
          (r = c.upload(c.d_fft, f.data(), f.size() * sizeof(GemmDesc))))
        return r;
    } else {
      c.plain_variant = 1;
    }
  }

  // partition access table for the interface kernel
  c.h_parts.assign(np, PartAccess());
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = c.topo.parts[i];
    PartAccess& A = c.h_parts[i];
    A.W = p.w; A.H = p.h; A.D = p.d;
    if (p.is_pml) {
      A.kind = 1;
      A.p[0] = c.pml[i].p[0]; A.p[1] = c.pml[i].p[1]; A.f = c.pml[i].f;
    } else {
      A.kind = 0;
      A.p[0] = A.p[1] = c.dct[i].p; A.f = c.dct[i].F;
    }
  }
  if ((r = c.dalloc((void**)&c.d_parts, np * sizeof(PartAccess)))) return r;
  if ((r = c.upload(c.d_parts, c.h_parts.data(), np * sizeof(PartAccess)))) return r;
  if ((r = build_boundary_descs(c, c.h_bnd, c.bnd_elems))) return r;
  if (!c.h_bnd.empty()) {
    if ((r = c.dalloc((void**)&c.d_bnd, c.h_bnd.size() * sizeof(BndDesc)))) return r;
    if ((r = c.upload(c.d_bnd, c.h_bnd.data(), c.h_bnd.size() * sizeof(BndDesc)))) return r;
    // every boundary in one launch: the elements of each start on a block,
    // and each force cell is written by its last writer in list order only
    std::vector<BndDesc> g = c.h_bnd;
    std::vector<uint8_t> mask;
    std::vector<long long> mstart;
    final_writer_masks(c.topo, g, c.bnd_elems, mask, mstart);
    uint8_t* dm = nullptr;
    if ((r = c.dalloc((void**)&dm, std::max<size_t>(mask.size(), 1))) || (r = c.upload(dm, mask.data(), mask.size())))
      return r;
    long long e0 = 0;
    for (size_t b = 0; b < g.size(); ++b) {
      g[b].mask = dm + mstart[b];
      g[b].elem0 = e0;
      e0 += ceil_div(c.bnd_elems[b], 256) * 256;
    }
    c.bnd_grp_total = e0;
    if ((r = c.dalloc((void**)&c.d_bnd_grp, g.size() * sizeof(BndDesc))) ||
        (r = c.upload(c.d_bnd_grp, g.data(), g.size() * sizeof(BndDesc))))
      return r;
  }

  // sources: ARD/partition.cpp:187-197 writes the sample into the owner's force volume
  std::vector<SourceDesc> src;
  for (const SourceDef& s : c.topo.sources) {
    if (s.part < 0) continue;  // outside every DCT partition: never assigned (ARD/simulation.cpp:37-48)
    const PartGeom& p = c.topo.parts[s.part];
    SourceDesc d{};
    d.part = s.part;
    d.cell = ((long long)(s.z - p.z0) * p.h + (s.y - p.y0)) * p.w + (s.x - p.x0);
    d.kind = s.wave.empty() ? 0 : 1;
    d.n = int(s.wave.size());
    if (d.kind == 1) {
      float* w;
      if ((r = c.dalloc((void**)&w, s.wave.size() * 4))) return r;
      if ((r = c.upload(w, s.wave.data(), s.wave.size() * 4))) return r;
      d.wave = w;
    }
    d.fptr = c.dct[s.part].F + d.cell;
    src.push_back(d);
  }
  c.nsrc = int(src.size());
  if (c.nsrc) {
    if ((r = c.dalloc((void**)&c.d_src, src.size() * sizeof(SourceDesc)))) return r;
    if ((r = c.upload(c.d_src, src.data(), src.size() * sizeof(SourceDesc)))) return r;
  }
  std::vector<RecDesc> rec;
  for (const RecorderDef& rd : c.topo.recorders) {
    const PartGeom& p = c.topo.parts[rd.part];
    RecDesc d{};
    d.p = c.dct[rd.part].p;
    d.cell = ((long long)rd.lz * p.h + rd.ly) * p.w + rd.lx;
    rec.push_back(d);
  }
  if (!rec.empty()) {
    if ((r = c.dalloc((void**)&c.d_rec, rec.size() * sizeof(RecDesc)))) return r;
    if ((r = c.upload(c.d_rec, rec.data(), rec.size() * sizeof(RecDesc)))) return r;
  }
  hipError_t e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

// One Simulation::Update per iteration (ARD/simulation.cpp:300-317): sources,
// every partition's update, then the boundaries in list order (the later
// writer wins on shared cells), then the recorder sample.
// One plain step of parity par: sources, the partitions in list order, PML
// slabs, boundaries in list order (the last writer wins on shared cells), then
// the recorder sample. With a timer, each phase is bracketed by events.
static int enqueue_plain(Ctx& c, int par, PhaseTimer* T) {
  const int nr = int(c.topo.recorders.size());
  auto begin = [&](int ph) { return T ? T->begin(c.stream, ph) : -1; };
  auto end = [&](int ev) { if (T) T->end(c.stream, ev); };
  int ev = begin(ARD_PH_SOURCES);
  if (launch_sources(c.d_src, c.nsrc, c.d_step, c.courant, c.stream)) return set_error(ARD_ERR_HIP, "sources");
  end(ev);
  if (c.plain_variant == 3) {
    // pass j of every partition in one launch (the partitions' passes are
    // independent; each has its own intermediates)
    int r = 0;
    for (int j = 0; j < 6; ++j) {
      if (j == 0) ev = begin(ARD_PH_DCT_FWD);
      if (j == 3) ev = begin(ARD_PH_DCT_INV);
      const Ctx::PassGroup& G = c.pass_group[par][j];
      r |= launch_gemm_fold(G.d, G.tiles, G.blocks, j % 3 == 0, j < 3 ? 1 : 2, c.stream, j == 2 ? 1 : 0);
      if (j == 2 || j == 5) end(ev);
    }
    if (r) return set_error(ARD_ERR_HIP, "DCT pass launch");
  } else {
    for (size_t i = 0; i < c.dct.size(); ++i) {
      const DctPlain& D = c.dct[i];
      if (!D.F) continue;
      const int off = D.gemm0 + par * 6;
      const bool fft = c.plain_variant == 4;
      // pass j: x (B and C contiguous along the axis), then y, then z (with
      // the modal update in the forward z pass)
      auto pass = [&](int j) {
        const bool kc = j % 3 == 0;
        const int ep = j == 2 ? 1 : 0;
        return fft ? launch_fft(c.d_fft + off + j, 1, D.fft_tiles[j], kc, kc, ep, c.fft_lds, c.stream)
                   : launch_gemm(c.d_gemm + off + j, 1, D.tiles[j], kc, kc, ep, c.stream, c.plain_variant);
      };
      int r = 0;
      ev = begin(ARD_PH_DCT_FWD);
      for (int j = 0; j < 3; ++j) r |= pass(j);
      end(ev);
      ev = begin(ARD_PH_DCT_INV);
      for (int j = 3; j < 6; ++j) r |= pass(j);
      end(ev);
      if (r) return set_error(ARD_ERR_HIP, "DCT pass launch");
    }
  }
  // every PML slab in one launch, then every boundary in one launch (its
  // last-writer masks give the list order's result: the reference's loop
  // writes each force cell last from its final writer, ARD/simulation.cpp:313-317)
  if (!c.h_pml.empty()) {
    ev = begin(ARD_PH_PML);
    if (launch_pml(c.d_pml_grp, int(c.h_pml.size()), c.pml_grp_blocks, par, c.prm.pml_layers, c.prm.h, c.prm.dt,
                   c.prm.c0, c.stream))
      return set_error(ARD_ERR_HIP, "PML launch");
    end(ev);
  }
  if (!c.h_bnd.empty()) {
    ev = begin(ARD_PH_INTERFACE);
    if (launch_interface(c.d_bnd_grp, int(c.h_bnd.size()), c.d_parts, par ^ 1, c.bnd_grp_total, c.stream))
      return set_error(ARD_ERR_HIP, "interface launch");
    end(ev);
  }
  ev = begin(ARD_PH_RECORD);
  if (enqueue_cubes(c, par, c.stream)) return set_error(ARD_ERR_HIP, "recorder cube launch");
  if (launch_record(c.d_rec, nr, c.d_trace, c.trace_cap, c.d_step, c.stream))
    return set_error(ARD_ERR_HIP, "recorder launch");
  end(ev);
  return ARD_OK;
}

// n steps: each a replay of the step graph of its parity (the launches of
// enqueue_plain captured once: no per-launch host overhead; rebuilt when the
// trace buffers move), or, with kernel timing on, direct launches.
int step_plain(Ctx& c, int n) {
  if (c.timer.on) {
    for (int s = 0; s < n; ++s) {
      int r = enqueue_plain(c, c.time_step & 1, &c.timer);
      if (r) return r;
      c.time_step++;
      if (c.timer.used > 2048 && (r = c.timer.collect())) return r;
    }
    return ARD_OK;
  }
  if (c.plain_graph_trace != c.d_trace || !c.graph_exec[0]) {
    for (int par = 0; par < 2; ++par) {
      if (c.graph_exec[par]) {
        hipGraphExecDestroy(c.graph_exec[par]);
        c.graph_exec[par] = nullptr;
      }
      hipGraph_t g;
      hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
      int r = enqueue_plain(c, par, nullptr);
      e = hipStreamEndCapture(c.stream, &g);
      if (r) return r;
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&c.graph_exec[par], g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    }
    c.plain_graph_trace = c.d_trace;
  }
  for (int s = 0; s < n; ++s) {
    hipError_t e = hipGraphLaunch(c.graph_exec[c.time_step & 1], c.stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    c.time_step++;
  }
  return ARD_OK;
}

int pressure_field_ptr(Ctx& c, int part, const float** out) {
  const PartGeom& p = c.topo.parts[part];
  if (p.is_pml) {
    *out = c.pml[part].p[c.time_step & 1];
    return ARD_OK;
  }
  if (c.prm.form == ARD_FORM_PLAIN) {
    *out = c.dct[part].p;
    return ARD_OK;
  }
  return batched_pressure_field(c, part, out);
}

int set_force_cell(Ctx& c, int part, int x, int y, int z, float f) {
  const PartGeom& p = c.topo.parts[part];
  float* dst = nullptr;
  const size_t idx = (size_t(z) * p.h + y) * p.w + x;
  if (p.is_pml) {
    dst = c.pml[part].f + idx;
  } else if (c.prm.form == ARD_FORM_PLAIN) {
    dst = c.dct[part].F + idx;
  } else {
    return batched_set_force(c, part, x, y, z, f);
  }
  return c.upload(dst, &f, 4);
}

}  // namespace ard
