// Batched face-shell form of the ARD step: every phase of Simulation::Update
// (ARD/simulation.cpp:300-317) for all partitions in one launch, and one step
// captured in a hipGraph (one graph per step parity, replayed alternately).
//
// Per step: PML update (beside) 2D DCT-II of the raw force face slabs (2
// grouped GEMM launches) -> modal pass (face spectra folded by the normal's
// parity, force spectrum expansion, modal update, reduction onto face planes
// and recorders) -> unfold -> 2D DCT-III
// to face-slab pressure (2 grouped GEMM launches) -> interface forcing with
// last-writer masks -> recorder samples and the next step's source samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ard_batched.h"
#include "ard_ctx_impl.h"
#include "ard_kernels.h"

namespace ard {

namespace {
long long ceil_div(long long a, long long b) { return (a + b - 1) / b; }
const double kPi = 3.14159265358979323846;

// 4-cell items when the plane length and the pointers allow 16-byte accesses
int vec_of(long long plane, std::initializer_list<const void*> ptrs) {
  if (plane % 4) return 1;
  for (const void* p : ptrs)
    if (reinterpret_cast<uintptr_t>(p) % 16) return 1;
  return 4;
}

}  // namespace

int apply_tiling(Ctx& c, int rpw, int kzc);
int autotune(Ctx& c);

struct BPart {
  // sizes of the three face-slab families: [6][D][H], [6][D][W], [6][H][W]
  long long plane[3] = {0, 0, 0};
  float* F[3] = {};   // raw face-slab force (written by the interface kernel and sources)
  float* S[3] = {};   // after the modal pass, the mode planes' face-pair sums
  float* U[3] = {};   // GEMM intermediate
  float* G[3] = {};   // folded force spectra
  float* P[3] = {};   // face-slab pressure
  float* T[3] = {};   // partial reductions [nchunk][6][plane]
  float* M[2] = {};   // modes [D][H][wp], rows padded to whole 128-byte lines
  int wp = 0;
  float* w2[3] = {};     // per-axis squared wavenumbers (c0 pi (k+1) / L)^2
  float gscale = 0.f;
  float* E[3] = {};   // Ex, Ey, Ez
  float* R[3] = {};   // Rx, Ry, Rz
  float* yz[2] = {};  // y, z: [n][8] = {E0, E1, E2, R0, R1, R2, u, 0}
  float* rec_part = nullptr;
  // recorders past the modal pass's ARD_MAX_REC_PER_PART: block partials of
  // k_xrec, [xnb][recs.size() - ARD_MAX_REC_PER_PART]
  float* xrec_part = nullptr;
  int xnb = 0;
  int rpw = 4, nbx = 0, nby = 0, nbz = 0, kz_chunk = 0;
  int nrec = 0;
  std::vector<int> recs;  // recorder ids owned
};

struct PointSrc {
  int part;
  int x, y, z;      // local cell
  int kind;         // 0 Gaussian, 1 waveform, 2 constant
  const float* wave = nullptr;
  int n = 0;
  float value = 0.f;
  float* slot = nullptr;   // value slot (interior) or face-slab cell
  bool in_shell = false;
  float* tab = nullptr;    // fx, fy, fz factor tables (interior points)
};

struct BatchedState {
  std::vector<BPart> parts;  // by partition id (DCT entries used)
  std::vector<PointSrc> srcs;
  struct ModalGroup {
    int rpw = 4;
    int ring = 0;  // planes of mode rows in flight per wave through LDS (0: one, in registers)
    bool pts = true;
    ModalDesc* d_desc[2] = {};  // same table for both parities (the kernel takes the parity)
    long long* d_start = nullptr;
    int n = 0;
    long long blocks = 0;
  };
  // Launch tables of the DCT partitions' phases: the four face passes
  // in every transform kernel, the modal pass and the unfold.
  struct Stage {
    UnfoldDesc* unfold = nullptr; long long* unfold_start = nullptr; int nunfold = 0; long long unfold_total = 0;
    GemmDesc* gemm[4] = {};        // FT1, FT2, IT1, IT2 (dense kernels, variants 0 and 1)
    int ngemm[4] = {}, gemm_tiles[4] = {};
    GemmDesc* gemm_fold[4] = {};   // the same passes even/odd folded (transform variant 3)
    int2* fold_tiles[4] = {};      // their tile tables (fold_tile_table)
    int gemm_fold_tiles[4] = {};
    GemmDesc* gemm_bf16[4] = {};   // the folded passes in BF16 split precision (variant 2), fold_tiles' tiles
    GemmDesc* gemm_direct[4] = {}; // the folded passes, operands loaded straight into registers (variant 5)
    int2* direct_tiles[4] = {};
    int gemm_direct_tiles[4] = {};
    GemmDesc* gemm_fft[4] = {};    // the same passes FFT-style (variant 4), when every length factors
    int gemm_fft_tiles[4] = {};
    ModalGroup modal;
    // point sources and recorders past the modal pass's four per partition
    XsrcDesc* xsrc = nullptr; long long* xsrc_start = nullptr; int nxsrc = 0; long long xsrc_blocks = 0;
    XrecDesc* xrec = nullptr; long long* xrec_start = nullptr; int nxrec = 0; long long xrec_blocks = 0;
  };
  Stage stage;
  // host copies of the per-partition descriptors in descriptor order (three
  // per DCT partition for unfold and each face pass; one modal)
  std::vector<int> dct_ids;
  std::vector<UnfoldDesc> h_unfold;
  std::vector<GemmDesc> h_gemm[4], h_gemm_fold[4], h_gemm_bf16[4], h_gemm_fft[4], h_gemm_direct[4];
  std::vector<ModalDesc> h_modal;
  size_t fft_lds[4] = {};
  bool has_fft = false;
  // the PML slabs in two launch groups (ARD_OPT_PML_SPLIT): group 0 beside
  // the forward face transforms, group 1 beside the unfold and the inverse
  // ones; npml counts both
  PmlDesc* d_pml[2] = {}; long long pml_blocks[2] = {}; int pml_n[2] = {}; int npml = 0;
  int pml_shape = 0;   // tile shape of the tiled PML update (k_pml_tiled), chosen at finalize
  int pml_split = 0;   // percent of the PML's tile blocks in group 0
  BndDesc* d_bnd = nullptr; long long bnd_total = 0; int nbnd = 0;
  PartAccess* d_parts = nullptr;
  BSourceDesc* d_src = nullptr; int nsrc = 0;
  BRecDesc* d_rec = nullptr;
  float* d_slots = nullptr;  // value slots of interior point sources
  int nslots = 0;
  // dense pressure evaluation on demand
  float* dense[3] = {};
  size_t dense_n = 0;
  GemmDesc* d_dense = nullptr;
  const float* graph_trace = nullptr;
  bool graphs_valid = false;
  // tiling of the modal pass and transform kernel, chosen at finalize
  int rpw = 4, kz_chunk = 48, gemm_variant = 1, tuned = 0;
  int ring = 0;  // modal pass: planes of mode rows in flight per wave through LDS (4 rows per wave)
  // second stream and events: in the step graph the PML slabs run beside the
  // forward face transforms, which they share no data with
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_fork2 = nullptr, ev_join = nullptr;
  // kMultiSteps steps captured in one graph, per starting parity: consecutive
  // graph launches leave the GPU idle a few microseconds between them
  static constexpr int kMultiSteps = 8;
  hipGraphExec_t graph_multi[2] = {nullptr, nullptr};
  ~BatchedState() {
    for (auto& g : graph_multi)
      if (g) hipGraphExecDestroy(g);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_fork2) hipEventDestroy(ev_fork2);
    if (ev_join) hipEventDestroy(ev_join);
    if (side) hipStreamDestroy(side);
  }
};

namespace {

int upload_vec(Ctx& c, void** dptr, const void* src, size_t bytes) {
  int r = c.dalloc(dptr, bytes);
  if (r) return r;
  return bytes ? c.upload(*dptr, src, bytes) : ARD_OK;
}

// 1D factor tables along one axis of length n for points at x0..x2 = 0,1,2:
// E[k][x] = 2 c_n(k, x) (force expansion, REDFT10 weight), R[k][x] = e(k) c_n(k, x)
// (pressure reduction, REDFT01 weight).
void axis_tables(int n, std::vector<float>& E, std::vector<float>& R) {
  E.resize(size_t(n) * 3);
  R.resize(size_t(n) * 3);
  for (int k = 0; k < n; ++k)
    for (int x = 0; x < 3; ++x) {
      const double cs = std::cos(kPi * k * (x + 0.5) / n);
      E[size_t(k) * 3 + x] = float(2.0 * cs);
      R[size_t(k) * 3 + x] = float((k == 0 ? 1.0 : 2.0) * cs);
    }
}

void point_table(int n, int x, bool forward, float* out) {
  for (int k = 0; k < n; ++k) {
    const double cs = std::cos(kPi * k * (x + 0.5) / n);
    out[k] = float(forward ? 2.0 * cs : (k == 0 ? 1.0 : 2.0) * cs);
  }
}

GemmDesc gemm_k_contig(const float* A, int L, const float* B, float* C, int rows, long long ld) {
  GemmDesc g{};
  g.b_elems = unsigned(rows * ld);
  g.A = A; g.B = B; g.C = C;
  g.M = L; g.K = L; g.N = rows; g.N1 = rows;
  g.bK = 1; g.b1 = ld; g.b2 = 0;
  g.cM = 1; g.c1 = ld; g.c2 = 0;
  g.alpha = 1.f;
  g.tiles_n = int(ceil_div(g.N, 64));
  return g;
}

// transform along the middle axis of a [6][L][inner] stack
GemmDesc gemm_n_contig(const float* A, int L, const float* B, float* C, int inner) {
  GemmDesc g{};
  g.b_elems = unsigned(6 * L * inner);
  g.A = A; g.B = B; g.C = C;
  g.M = L; g.K = L; g.N = 6 * inner; g.N1 = inner;
  g.bK = inner; g.b1 = 1; g.b2 = (long long)L * inner;
  g.cM = inner; g.c1 = 1; g.c2 = (long long)L * inner;
  g.alpha = 1.f;
  g.tiles_n = int(ceil_div(g.N, 64));
  return g;
}

// Even/odd folded cosine matrices of an axis of length L (k_gemm_mfma_fold),
// zero-padded so that the kernel's operand tiles (64 rows forward, 32 rows of
// each matrix inverse, 16 columns) never leave them: columns to
// Kp = 16 ceil(ceil(L/2) / 16);
// forward [fold_rows_fwd(L)][Kp]: rows m < ceil(L/2) = REDFT10 row 2m, zero
// rows up to Mh = fold_even_rows(L) (a multiple of 32: each wave's 32 rows of a
// tile are all even or all odd), then the odd rows 2m+1, zero rows to a
// multiple of 64; inverse [2][fold_rows_inv(L)][Kp]: REDFT01 columns 2k' (E)
// and 2k'+1 (O) of rows j, zero rows to a multiple of 32. For odd L the forward operand's middle element is
// x[j] + x[L-1-j] = 2 x[j], so the even rows' middle column is halved.
// FP64, stored FP32; cached per length in the context.
int fold_kp(int L) { return int(ceil_div((L + 1) / 2, 16) * 16); }
int fold_even_rows(int L) { return int(ceil_div((L + 1) / 2, 32) * 32); }
int fold_rows_fwd(int L) { return int(ceil_div(fold_even_rows(L) + L / 2, 64) * 64); }
int fold_rows_inv(int L) { return int(ceil_div((L + 1) / 2, 32) * 32); }

void fold_table_host(int L, std::vector<float>& f, std::vector<float>& v) {
  const int Kh = (L + 1) / 2, Me = (L + 1) / 2, Mo = L / 2, Kp = fold_kp(L);
  const int Mh = fold_even_rows(L), Rf = fold_rows_fwd(L), Ri = fold_rows_inv(L);
  f.assign(size_t(Rf) * Kp, 0.f);
  v.assign(size_t(2) * Ri * Kp, 0.f);
  {
    auto c2 = [&](int k, int j) { return 2.0 * std::cos(kPi * k * (j + 0.5) / L); };
    auto c3 = [&](int j, int k) { return k == 0 ? 1.0 : 2.0 * std::cos(kPi * k * (j + 0.5) / L); };
    for (int j = 0; j < Kh; ++j) {
      const double mid = (L % 2 == 1 && j == L / 2) ? 0.5 : 1.0;
      for (int m = 0; m < Me; ++m) f[size_t(m) * Kp + j] = float(mid * c2(2 * m, j));
      for (int m = 0; m < Mo; ++m) f[size_t(Mh + m) * Kp + j] = float(c2(2 * m + 1, j));
      for (int k = 0; k < Kh; ++k) {
        if (2 * k < L) v[size_t(j) * Kp + k] = float(c3(j, 2 * k));
        if (2 * k + 1 < L) v[size_t(Ri + j) * Kp + k] = float(c3(j, 2 * k + 1));
      }
    }
  }
}

// BF16 round to nearest even of a finite float, as the device's bf16_rn
uint16_t bf16_host(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// The folded tables' BF16 high and low parts (x = hi + lo), for
// k_gemm_bf16x3_fold: out = {forward hi, forward lo, inverse hi, inverse lo}.
int fold_tables_bf16(Ctx& c, int L, const void* out[4]) {
  auto it = c.fold_bf16.find(L);
  if (it == c.fold_bf16.end()) {
    std::vector<float> f, v;
    fold_table_host(L, f, v);
    std::array<void*, 4> d{};
    int r;
    int q = 0;
    for (const std::vector<float>* t : {&f, &v}) {
      std::vector<uint16_t> hi(t->size()), lo(t->size());
      for (size_t i = 0; i < t->size(); ++i) {
        hi[i] = bf16_host((*t)[i]);
        const uint32_t hb = uint32_t(hi[i]) << 16;
        float hf;
        std::memcpy(&hf, &hb, 4);
        lo[i] = bf16_host((*t)[i] - hf);
      }
      if ((r = upload_vec(c, &d[q], hi.data(), hi.size() * 2)) || (r = upload_vec(c, &d[q + 1], lo.data(), lo.size() * 2)))
        return r;
      q += 2;
    }
    it = c.fold_bf16.emplace(L, d).first;
  }
  for (int i = 0; i < 4; ++i) out[i] = it->second[i];
  return ARD_OK;
}

int fold_tables(Ctx& c, int L, const float** fwd, const float** inv) {
  auto it = c.fold_tables.find(L);
  if (it == c.fold_tables.end()) {
    std::vector<float> f, v;
    fold_table_host(L, f, v);
    float *df, *di;
    int r;
    if ((r = upload_vec(c, (void**)&df, f.data(), f.size() * 4)) || (r = upload_vec(c, (void**)&di, v.data(), v.size() * 4)))
      return r;
    it = c.fold_tables.emplace(L, std::make_pair(df, di)).first;
  }
  *fwd = it->second.first;
  *inv = it->second.second;
  return ARD_OK;
}

// Tile table of a folded transform launch: every (descriptor, tile) pair,
// tiles of tile_m rows (64 for k_gemm_mfma_fold, 32 pair rows for
// k_fold_direct) by the descriptor's tiles_n column tiles.
// A panel (one n tile of a descriptor, all its m tiles) shares its B operand,
// so a panel goes to one XCD (block b runs on XCD b % 8); panels are dealt
// longest K first to the XCD with the least work so far, and each XCD's
// blocks run longest first.
int fold_tile_table(Ctx& c, const std::vector<GemmDesc>& v, int2** d, int tile_m) {
  struct Panel { int desc, nt, tm; };
  std::vector<Panel> panels;
  for (int i = 0; i < int(v.size()); ++i) {
    const int tm = int(ceil_div(v[i].M, tile_m));
    for (int nt = 0; nt < v[i].tiles_n; ++nt) panels.push_back(Panel{i, nt, tm});
  }
  std::stable_sort(panels.begin(), panels.end(), [&](const Panel& a, const Panel& b) { return v[a.desc].K > v[b.desc].K; });
  constexpr int kXcd = 8;
  std::vector<int2> lists[kXcd];
  long long work[kXcd] = {};
  for (const Panel& p : panels) {
    const int x = int(std::min_element(work, work + kXcd) - work);
    for (int m = 0; m < p.tm; ++m) lists[x].push_back(int2{p.desc, p.nt * p.tm + m});
    work[x] += (long long)p.tm * v[p.desc].K;
  }
  std::vector<int2> t;
  size_t pos[kXcd] = {};
  size_t left = 0;
  for (auto& l : lists) left += l.size();
  while (left) {  // block b from XCD b % 8's list; an exhausted list yields to the longest remaining one
    int x = int(t.size() % kXcd);
    if (pos[x] == lists[x].size()) {
      size_t best = 0;
      for (int y = 0; y < kXcd; ++y)
        if (lists[y].size() - pos[y] > best) best = lists[y].size() - pos[y], x = y;
    }
    t.push_back(lists[x][pos[x]++]);
    --left;
  }
  return upload_vec(c, (void**)d, t.data(), t.size() * sizeof(int2));
}

// One face-transform launch k (0, 1 forward; 2, 3 inverse) of a group of
// partitions with transform kernel v: 3 the even/odd folded matrix-core
// passes (5 the same with operands loaded straight into registers), 4 the
// FFT-style passes, else the dense passes of launch_gemm (0, 1, 2).
int launch_face(const BatchedState& B, const BatchedState::Stage& S, int k, int v, hipStream_t st) {
  if (v == 3)
    return launch_gemm_fold(S.gemm_fold[k], S.fold_tiles[k], S.gemm_fold_tiles[k], k % 2 == 0, k < 2 ? 1 : 2, st);
  if (v == 5)
    return launch_fold_direct(S.gemm_direct[k], S.direct_tiles[k], S.gemm_direct_tiles[k], k % 2 == 0, k < 2 ? 1 : 2,
                              st);
  if (v == 4)
    return launch_fft(S.gemm_fft[k], S.ngemm[k], S.gemm_fft_tiles[k], k % 2 == 0, k % 2 == 0, 0, B.fft_lds[k], st);
  if (v == 2)
    return launch_gemm_bf16_fold(S.gemm_bf16[k], S.fold_tiles[k], S.gemm_fold_tiles[k], k % 2 == 0, k < 2 ? 1 : 2,
                                 st);
  return launch_gemm(S.gemm[k], S.ngemm[k], S.gemm_tiles[k], k % 2 == 0, k % 2 == 0, 0, st, v);
}

int upload_gemms(Ctx& c, std::vector<GemmDesc>& v, GemmDesc** d, int* tiles) {
  int t = 0;
  for (auto& g : v) {
    g.tile0 = t;
    t += int(ceil_div(g.M, 64)) * g.tiles_n;
  }
  *tiles = t;
  return upload_vec(c, (void**)d, v.data(), v.size() * sizeof(GemmDesc));
}

// Uploads the stage's launch tables that depend on the geometry only: the
// four face passes in every transform kernel.
int build_stage(Ctx& c) {
  BatchedState& B = *c.bs;
  BatchedState::Stage& S = B.stage;
  int r;
  if (B.dct_ids.empty()) return ARD_OK;  // nothing to launch
  std::vector<long long> st;
  long long t = 0;
  for (int k = 0; k < 4; ++k) {
    std::vector<GemmDesc> g = B.h_gemm[k];
    S.ngemm[k] = int(g.size());
    if ((r = upload_gemms(c, g, &S.gemm[k], &S.gemm_tiles[k]))) return r;
    std::vector<GemmDesc> gf = B.h_gemm_fold[k];
    if ((r = upload_gemms(c, gf, &S.gemm_fold[k], &S.gemm_fold_tiles[k]))) return r;
    if ((r = fold_tile_table(c, gf, &S.fold_tiles[k], 64))) return r;
    {
      const std::vector<GemmDesc>& gd = B.h_gemm_direct[k];
      int t = 0;
      for (const GemmDesc& x : gd) t += int(ceil_div(x.M, 32)) * x.tiles_n;
      S.gemm_direct_tiles[k] = t;
      if ((r = upload_vec(c, (void**)&S.gemm_direct[k], gd.data(), gd.size() * sizeof(GemmDesc))) ||
          (r = fold_tile_table(c, gd, &S.direct_tiles[k], 32)))
        return r;
    }
    const std::vector<GemmDesc>& gb = B.h_gemm_bf16[k];
    if ((r = upload_vec(c, (void**)&S.gemm_bf16[k], gb.data(), gb.size() * sizeof(GemmDesc)))) return r;
    if (B.has_fft) {
      std::vector<GemmDesc> ff = B.h_gemm_fft[k];
      int tf = 0;
      for (GemmDesc& x : ff) {
        x.tile0 = tf;
        tf += fft_tiles(x);
      }
      S.gemm_fft_tiles[k] = tf;
      if ((r = upload_vec(c, (void**)&S.gemm_fft[k], ff.data(), ff.size() * sizeof(GemmDesc)))) return r;
    }
  }
  return ARD_OK;
}

// Uploads the stage's tables that depend on the modal tiling and the point
// forces: the modal pass and the unfold.
int build_stage_tiling(Ctx& c) {
  BatchedState& B = *c.bs;
  BatchedState::Stage& S = B.stage;
  int r;
  if (B.dct_ids.empty()) return ARD_OK;
  std::vector<long long> su;
  long long t = 0;
  for (auto& x : B.h_unfold) { su.push_back(t); t += ceil_div(3 * x.plane / x.vec, 256) * 256; }
  S.nunfold = int(B.h_unfold.size());
  S.unfold_total = t;
  if ((r = upload_vec(c, (void**)&S.unfold, B.h_unfold.data(), B.h_unfold.size() * sizeof(UnfoldDesc)))) return r;
  if ((r = upload_vec(c, (void**)&S.unfold_start, su.data(), su.size() * 8))) return r;
  // one launch for all partitions: the point terms are uniform branches, and
  // one grid avoids the tail of a second, smaller launch
  BatchedState::ModalGroup g;
  g.rpw = B.rpw;
  g.ring = g.rpw == 4 ? B.ring : 0;
  g.pts = true;
  std::vector<long long> starts;
  long long blocks = 0;
  for (const ModalDesc& d : B.h_modal) {
    starts.push_back(blocks);
    blocks += (long long)d.nbx * d.nby * d.nbz;
  }
  g.n = int(B.h_modal.size());
  g.blocks = blocks;
  if ((r = upload_vec(c, (void**)&g.d_desc[0], B.h_modal.data(), B.h_modal.size() * sizeof(ModalDesc)))) return r;
  g.d_desc[1] = g.d_desc[0];
  if ((r = upload_vec(c, (void**)&g.d_start, starts.data(), starts.size() * 8))) return r;
  S.modal = g;
  return ARD_OK;
}

// The grouped PML launch's descriptor table for tile shape `shape` (the
// blocks of each slab depend on it): each slab's first block and the total.
int build_pml_table(Ctx& c, int shape, int split) {
  BatchedState& B = *c.bs;
  const int L = c.prm.pml_layers;
  std::vector<PmlDesc> pd = c.h_pml;
  std::vector<long long> nb(pd.size());
  long long total = 0;
  for (size_t i = 0; i < pd.size(); ++i) total += nb[i] = pml_slab_blocks(pd[i], L, shape);
  // group 0 takes the leading slabs until it holds split % of the blocks
  std::vector<PmlDesc> g[2];
  long long acc = 0;
  for (size_t i = 0; i < pd.size(); ++i) {
    g[acc * 100 < split * total ? 0 : 1].push_back(pd[i]);
    acc += nb[i];
  }
  for (int k = 0; k < 2; ++k) {
    long long b0 = 0;
    for (PmlDesc& d : g[k]) {
      d.block0 = b0;
      b0 += pml_slab_blocks(d, L, shape);
    }
    B.pml_blocks[k] = b0;
    B.pml_n[k] = int(g[k].size());
    if (!g[k].empty()) {
      const int r = c.upload(B.d_pml[k], g[k].data(), g[k].size() * sizeof(PmlDesc));
      if (r) return r;
    }
  }
  B.pml_shape = shape;
  B.pml_split = split;
  B.graphs_valid = false;
  return ARD_OK;
}

namespace {
int launch_pml_group(Ctx& c, const BatchedState& B, int k, int par, hipStream_t s) {
  return launch_pml(B.d_pml[k], B.pml_n[k], B.pml_blocks[k], par, c.prm.pml_layers, c.prm.h, c.prm.dt, c.prm.c0, s,
                    B.pml_shape);
}
}  // namespace

// The source samples of the current step (*d_step) into their cells.
int write_sources(Ctx& c) {
  BatchedState& B = *c.bs;
  if (launch_bsources(B.d_src, B.nsrc, c.d_step, c.courant, c.stream)) return set_error(ARD_ERR_HIP, "sources");
  hipError_t e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int build_sources(Ctx& c) {
  BatchedState& B = *c.bs;
  int r;
  std::vector<BSourceDesc> sd;
  B.nslots = 0;
  for (auto& s : B.srcs) {
    const PartGeom& p = c.topo.parts[s.part];
    long long off;
    if (shell_of(p.w, p.h, p.d, s.x, s.y, s.z, &off) < 0) B.nslots++;
  }
  if (B.nslots) {
    if ((r = c.dalloc((void**)&B.d_slots, B.nslots * 4))) return r;
    if ((r = c.dzero(B.d_slots, B.nslots * 4))) return r;
  }
  int slot = 0;
  for (auto& s : B.srcs) {
    const PartGeom& p = c.topo.parts[s.part];
    BPart& P = B.parts[s.part];
    long long off;
    const int w = shell_of(p.w, p.h, p.d, s.x, s.y, s.z, &off);
    if (w >= 0) {
      s.in_shell = true;
      s.slot = P.F[w] + off;
    } else {
      s.in_shell = false;
      s.slot = B.d_slots + slot++;
      if (!s.tab) {
        std::vector<float> t(size_t(p.w) + p.h + p.d);
        point_table(p.w, s.x, true, t.data());
        point_table(p.h, s.y, true, t.data() + p.w);
        point_table(p.d, s.z, true, t.data() + p.w + p.h);
        if ((r = upload_vec(c, (void**)&s.tab, t.data(), t.size() * 4))) return r;
      }
    }
    BSourceDesc d{};
    d.kind = s.kind; d.wave = s.wave; d.n = s.n; d.value = s.value; d.dst = s.slot;
    sd.push_back(d);
  }
  B.nsrc = int(sd.size());
  if (B.nsrc && (r = upload_vec(c, (void**)&B.d_src, sd.data(), sd.size() * sizeof(BSourceDesc)))) return r;
  // the current step's samples (each step's recorder kernel writes the next one's)
  return write_sources(c);
}

int build_modal(Ctx& c) {
  BatchedState& B = *c.bs;
  int r;
  B.h_modal.clear();
  std::vector<XsrcDesc> xs;
  std::vector<XrecDesc> xr;
  for (int i = 0; i < int(c.topo.parts.size()); ++i) {
    const PartGeom& p = c.topo.parts[i];
    if (p.is_pml) continue;
    BPart& P = B.parts[i];
    ModalDesc d{};
    d.W = p.w; d.H = p.h; d.D = p.d;
    d.nbx = P.nbx; d.nby = P.nby; d.nbz = P.nbz; d.kz_chunk = P.kz_chunk;
    d.dt = c.prm.dt;
    d.gscale = P.gscale;
    d.wx2 = P.w2[0]; d.wy2 = P.w2[1]; d.wz2 = P.w2[2];
    d.M[0] = P.M[0]; d.M[1] = P.M[1];
    d.Wp = P.wp;
    d.Ex = P.E[0]; d.Rx = P.R[0]; d.Ey = P.E[1]; d.Ry = P.R[1]; d.Ez = P.E[2]; d.Rz = P.R[2];
    d.Gx = P.G[0]; d.Gy = P.G[1]; d.Gz = P.G[2];
    d.yt = P.yz[0]; d.zt = P.yz[1];
    d.Tx = P.T[0]; d.Ty = P.T[1]; d.Tz = P.T[2];
    std::vector<SrcTab> st;
    for (auto& s : B.srcs)
      if (s.part == i && !s.in_shell) st.push_back(SrcTab{s.tab, s.tab + p.w, s.tab + p.w + p.h, s.slot});
    if (int(st.size()) > ARD_MAX_SRC_PER_PART) {  // the rest: k_xsrc
      XsrcDesc x{};
      x.M[0] = P.M[0]; x.M[1] = P.M[1];
      x.W = p.w; x.H = p.h; x.D = p.d; x.Wp = P.wp;
      x.wx2 = P.w2[0]; x.wy2 = P.w2[1]; x.wz2 = P.w2[2];
      x.gscale = P.gscale;
      x.n = int(st.size()) - ARD_MAX_SRC_PER_PART;
      SrcTab* ds;
      if ((r = upload_vec(c, (void**)&ds, st.data() + ARD_MAX_SRC_PER_PART, size_t(x.n) * sizeof(SrcTab)))) return r;
      x.src = ds;
      xs.push_back(x);
      st.resize(ARD_MAX_SRC_PER_PART);
    }
    d.nsrc = int(st.size());
    if (d.nsrc) {
      SrcTab* ds;
      if ((r = upload_vec(c, (void**)&ds, st.data(), st.size() * sizeof(SrcTab)))) return r;
      d.src = ds;
    }
    std::vector<RecTab> rt;
    for (int rid : P.recs) {
      const RecorderDef& rd = c.topo.recorders[rid];
      std::vector<float> t(size_t(p.w) + p.h + p.d);
      point_table(p.w, rd.lx, false, t.data());
      point_table(p.h, rd.ly, false, t.data() + p.w);
      point_table(p.d, rd.lz, false, t.data() + p.w + p.h);
      float* dt;
      if ((r = upload_vec(c, (void**)&dt, t.data(), t.size() * 4))) return r;
      rt.push_back(RecTab{dt, dt + p.w, dt + p.w + p.h});
    }
    if (int(rt.size()) > ARD_MAX_REC_PER_PART) {  // the rest: k_xrec
      XrecDesc x{};
      x.M[0] = P.M[0]; x.M[1] = P.M[1];
      x.W = p.w; x.H = p.h; x.D = p.d; x.Wp = P.wp;
      x.n = int(rt.size()) - ARD_MAX_REC_PER_PART;
      RecTab* dr;
      if ((r = upload_vec(c, (void**)&dr, rt.data() + ARD_MAX_REC_PER_PART, size_t(x.n) * sizeof(RecTab)))) return r;
      x.rec = dr;
      x.part = P.xrec_part;
      xr.push_back(x);
      rt.resize(ARD_MAX_REC_PER_PART);
    }
    d.nrec = int(rt.size());
    if (d.nrec) {
      RecTab* dr;
      if ((r = upload_vec(c, (void**)&dr, rt.data(), rt.size() * sizeof(RecTab)))) return r;
      d.rec = dr;
    }
    d.rec_part = P.rec_part;
    {
      // y and z factors of the points, 8 per wavenumber (sources, then
      // recorders); all zero for a partition without points
      std::vector<float> ty(size_t(p.h) * 8, 0.f), tz(size_t(p.d) * 8, 0.f), col(size_t(std::max(p.h, p.d)));
      int q = 0;
      for (auto& s : B.srcs) {
        if (s.part != i || s.in_shell || q == ARD_MAX_SRC_PER_PART) continue;
        point_table(p.h, s.y, true, col.data());
        for (int k = 0; k < p.h; ++k) ty[size_t(k) * 8 + q] = col[size_t(k)];
        point_table(p.d, s.z, true, col.data());
        for (int k = 0; k < p.d; ++k) tz[size_t(k) * 8 + q] = col[size_t(k)];
        ++q;
      }
      q = 0;
      for (int rid : P.recs) {
        if (q == ARD_MAX_REC_PER_PART) break;
        const RecorderDef& rd = c.topo.recorders[rid];
        point_table(p.h, rd.ly, false, col.data());
        for (int k = 0; k < p.h; ++k) ty[size_t(k) * 8 + 4 + q] = col[size_t(k)];
        point_table(p.d, rd.lz, false, col.data());
        for (int k = 0; k < p.d; ++k) tz[size_t(k) * 8 + 4 + q] = col[size_t(k)];
        ++q;
      }
      float *dy, *dz;
      if ((r = upload_vec(c, (void**)&dy, ty.data(), ty.size() * 4))) return r;
      if ((r = upload_vec(c, (void**)&dz, tz.data(), tz.size() * 4))) return r;
      d.pty = dy;
      d.ptz = dz;
    }
    B.h_modal.push_back(d);
  }
  // the extra point passes: one launch each over every partition that has them
  BatchedState::Stage& S = B.stage;
  S.nxsrc = int(xs.size());
  S.nxrec = int(xr.size());
  S.xsrc_blocks = S.xrec_blocks = 0;
  if (S.nxsrc) {
    std::vector<long long> st;
    for (const XsrcDesc& x : xs) { st.push_back(S.xsrc_blocks); S.xsrc_blocks += ceil_div((long long)x.D * x.H, kXRows); }
    if ((r = upload_vec(c, (void**)&S.xsrc, xs.data(), xs.size() * sizeof(XsrcDesc))) ||
        (r = upload_vec(c, (void**)&S.xsrc_start, st.data(), st.size() * 8)))
      return r;
  }
  if (S.nxrec) {
    std::vector<long long> st;
    for (const XrecDesc& x : xr) { st.push_back(S.xrec_blocks); S.xrec_blocks += ceil_div((long long)x.D * x.H, kXRows); }
    if ((r = upload_vec(c, (void**)&S.xrec, xr.data(), xr.size() * sizeof(XrecDesc))) ||
        (r = upload_vec(c, (void**)&S.xrec_start, st.data(), st.size() * 8)))
      return r;
  }
  return build_stage_tiling(c);
}

// Last-writer masks of the grouped interface launch (final_writer_masks),
// uploaded with each boundary's element offset (a multiple of the 256-thread
// block, which k_interface's per-block descriptor lookup relies on).
int build_masks(Ctx& c, std::vector<BndDesc>& descs, const std::vector<long long>& elems) {
  BatchedState& B = *c.bs;
  std::vector<uint8_t> mask;
  std::vector<long long> mstart;
  final_writer_masks(c.topo, descs, elems, mask, mstart);
  uint8_t* dm = nullptr;
  int r = upload_vec(c, (void**)&dm, mask.data(), mask.size());
  if (r) return r;
  long long e0 = 0;
  for (int b = 0; b < int(descs.size()); ++b) {
    descs[b].mask = dm + mstart[b];
    descs[b].elem0 = e0;
    e0 += ceil_div(elems[b], 256) * 256;
  }
  B.bnd_total = e0;
  return ARD_OK;
}

}  // namespace

// The even/odd folded twin of a dense per-axis pass of length L = g.M (the
// descriptor k_gemm_mfma_fold takes): the folded tables of L, K halved,
// padded extents, rows as the batched form's face passes lay them out.
int fold_twin(Ctx& c, GemmDesc& g, bool inverse) {
  const int L = g.M;
  const float *ff, *fi;
  const int r = fold_tables(c, L, &ff, &fi);
  if (r) return r;
  g.L = L;
  g.K = (L + 1) / 2;
  g.Kp = fold_kp(L);
  g.Ap = inverse ? fold_rows_inv(L) : fold_rows_fwd(L);
  if (!inverse) {
    g.A = ff;
    g.Mh = fold_even_rows(L);
    g.M = g.Mh + L / 2;
  } else {
    g.A = fi;
    g.M = int(64 * ceil_div((L + 1) / 2, 32));
  }
  g.tiles_n = int(ceil_div(g.N, 64));
  return ARD_OK;
}

// The tile table of a launch of folded passes (64-row tiles).
int fold_pass_tiles(Ctx& c, const std::vector<GemmDesc>& v, int2** d) { return fold_tile_table(c, v, d, 64); }


// Sets the modal-pass tiling of every DCT partition (rows per wave, kz planes
// per block) and rebuilds the descriptors that depend on it.
int apply_tiling(Ctx& c, int rpw, int kzc) {
  BatchedState& B = *c.bs;
  int r;
  B.rpw = rpw;
  B.kz_chunk = kzc;
  std::vector<UnfoldDesc>& unfolds = B.h_unfold;
  unfolds.clear();
  for (int i : B.dct_ids) {
    const PartGeom& p = c.topo.parts[size_t(i)];
    BPart& P = B.parts[size_t(i)];
    P.rpw = rpw;
    P.nbx = int(ceil_div(p.w, 64));
    P.nby = int(ceil_div(p.h, 4 * rpw));
    P.nbz = int(ceil_div(p.d, kzc));
    P.kz_chunk = int(ceil_div(ceil_div(p.d, P.nbz), 2) * 2);  // even: each block starts on an even plane
    P.nbz = int(ceil_div(p.d, P.kz_chunk));
    for (int w = 0; w < 3; ++w)
      unfolds.push_back(UnfoldDesc{P.T[w], P.S[w], P.plane[w], w == 0 ? P.nbx : (w == 1 ? P.nby : P.nbz),
                                   vec_of(P.plane[w], {P.T[w], P.S[w]})});
  }
  if ((r = build_modal(c))) return r;  // and the stage's modal and unfold tables
  std::vector<BRecDesc> rd;
  for (size_t q = 0; q < c.topo.recorders.size(); ++q) {
    const BPart& P = B.parts[c.topo.recorders[q].part];
    const int slot = int(std::find(P.recs.begin(), P.recs.end(), int(q)) - P.recs.begin());
    if (slot < ARD_MAX_REC_PER_PART)
      rd.push_back(BRecDesc{P.rec_part + slot, P.nbx * P.nby * P.nbz, ARD_MAX_REC_PER_PART});
    else  // k_xrec's partials
      rd.push_back(BRecDesc{P.xrec_part + (slot - ARD_MAX_REC_PER_PART), P.xnb,
                            int(P.recs.size()) - ARD_MAX_REC_PER_PART});
  }
  if (!rd.empty() && (r = upload_vec(c, (void**)&B.d_rec, rd.data(), rd.size() * sizeof(BRecDesc)))) return r;
  B.graphs_valid = false;
  hipError_t e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

int finalize_batched(Ctx& c) {
  int r;
  if ((r = setup_common(c))) return r;
  c.bs = new BatchedState();
  BatchedState& B = *c.bs;
  B.gemm_variant = g_gemm_variant;
  {
    hipError_t e = hipStreamCreateWithFlags(&B.side, hipStreamNonBlocking);  // default priority
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B.ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B.ev_fork2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B.ev_join, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "side stream");
  }
  const int np = int(c.topo.parts.size());
  B.parts.assign(np, BPart());
  for (size_t i = 0; i < c.topo.recorders.size(); ++i) B.parts[c.topo.recorders[i].part].recs.push_back(int(i));
  std::vector<GemmDesc>* gm = B.h_gemm;
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = c.topo.parts[i];
    if (p.is_pml) continue;
    B.dct_ids.push_back(i);
    BPart& P = B.parts[i];
    const long long W = p.w, H = p.h, D = p.d, N = W * H * D;
    if (D * H * ceil_div(W, 32) * 32 >= (1LL << 30))  // 32-bit cell and byte offsets in the modal pass
      return set_error(ARD_ERR_UNSUPPORTED, "a DCT partition of 2^30 or more cells (split it)");
    P.plane[0] = D * H; P.plane[1] = D * W; P.plane[2] = H * W;
    // partial-reduction buffers are sized for the finest tiling the
    // autotuner may pick: 2 rows per wave, 16 kz planes per block
    P.rpw = 2;
    P.nbx = int(ceil_div(W, 64));
    P.nby = int(ceil_div(H, 8));
    P.kz_chunk = 16;
    P.nbz = int(ceil_div(D, 16));
    const long long shell = 6 * (P.plane[0] + P.plane[1] + P.plane[2]);
    const long long tsz = 6 * (P.nbx * P.plane[0] + P.nby * P.plane[1] + P.nbz * P.plane[2]);
    P.wp = int(ceil_div(W, 32) * 32);
    const long long NP = D * H * P.wp;  // padded mode array
    const long long total = 5 * shell + tsz + 64 + 2 * NP;
    float* blk;
    if ((r = c.dalloc((void**)&blk, size_t(total) * 4))) return r;
    if ((r = c.dzero(blk, size_t(total) * 4))) return r;
    float* q = blk;
    for (float** arr : {P.F, P.S, P.U, P.G, P.P})
      for (int w = 0; w < 3; ++w) { arr[w] = q; q += 6 * P.plane[w]; }
    P.T[0] = q; q += 6 * P.nbx * P.plane[0];
    P.T[1] = q; q += 6 * P.nby * P.plane[1];
    P.T[2] = q; q += 6 * P.nbz * P.plane[2];
    q = blk + ceil_div(q - blk, 64) * 64;  // 256-byte aligned
    P.M[0] = q; q += NP;
    P.M[1] = q; q += NP;
    // modal coefficients are formed in the kernel from per-axis terms
    // u = (w dt / 2)^2 of the 1-based wavenumbers (ARD/dct_partition.cpp:24-33),
    // tabulated in FP64; the kernel's series for sin^2 holds for u <= 3, past
    // the scheme's stability limit (u = pi^2 / 4 at dt = h / (c0 sqrt 3))
    std::vector<float> uh[3];
    {
      const double pi = 3.14159265358979323846;
      const double hdt = 0.5 * double(c.prm.dt);
      const int n3[3] = {p.w, p.h, p.d};
      double umax = 0.0;
      for (int ax = 0; ax < 3; ++ax) {
        std::vector<float> t(static_cast<size_t>(n3[ax]));
        const double L = n3[ax] * double(c.prm.h);
        for (int k = 0; k < n3[ax]; ++k) {
          const double w = double(c.prm.c0) * pi * (k + 1) / L;
          t[size_t(k)] = float(w * w * hdt * hdt);
        }
        umax += double(t.back());
        if ((r = upload_vec(c, (void**)&P.w2[ax], t.data(), t.size() * 4))) return r;
        uh[ax] = t;
      }
      if (umax > 3.0)
        return set_error(ARD_ERR_UNSUPPORTED,
                         "time step beyond the stability limit dt <= h / (c0 sqrt 3) (batched form; use the plain form)");
      const double Nd = double(N);
      P.gscale = float(double(0.999f) * 4.0 / (2.0 * std::sqrt(2.0 * Nd)) / std::sqrt(2.0 * Nd) * hdt * hdt);
    }
    const int dims[3] = {p.w, p.h, p.d};
    for (int ax = 0; ax < 3; ++ax) {
      std::vector<float> E, R;
      axis_tables(dims[ax], E, R);
      if ((r = upload_vec(c, (void**)&P.E[ax], E.data(), E.size() * 4))) return r;
      if ((r = upload_vec(c, (void**)&P.R[ax], R.data(), R.size() * 4))) return r;
      if (ax == 0) continue;
      // y and z: per wavenumber {E0, E1, E2, R0, R1, R2, u, 0}, one scalar load
      const std::vector<float>& u = uh[ax];
      std::vector<float> t(size_t(dims[ax]) * 8, 0.f);
      for (int k = 0; k < dims[ax]; ++k) {
        for (int j = 0; j < 3; ++j) {
          t[size_t(k) * 8 + j] = E[size_t(k) * 3 + j];
          t[size_t(k) * 8 + 3 + j] = R[size_t(k) * 3 + j];
        }
        t[size_t(k) * 8 + 6] = u[size_t(k)];
      }
      if ((r = upload_vec(c, (void**)&P.yz[ax - 1], t.data(), t.size() * 4))) return r;
    }
    if ((r = c.dalloc((void**)&P.rec_part, size_t(P.nbx) * P.nby * P.nbz * ARD_MAX_REC_PER_PART * 4))) return r;
    if (int(P.recs.size()) > ARD_MAX_REC_PER_PART) {
      P.xnb = int(ceil_div(D * H, kXRows));
      const size_t nx = P.recs.size() - ARD_MAX_REC_PER_PART;
      if ((r = c.dalloc((void**)&P.xrec_part, size_t(P.xnb) * nx * 4))) return r;
    }
    const float *fw, *iw, *fh, *ih, *fd, *id;
    if ((r = c.cos_tables(p.w, &fw, &iw)) || (r = c.cos_tables(p.h, &fh, &ih)) || (r = c.cos_tables(p.d, &fd, &id)))
      return r;
    // x faces [6][D][H]: along y (contiguous) then z; y faces [6][D][W]: x then z; z faces [6][H][W]: x then y.
    // The forward passes transform the six raw force slabs of each face pair
    // (the modal pass folds their spectra: the transforms are linear)
    gm[0].push_back(gemm_k_contig(fh, p.h, P.F[0], P.U[0], int(6 * D), H));
    gm[0].push_back(gemm_k_contig(fw, p.w, P.F[1], P.U[1], int(6 * D), W));
    gm[0].push_back(gemm_k_contig(fw, p.w, P.F[2], P.U[2], int(6 * H), W));
    {
      // x-face spectra are stored [D][H][6] (the six slabs of one
      // (kz, ky) adjacent), so the modal pass reads a row group with one
      // scalar load
      GemmDesc g = gemm_n_contig(fd, p.d, P.U[0], P.G[0], p.h);
      g.cM = 6 * H; g.c1 = 6; g.c2 = 1;
      gm[1].push_back(g);
    }
    gm[1].push_back(gemm_n_contig(fd, p.d, P.U[1], P.G[1], p.w));
    gm[1].push_back(gemm_n_contig(fh, p.h, P.U[2], P.G[2], p.w));
    gm[2].push_back(gemm_k_contig(ih, p.h, P.S[0], P.U[0], int(6 * D), H));
    gm[2].push_back(gemm_k_contig(iw, p.w, P.S[1], P.U[1], int(6 * D), W));
    gm[2].push_back(gemm_k_contig(iw, p.w, P.S[2], P.U[2], int(6 * H), W));
    gm[3].push_back(gemm_n_contig(id, p.d, P.U[0], P.P[0], p.h));
    gm[3].push_back(gemm_n_contig(id, p.d, P.U[1], P.P[1], p.w));
    gm[3].push_back(gemm_n_contig(ih, p.h, P.U[2], P.P[2], p.w));
  }
  for (int k = 0; k < 4; ++k) {
    // the folded twin of each pass: forward passes (k = 0, 1) fold the input,
    // inverse passes (k = 2, 3) produce the output pair
    // (inverse: 64 tile rows, 32 of A_E and 32 of A_O, per 32 outputs j)
    std::vector<GemmDesc> gf = gm[k];
    for (GemmDesc& g : gf)
      if ((r = fold_twin(c, g, k >= 2))) return r;
    B.h_gemm_fold[k] = gf;
    // k_fold_direct's twins: M counts pair rows, tiles of 32 pair rows x 128 columns
    std::vector<GemmDesc> gd = gf;
    for (GemmDesc& g : gd) {
      g.M = (g.L + 1) / 2;
      g.tiles_n = int(ceil_div(g.N, 128));
    }
    B.h_gemm_direct[k] = gd;
    // and their BF16 split twins: the tables' high and low parts
    std::vector<GemmDesc> gb = gf;
    for (GemmDesc& g : gb) {
      const void* t[4];
      if ((r = fold_tables_bf16(c, g.L, t))) return r;
      g.A = static_cast<const float*>(k < 2 ? t[0] : t[2]);
      g.Alo = k < 2 ? t[1] : t[3];
    }
    B.h_gemm_bf16[k] = gb;
  }
  // the FFT-style twins of the dense passes (variant 4): only if every face
  // axis length factors into radices up to 13
  B.has_fft = true;
  for (int k = 0; k < 4 && B.has_fft; ++k) {
    B.h_gemm_fft[k].clear();
    for (GemmDesc g : gm[k]) {
      if (fft_setup(c, g, k >= 2) != ARD_OK) { B.has_fft = false; break; }
      B.fft_lds[k] = std::max(B.fft_lds[k], fft_lds_bytes(g));
      B.h_gemm_fft[k].push_back(g);
    }
  }
  if (B.gemm_variant == 4 && !B.has_fft) B.gemm_variant = 1;
  if ((r = build_stage(c))) return r;
  // PML slabs: one grouped launch
  B.npml = int(c.h_pml.size());
  for (int k = 0; k < 2 && B.npml; ++k)
    if ((r = c.dalloc((void**)&B.d_pml[k], c.h_pml.size() * sizeof(PmlDesc)))) return r;
  if ((r = build_pml_table(c, 0, B.pml_split))) return r;
  // partition access for the interface kernel
  std::vector<PartAccess> pa(np);
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = c.topo.parts[i];
    PartAccess& A = pa[i];
    A.W = p.w; A.H = p.h; A.D = p.d;
    if (p.is_pml) {
      A.kind = 1;
      A.p[0] = c.pml[i].p[0]; A.p[1] = c.pml[i].p[1]; A.f = c.pml[i].f;
    } else {
      A.kind = 2;
      for (int w = 0; w < 3; ++w) { A.ps[w] = B.parts[i].P[w]; A.fs[w] = B.parts[i].F[w]; }
    }
  }
  if ((r = upload_vec(c, (void**)&B.d_parts, pa.data(), pa.size() * sizeof(PartAccess)))) return r;
  std::vector<BndDesc> bd;
  std::vector<long long> be;
  if ((r = build_boundary_descs(c, bd, be))) return r;
  if ((r = build_masks(c, bd, be))) return r;
  B.nbnd = int(bd.size());
  if (B.nbnd && (r = upload_vec(c, (void**)&B.d_bnd, bd.data(), bd.size() * sizeof(BndDesc)))) return r;
  // sources
  for (const SourceDef& s : c.topo.sources) {
    if (s.part < 0) continue;
    const PartGeom& p = c.topo.parts[s.part];
    PointSrc ps;
    ps.part = s.part;
    ps.x = s.x - p.x0; ps.y = s.y - p.y0; ps.z = s.z - p.z0;
    ps.kind = s.wave.empty() ? 0 : 1;
    if (ps.kind == 1) {
      float* w;
      if ((r = upload_vec(c, (void**)&w, s.wave.data(), s.wave.size() * 4))) return r;
      ps.wave = w;
      ps.n = int(s.wave.size());
    }
    B.srcs.push_back(ps);
  }
  if ((r = build_sources(c))) return r;
  if ((r = apply_tiling(c, 4, 48))) return r;
  if (c.prm.autotune && (r = autotune(c))) return r;
  hipError_t e = hipStreamSynchronize(c.stream);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipStreamSynchronize");
}

namespace {

int enqueue_step(Ctx& c, int par, PhaseTimer* T) {
  BatchedState& B = *c.bs;
  hipStream_t st = c.stream;
  int r = 0;
  auto phase = [&](int ph, hipStream_t s, auto&& launch) {
    const int ev = T ? T->begin(s, ph) : -1;
    r |= launch();
    if (T) T->end(s, ev);
  };
  auto ok = [&](hipError_t e) { r |= e != hipSuccess; };
  // the step's source samples are in place: written by the previous step's
  // recorder kernel (or at set-up)
  // captured: PML on the side stream, joined before the interface (which
  // reads the new PML pressure and writes PML forces); timed: all in order on
  // one stream, phase by phase over all partitions.
  const bool fork = !T && B.side;
  const BatchedState::Stage& S = B.stage;
  auto face = [&](int k) {
    static const int ph[4] = {ARD_PH_FACE_FWD1, ARD_PH_FACE_FWD2, ARD_PH_FACE_INV1, ARD_PH_FACE_INV2};
    phase(ph[k], st, [&] { return launch_face(B, S, k, B.gemm_variant, st); });
  };
  // PML group 0 beside the forward transforms (forked at the step's start:
  // it depends on the sources only). The first forward pass is enqueued
  // first, so that its blocks are dispatched before the PML slabs' fill the
  // CUs
  const bool early = fork && B.pml_n[0] > 0;
  if (early) ok(hipEventRecord(B.ev_fork, st));
  face(0);
  if (early) {
    ok(hipStreamWaitEvent(B.side, B.ev_fork, 0));
    r |= launch_pml_group(c, B, 0, par, B.side);
  } else if (!fork) {
    // timed: both groups here, in one phase
    phase(ARD_PH_PML, st, [&] { return launch_pml_group(c, B, 0, par, st) | launch_pml_group(c, B, 1, par, st); });
  }
  face(1);
  {
    // with the extra point passes (partitions of more than four interior
    // point sources or recorders) before and after it
    const BatchedState::ModalGroup& g = S.modal;
    phase(ARD_PH_MODAL, st, [&] {
      int e = launch_xsrc(S.xsrc, S.xsrc_start, S.nxsrc, S.xsrc_blocks, par, st);
      e |= launch_modal(g.rpw, g.pts, g.ring, g.d_desc[par], g.d_start, g.n, g.blocks, par, st);
      e |= launch_xrec(S.xrec, S.xrec_start, S.nxrec, S.xrec_blocks, par, st);
      return e;
    });
  }
  if (fork) {
    // PML group 1 beside the unfold and the inverse transforms (by default
    // the whole PML: on the hall 0.606 ms per step, against 0.632 with the
    // whole PML beside the forward transforms and 0.641-0.645 with it split
    // between the two, 25-75 % beside the forward ones: there it outlasts
    // them and slows the memory-bound modal pass). The fork also keeps the
    // side stream in the capture when group 1 is empty. The unfold is
    // enqueued before the PML's launch (the PML still depends on the modal
    // pass only), so the graph holds its node first and its blocks take the
    // CUs first: 0.620 -> 0.606 ms.
    ok(hipEventRecord(B.ev_fork2, st));
  }
  phase(ARD_PH_UNFOLD, st, [&] { return launch_unfold(S.unfold, S.unfold_start, S.nunfold, S.unfold_total, st); });
  if (fork) {
    ok(hipStreamWaitEvent(B.side, B.ev_fork2, 0));
    if (B.pml_n[1]) r |= launch_pml_group(c, B, 1, par, B.side);
    ok(hipEventRecord(B.ev_join, B.side));
  }
  face(2);
  face(3);
  // the PML stencil overlaps the forward face transforms and is joined
  // here, at its first consumer (long finished by then, so the join costs the
  // step no wait)
  if (fork) ok(hipStreamWaitEvent(st, B.ev_join, 0));
  // the interface reads the PML pressure just written, buffer par ^ 1
  phase(ARD_PH_INTERFACE, st, [&] { return launch_interface(B.d_bnd, B.nbnd, B.d_parts, par ^ 1, B.bnd_total, st); });
  phase(ARD_PH_RECORD, st, [&] {
    if (enqueue_cubes(c, par, st)) return 1;
    return launch_brecord(B.d_rec, int(c.topo.recorders.size()), c.d_trace, c.trace_cap, c.d_step, B.d_src, B.nsrc,
                          c.courant, st);
  });
  return r ? set_error(ARD_ERR_HIP, "batched step launch") : ARD_OK;
}

// Captures nsteps steps from step parity par into an executable graph.
int capture_steps(Ctx& c, int par, int nsteps, hipGraphExec_t* ex) {
  hipGraph_t g;
  hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
  int r = ARD_OK;
  for (int s = 0; s < nsteps && !r; ++s) r = enqueue_step(c, (par + s) & 1, nullptr);
  e = hipStreamEndCapture(c.stream, &g);
  if (r) return r;
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  return e == hipSuccess ? ARD_OK : hip_fail(e, "hipGraphInstantiate");
}

int build_graphs(Ctx& c) {
  BatchedState& B = *c.bs;
  // per parity: one step, and kMultiSteps steps (an even count, so the graph
  // ends at the parity it starts with)
  for (int multi = 0; multi < 2; ++multi) {
    for (int par = 0; par < 2; ++par) {
      hipGraphExec_t& ex = multi ? B.graph_multi[par] : c.graph_exec[par];
      if (ex) {
        hipGraphExecDestroy(ex);
        ex = nullptr;
      }
      int r = capture_steps(c, par, multi ? BatchedState::kMultiSteps : 1, &ex);
      if (r) return r;
    }
  }
  B.graph_trace = c.d_trace;
  B.graphs_valid = true;
  return ARD_OK;
}

}  // namespace


// The autotuner: times the modal pass with the unfold kernel that sums its
// partial reductions (the tiling sets how many partials there are) for each
// tiling candidate, and the face transforms for each kernel variant, on this
// scene; keeps the fastest. Runs at finalize, when the state is all zeros, so
// the timed launches change nothing.
int autotune(Ctx& c) {
  BatchedState& B = *c.bs;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return set_error(ARD_ERR_HIP, "hipEventCreate");
  auto time_ms = [&](auto&& launch, int reps) {
    launch();
    hipEventRecord(e0, c.stream);
    for (int i = 0; i < reps; ++i) launch();
    hipEventRecord(e1, c.stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    return ms / reps;
  };
  // Every candidate is timed in two rounds (three in the transforms') and
  // keeps its fastest: the first launches after set-up run on a GPU whose
  // clocks are still rising, which would favour the candidates timed last.
  struct Tiling { int rpw, kzc, ring; };
  std::vector<Tiling> tilings;
  for (int rpw : {2, 4, 6, 8})
    for (int kzc : {16, 32, 48, 96})
      for (int ring : {0, 2})
        if (!ring || rpw == 4) tilings.push_back(Tiling{rpw, kzc, ring});
  std::vector<float> tms(tilings.size(), 1e30f);
  int tried = 0;
  int r = ARD_OK;
  for (int round = 0; round < 2 && !r; ++round) {
    for (size_t i = 0; i < tilings.size(); ++i) {
      B.ring = tilings[i].ring;
      if ((r = apply_tiling(c, tilings[i].rpw, tilings[i].kzc))) break;
      const BatchedState::Stage& S = B.stage;
      const float ms = time_ms([&] {
        const BatchedState::ModalGroup& g = S.modal;
        launch_modal(g.rpw, g.pts, g.ring, g.d_desc[0], g.d_start, g.n, g.blocks, 0, c.stream);
        launch_unfold(S.unfold, S.unfold_start, S.nunfold, S.unfold_total, c.stream);
      }, 3);
      tms[i] = std::min(tms[i], ms);
    }
  }
  tried += int(tilings.size());
  const size_t bt = size_t(std::min_element(tms.begin(), tms.end()) - tms.begin());
  const int best_rpw = tilings[bt].rpw, best_kzc = tilings[bt].kzc;
  B.ring = tilings[bt].ring;
  // the FP32 transform kernels compete on speed; BF16 split precision, when
  // requested, is kept (a precision choice, not a speed one). They are timed
  // as the step graph runs them, with the PML slabs beside them on the
  // second stream where the step places them (the
  // kernels share the CUs, and one that issues more memory instructions
  // loses more to the PML than it gains alone).
  if (B.gemm_variant != 2) {
    std::vector<int> cand;
    for (int v : {0, 1, 3, 4, 5})
      if (v != 4 || B.has_fft) cand.push_back(v);
    const BatchedState::Stage& S = B.stage;
    hipStream_t st = c.stream;
    int e = 0;
    // the step's transforms with the PML beside them as the step runs it
    auto faces = [&](int v) {
      const bool fork = B.side && B.npml > 0;
      if (fork) e |= hipEventRecord(B.ev_fork, st) != hipSuccess;
      e |= launch_face(B, S, 0, v, st);
      if (fork) {
        e |= hipStreamWaitEvent(B.side, B.ev_fork, 0) != hipSuccess;
        e |= launch_pml_group(c, B, 0, 0, B.side);
      }
      e |= launch_face(B, S, 1, v, st);
      // group 1 forked here and enqueued after the next launch on the main
      // stream, as the step enqueues it after the unfold
      if (fork && B.pml_n[1]) e |= hipEventRecord(B.ev_fork2, st) != hipSuccess;
      e |= launch_face(B, S, 2, v, st);
      if (fork) {
        if (B.pml_n[1]) {
          e |= hipStreamWaitEvent(B.side, B.ev_fork2, 0) != hipSuccess;
          e |= launch_pml_group(c, B, 1, 0, B.side);
        }
        e |= hipEventRecord(B.ev_join, B.side) != hipSuccess;
      }
      e |= launch_face(B, S, 3, v, st);
      if (fork) e |= hipStreamWaitEvent(st, B.ev_join, 0) != hipSuccess;
    };
    std::vector<float> gms(cand.size(), 1e30f);
    for (int round = 0; round < 3; ++round)
      for (size_t i = 0; i < cand.size(); ++i) gms[i] = std::min(gms[i], time_ms([&] { faces(cand[i]); }, 3));
    tried += int(cand.size());
    B.gemm_variant = cand[size_t(std::min_element(gms.begin(), gms.end()) - gms.begin())];
    // then the PML tile shapes, timed the same way with the chosen transforms
    if (B.npml && B.side) {
      float pms[2] = {1e30f, 1e30f};
      for (int round = 0; round < 3 && !e; ++round)
        for (int sh = 0; sh < 2 && !e; ++sh) {
          e |= build_pml_table(c, sh, B.pml_split) != ARD_OK;
          if (!e) pms[sh] = std::min(pms[sh], time_ms([&] { faces(B.gemm_variant); }, 3));
        }
      tried += 2;
      if (!e) e |= build_pml_table(c, pms[1] < pms[0] ? 1 : 0, B.pml_split) != ARD_OK;
    }
    if (e && !r) r = set_error(ARD_ERR_HIP, "autotuner: transform timing");
  }
  // The timed launches ran on the step-0 state: zero fields plus the step-0
  // source samples. The modal pass wrote M+ = -.999 M- + g F over the
  // previous modes M[1] (a residue of the interior point sources' samples,
  // since the -.999 keeps pairs of launches from cancelling), and the PML
  // update its next pressure and phi; both go back to zero (the PML force it
  // cleared held no samples: sources lie in DCT partitions), as does every
  // buffer the step writes before reading.
  if (!r) {
    for (int i : B.dct_ids) {
      const BPart& P = B.parts[size_t(i)];
      const PartGeom& p = c.topo.parts[size_t(i)];
      if ((r = c.dzero(P.M[1], size_t(p.d) * p.h * P.wp * 4))) break;
    }
    for (size_t i = 0; i < c.pml.size() && !r; ++i)
      if (c.pml[i].p[0]) r = c.dzero(c.pml[i].p[0], size_t(c.topo.parts[i].cells()) * 8 * 4);
  }
  if (!r) r = apply_tiling(c, best_rpw, best_kzc);
  B.graphs_valid = false;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (r) return r;
  B.tuned = tried;
  return ARD_OK;
}

int batched_set_option(Ctx& c, int option, int value) {
  if (!c.bs) return set_error(ARD_ERR_STATE, "not a finalized batched context");
  BatchedState& B = *c.bs;
  if (option == ARD_OPT_RING) {
    if (value != 0 && value != 2) return set_error(ARD_ERR_INVALID, "ring depth must be 0 or 2");
    B.ring = value;
    B.graphs_valid = false;
    return build_stage_tiling(c);
  }
  if (option == ARD_OPT_PML_SHAPE) {
    if (value != 0 && value != 1) return set_error(ARD_ERR_INVALID, "PML tile shape must be 0 or 1");
    return build_pml_table(c, value, B.pml_split);
  }
  if (option == ARD_OPT_PML_SPLIT) {
    if (value < 0 || value > 100) return set_error(ARD_ERR_INVALID, "PML split must be a percentage in [0, 100]");
    return build_pml_table(c, B.pml_shape, value);
  }
  return set_error(ARD_ERR_INVALID, "unknown batched option");
}

int batched_option(Ctx& c, int option, int* value) {
  if (!c.bs) return set_error(ARD_ERR_STATE, "not a finalized batched context");
  if (option == ARD_OPT_RING) *value = c.bs->rpw == 4 ? c.bs->ring : 0;
  else if (option == ARD_OPT_PML_SHAPE) *value = c.bs->pml_shape;
  else if (option == ARD_OPT_PML_SPLIT) *value = c.bs->pml_split;
  else return set_error(ARD_ERR_INVALID, "unknown batched option");
  return ARD_OK;
}

int batched_set_tiling(Ctx& c, int rpw, int kzc) {
  if (!c.bs) return set_error(ARD_ERR_STATE, "not a finalized batched context");
  if (rpw != 2 && rpw != 4 && rpw != 6 && rpw != 8) return set_error(ARD_ERR_INVALID, "rows per wave must be 2, 4, 6 or 8");
  // the partial-reduction buffers hold the finest tiling: 16 kz planes per block
  if (kzc < 16 || kzc > 4096) return set_error(ARD_ERR_INVALID, "kz planes per block must be in [16, 4096]");
  return apply_tiling(c, rpw, kzc);
}

int batched_config(Ctx& c, int* out4) {
  if (!c.bs) return set_error(ARD_ERR_STATE, "not a finalized batched context");
  out4[0] = c.bs->rpw;
  out4[1] = c.bs->kz_chunk;
  out4[2] = c.bs->gemm_variant;
  out4[3] = c.bs->tuned;
  return ARD_OK;
}

int step_batched(Ctx& c, int n) {
  BatchedState& B = *c.bs;
  if (c.timer.on) {  // direct launches between events
    for (int s = 0; s < n; ++s) {
      int r = enqueue_step(c, c.time_step & 1, &c.timer);
      if (r) return r;
      c.time_step++;
      if (c.timer.used > 2048 && (r = c.timer.collect())) return r;
    }
    return ARD_OK;
  }
  if (!B.graphs_valid || B.graph_trace != c.d_trace) {
    int r = build_graphs(c);
    if (r) return r;
  }
  for (int s = 0; s < n;) {
    const int par = c.time_step & 1;
    const int k = n - s >= BatchedState::kMultiSteps ? BatchedState::kMultiSteps : 1;
    hipError_t e = hipGraphLaunch(k > 1 ? B.graph_multi[par] : c.graph_exec[par], c.stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    c.time_step += k;
    s += k;
  }
  return ARD_OK;
}

// Dense pressure of a DCT partition from its current modes: the three
// DCT-III passes of the plain form, on demand.
int batched_pressure_field(Ctx& c, int part, const float** out) {
  BatchedState& B = *c.bs;
  const PartGeom& p = c.topo.parts[part];
  const size_t N = size_t(p.cells());
  int r;
  if (B.dense_n < N) {
    for (int k = 0; k < 3; ++k)
      if ((r = c.dalloc((void**)&B.dense[k], N * 4))) return r;
    B.dense_n = N;
    if (!B.d_dense && (r = c.dalloc((void**)&B.d_dense, 3 * sizeof(GemmDesc)))) return r;
  }
  const float *fw, *iw, *fh, *ih, *fd, *id;
  if ((r = c.cos_tables(p.w, &fw, &iw)) || (r = c.cos_tables(p.h, &fh, &ih)) || (r = c.cos_tables(p.d, &fd, &id)))
    return r;
  const long long W = p.w, H = p.h, D = p.d;
  const float* M = B.parts[part].M[c.time_step & 1];
  GemmDesc g[3];
  g[0] = gemm_k_contig(iw, p.w, M, B.dense[0], int(H * D), W);
  g[0].b1 = B.parts[part].wp;  // padded mode rows
  // along y per z-plane: n = (x, z)
  g[1] = GemmDesc{};
  g[1].A = ih; g[1].B = B.dense[0]; g[1].C = B.dense[1];
  g[1].M = p.h; g[1].K = p.h; g[1].N = int(W * D); g[1].N1 = p.w;
  g[1].bK = W; g[1].b1 = 1; g[1].b2 = W * H; g[1].cM = W; g[1].c1 = 1; g[1].c2 = W * H; g[1].alpha = 1.f;
  g[1].tiles_n = int(ceil_div(g[1].N, 64));
  g[2] = GemmDesc{};
  g[2].A = id; g[2].B = B.dense[1]; g[2].C = B.dense[2];
  g[2].M = p.d; g[2].K = p.d; g[2].N = int(W * H); g[2].N1 = g[2].N;
  g[2].bK = W * H; g[2].b1 = 1; g[2].cM = W * H; g[2].c1 = 1; g[2].alpha = 1.f;
  g[2].tiles_n = int(ceil_div(g[2].N, 64));
  for (auto& x : g) x.tile0 = 0;
  if ((r = c.upload(B.d_dense, g, sizeof g))) return r;
  r = launch_gemm(B.d_dense + 0, 1, int(ceil_div(g[0].M, 64)) * g[0].tiles_n, true, true, 0, c.stream);
  r |= launch_gemm(B.d_dense + 1, 1, int(ceil_div(g[1].M, 64)) * g[1].tiles_n, false, false, 0, c.stream);
  r |= launch_gemm(B.d_dense + 2, 1, int(ceil_div(g[2].M, 64)) * g[2].tiles_n, false, false, 0, c.stream);
  if (r) return set_error(ARD_ERR_HIP, "dense evaluation launch");
  *out = B.dense[2];
  return ARD_OK;
}

// The mode buffers of a DCT partition (by step parity) and their padded row length.
int batched_modes(Ctx& c, int part, const float* M[2], int* wp) {
  const BPart& P = c.bs->parts[size_t(part)];
  M[0] = P.M[0];
  M[1] = P.M[1];
  *wp = P.wp;
  return ARD_OK;
}

void batched_state_arrays(Ctx& c, int part, std::vector<std::pair<const float*, long long>>& out) {
  const BPart& P = c.bs->parts[size_t(part)];
  const PartGeom& g = c.topo.parts[size_t(part)];
  const long long np = (long long)g.d * g.h * P.wp;  // padded mode rows (the padding holds zeros)
  out.emplace_back(P.M[0], np);
  out.emplace_back(P.M[1], np);
  for (int w = 0; w < 3; ++w) {
    out.emplace_back(P.F[w], 6 * P.plane[w]);
    out.emplace_back(P.P[w], 6 * P.plane[w]);
  }
}

// Partition::set_force on a DCT partition: a face-slab cell is written in
// place; an interior cell becomes a constant point force (forces of DCT
// partitions persist, ARD/dct_partition.cpp:48-73).
int batched_set_force(Ctx& c, int part, int x, int y, int z, float f) {
  BatchedState& B = *c.bs;
  const PartGeom& p = c.topo.parts[part];
  long long off;
  const int w = shell_of(p.w, p.h, p.d, x, y, z, &off);
  if (w >= 0) {
    // a face-slab cell; if it is a source's, the source overwrites it before
    // the next update, as in the reference (ARD/simulation.cpp:306-312)
    const int r = c.upload(B.parts[part].F[w] + off, &f, 4);
    return r ? r : write_sources(c);
  }
  for (auto& s : B.srcs)
    if (s.part == part && s.x == x && s.y == y && s.z == z) {
      // a source cell: the source overwrites it before the next update anyway
      if (s.kind != 2) return ARD_OK;
      s.value = f;
      int r = build_sources(c);
      if (r) return r;
      B.graphs_valid = false;
      return build_modal(c);
    }
  PointSrc ps;
  ps.part = part; ps.x = x; ps.y = y; ps.z = z; ps.kind = 2; ps.value = f;
  B.srcs.push_back(ps);
  int r = build_sources(c);
  if (r) return r;
  B.graphs_valid = false;
  return build_modal(c);
}

void destroy_batched(Ctx& c) {
  delete c.bs;
  c.bs = nullptr;
}

// Self-test of the folded transform kernels (ard_selftest_fold, _bf16,
// _direct): one transform of length L over N columns, built exactly as the
// batched form builds its face passes (fold_tables, a folded GemmDesc,
// fold_tile_table), in a temporary context of its own. kind: 0 LDS-staged
// FP32, 1 BF16 split, 2 direct-to-register FP32.
int fold_selftest(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y, int kind) {
  const bool bf16 = kind == 1;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  Ctx c;
  c.device = device;
  if ((e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess)
    return hip_fail(e, "hipStreamCreate");
  int r;
  const float *ff, *fi;
  if ((r = fold_tables(c, L, &ff, &fi))) return r;
  const size_t n = size_t(L) * N;
  float *dX, *dY;
  // the kernel's 16-byte B loads may end up to 3 floats past the array
  if ((r = c.dalloc((void**)&dX, (n + 16) * 4)) || (r = c.dalloc((void**)&dY, n * 4))) return r;
  if ((r = c.dzero(dX, (n + 16) * 4)) || (r = c.upload(dX, X, n * 4))) return r;
  GemmDesc g{};
  g.B = dX; g.C = dY;
  g.N = N; g.N1 = N;
  if (b_layout == 0) { g.bK = N; g.b1 = 1; g.cM = N; g.c1 = 1; }   // X, Y [L][N]
  else { g.bK = 1; g.b1 = L; g.cM = 1; g.c1 = L; }                 // X, Y [N][L]
  g.alpha = 1.f;
  g.b_elems = unsigned(n);
  g.L = L;
  g.K = (L + 1) / 2;
  g.Kp = fold_kp(L);
  g.Ap = inverse ? fold_rows_inv(L) : fold_rows_fwd(L);
  if (!inverse) {
    g.A = ff;
    g.Mh = fold_even_rows(L);
    g.M = g.Mh + L / 2;
  } else {
    g.A = fi;
    g.M = int(64 * ceil_div((L + 1) / 2, 32));
  }
  if (bf16) {
    const void* t[4];
    if ((r = fold_tables_bf16(c, L, t))) return r;
    g.A = static_cast<const float*>(inverse ? t[2] : t[0]);
    g.Alo = inverse ? t[3] : t[1];
  }
  g.tiles_n = int(ceil_div(N, 64));
  const int tile_m = kind == 2 ? 32 : 64;
  if (kind == 2) {  // pair rows, tiles of 32 x 128
    g.M = (L + 1) / 2;
    g.tiles_n = int(ceil_div(N, 128));
  }
  std::vector<GemmDesc> v{g};
  GemmDesc* dd;
  int2* dt;
  if ((r = upload_vec(c, (void**)&dd, v.data(), sizeof(GemmDesc))) || (r = fold_tile_table(c, v, &dt, tile_m))) return r;
  const int tiles = int(ceil_div(g.M, tile_m)) * g.tiles_n;
  const bool kc = b_layout == 1;
  const int fold = inverse ? 2 : 1;
  const int lr = kind == 2   ? launch_fold_direct(dd, dt, tiles, kc, fold, c.stream)
                 : kind == 1 ? launch_gemm_bf16_fold(dd, dt, tiles, kc, fold, c.stream)
                             : launch_gemm_fold(dd, dt, tiles, kc, fold, c.stream);
  if (lr)
    return set_error(ARD_ERR_HIP, "folded transform launch");
  if ((e = hipMemcpyAsync(Y, dY, n * 4, hipMemcpyDeviceToHost, c.stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c.stream)) != hipSuccess)
    return hip_fail(e, "folded transform result");
  return ARD_OK;
}

}  // namespace ard
