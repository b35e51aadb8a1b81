// Kernels of the batched face-shell form of the ARD step.
//
// A DCT partition's force field is nonzero only within 3 cells of its faces
// (interface forcing, ARD/boundary.cpp:30-173) and at source cells
// (ARD/partition.cpp:187-197); its pressure is read only there and at the
// recorder cells (ARD/recorder.cpp:40-57). The 3D DCT-II of the force and the
// DCT-III to pressure (ARD/dct_volume.cpp:38-68) are therefore evaluated as
//   * 2D transforms of the six 3-cell face slabs (small GEMMs), and
//   * one streaming pass over the modes (k_modal) that expands the face
//     transforms into the force spectrum, applies the modal update
//     (ARD/dct_partition.cpp:66-70) and reduces the new modes back onto the
//     face slabs and recorder points,
// which computes the same quantities as the dense transforms with O(1) work
// per cell instead of O(W+H+D).
//
// Notation for one partition (W x H x D cells, modes stored times the inverse
// scale 1/sqrt(2N)): c_n(k, x) = cos(pi k (x + 1/2) / n), e(k) = 1 for k = 0
// else 2. For x in {0,1,2} the mirror cell n-1-x has c_n(k, n-1-x) =
// (-1)^k c_n(k, x), so each face pair is carried as a sum (even k) and a
// difference (odd k) plane.
#include <type_traits>

#include "ard_batched.h"

namespace ard {

namespace {
__device__ __forceinline__ int find_desc_ll(const long long* __restrict__ starts, int n, long long key) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
}  // namespace

typedef __attribute__((address_space(1))) v4f gv4f;

// ---------------------------------------------------------------------------
// Unfold: sum the per-chunk partial reductions and turn the (even, odd) pair
// back into the face pair: x -> even + odd, n-1-x -> even - odd.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_unfold(const UnfoldDesc* __restrict__ d,
                                                const long long* __restrict__ starts, int n) {
  const long long b0 = (long long)blockIdx.x * 256;
  const int i = find_desc_ll(starts, n, b0);  // block-uniform
  const UnfoldDesc& U = d[i];
  const long long it = b0 - starts[i] + threadIdx.x;
  const long long pl = U.plane;
  if (it * U.vec >= 3 * pl) return;
  const long long cell = it * U.vec;
  const int q = int(cell / pl);
  const long long o = cell - q * pl;
  const gfloat* p = ARD_CG(U.part) + q * pl + o;
  gfloat* dst = ARD_G(U.dst);
  const long long cs = 6 * pl;  // chunk stride
  const int nc = U.nchunk;
  if (U.vec == 4) {
    v4f ev = v4f{0.f, 0.f, 0.f, 0.f}, od = ev;
    int c = 0;
    for (; c + 1 < nc; c += 2) {  // two chunks' loads in flight per iteration
      const v4f e0 = *reinterpret_cast<const gv4f*>(p + c * cs);
      const v4f o0 = *reinterpret_cast<const gv4f*>(p + c * cs + 3 * pl);
      const v4f e1 = *reinterpret_cast<const gv4f*>(p + (c + 1) * cs);
      const v4f o1 = *reinterpret_cast<const gv4f*>(p + (c + 1) * cs + 3 * pl);
      ev.x += e0.x; ev.y += e0.y; ev.z += e0.z; ev.w += e0.w;
      od.x += o0.x; od.y += o0.y; od.z += o0.z; od.w += o0.w;
      ev.x += e1.x; ev.y += e1.y; ev.z += e1.z; ev.w += e1.w;
      od.x += o1.x; od.y += o1.y; od.z += o1.z; od.w += o1.w;
    }
    if (c < nc) {
      const v4f e0 = *reinterpret_cast<const gv4f*>(p + c * cs);
      const v4f o0 = *reinterpret_cast<const gv4f*>(p + c * cs + 3 * pl);
      ev.x += e0.x; ev.y += e0.y; ev.z += e0.z; ev.w += e0.w;
      od.x += o0.x; od.y += o0.y; od.z += o0.z; od.w += o0.w;
    }
    *reinterpret_cast<gv4f*>(dst + q * pl + o) = v4f{ev.x + od.x, ev.y + od.y, ev.z + od.z, ev.w + od.w};
    *reinterpret_cast<gv4f*>(dst + (5 - q) * pl + o) =
        v4f{ev.x - od.x, ev.y - od.y, ev.z - od.z, ev.w - od.w};
  } else {
    float ev = 0.f, od = 0.f;
    for (int c = 0; c < nc; ++c) {
      ev += p[c * cs];
      od += p[c * cs + 3 * pl];
    }
    dst[q * pl + o] = ev + od;
    dst[(5 - q) * pl + o] = ev - od;
  }
}

// ---------------------------------------------------------------------------
// Source samples: GaussianSource (ARD/gaussian_source.cpp:16-20), a waveform,
// or a constant (a force set through set_force at an interior cell). Shell
// cells receive the value in the face slab; interior cells in a value slot
// read by k_modal.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float bsource_value(const BSourceDesc& s, int t, float courant) {
  if (s.kind == 0) {
    const float x = 3.14159265358979f * ((2.f * courant * float(t)) / 6.f - 2.f);
    return 1e9f * expf(-(x * x));
  }
  if (s.kind == 1) return t < s.n ? s.wave[t] : 0.f;
  return s.value;
}

// the samples of the current step (*d_step): at set-up, and whenever the
// sources or a source cell's force change between steps; within a step the
// recorder kernel writes the next step's
__global__ void k_bsources(const BSourceDesc* __restrict__ s, int n, const int* __restrict__ d_step,
                           float courant) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  *s[i].dst = bsource_value(s[i], *d_step, courant);
}

// ---------------------------------------------------------------------------
// The modal pass. Block = 4 waves; lane <-> kx (64 per block), wave w owns
// rows ky = kyb + w RPW + r (r < RPW), the block walks kz over its chunk.
// Per mode:   Fh = sum_i Ex[kx][i] Gx[par kx][i][kz][ky] + (y, z alike) + sources
//             M+ = a M - b M- + g Fh           (stored over M-)
// and reduces M+ onto the face planes and recorders:
//   Tx[par kx][i][kz][ky] += Rx[kx][i] M+   (across lanes)
//   Ty[par ky][j][kz][kx] += Ry[ky][j] M+   (across rows and waves)
//   Tz[par kz][l][ky][kx] += Rz[kz][l] M+   (along the kz walk)
// Partial sums go to per-chunk slots, summed by k_unfold.
// ---------------------------------------------------------------------------
// sin^2(x) / x^2 as a polynomial in u = x^2, for 0 <= u <= 3 (the stability
// limit of the scheme puts x = w dt / 2 at most at pi / 2, u <= 2.47): the
// series of (1 - cos 2x) / (2 x^2), c_n = (-1)^(n+1) 2^(2n-1) / (2n)!, to
// n = 10; truncation error < 3e-11, FP32 Horner error < 4e-7 relative.
__device__ __forceinline__ float sinc2_poly(float u) {
  float p = -2.1549902e-13f;
  p = fmaf(p, u, 2.0472407e-11f);
  p = fmaf(p, u, -1.5661391e-9f);
  p = fmaf(p, u, 9.3968346e-8f);
  p = fmaf(p, u, -4.2755596e-6f);
  p = fmaf(p, u, 1.4109348e-4f);
  p = fmaf(p, u, -3.1746032e-3f);
  p = fmaf(p, u, 4.4444446e-2f);
  p = fmaf(p, u, -3.3333334e-1f);
  return fmaf(p, u, 1.0f);
}
// streamed once per step (the mode state): non-temporal ("nt") loads
__device__ __forceinline__ float bload_nt(brsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 2));
}

// Modal pass of the batched form, one block per (64 kx) x (4 RPW ky) x (kz
// chunk) tile of one partition. For every mode:
//   force spectrum   fh = sum of the face-slab expansions (x, y, z faces) + point forces
//   modal update     M+ = a M - .999 M- + g fh         (ARD/dct_partition.cpp:66-70)
//   reductions       Tx, Ty, Tz: M+ projected onto the face slabs (x, y, z), and recorders
// Wave w owns rows ky0 = kyb + w RPW .. ky0 + RPW - 1 (parity r & 1 known at
// compile time) and walks the chunk's kz planes in pairs (parity known), with
// the next plane's loads in flight during the current one. z sums stay in
// registers. The chunk goes in batches of KB planes:
//   planes   each wave: new modes of its rows -> global and the LDS mode tile
//   x sums   each wave from its own rows of the tile (no barrier), with the
//            recorder sums
//   barrier
//   y sums   of the batch from the tile (one (plane, kx) per thread, over the
//            block's rows); the next batch's per-plane factors (x-face
//            spectra, z tables, point z factors), loaded a batch ahead, to LDS
//   barrier
// so the plane loop reads no scalar memory and no LDS atomics (float atomics
// on LDS serialise), and the loads of the next plane stay in flight across
// the batch's sums and barriers.
// At most 4 rows per wave the kernel fits 3 waves per SIMD (168 VGPRs); the
// attribute keeps the compiler there. Taller tiles hold more and get 2.
//
// RING > 0 (4 rows per wave): the mode rows (M, M-) of the next two planes
// stream into a per-wave LDS ring by direct-to-LDS loads (buffer_load_dwordx4
// ... lds: lane l's 16 bytes land at 16 l, so one instruction moves the wave's
// 4 rows x 64 kx of one array), instead of one plane ahead in registers: two
// planes in flight per wave at no register cost. The ring's two slots (even
// and odd planes) are separate LDS arrays, so that the compiler's own
// tracking of direct-to-LDS loads (per LDS variable) never makes a read of one
// slot wait for the loads into the other; the kernel waits for a slot's loads
// itself, by count (see the plane step). Block barriers inside the plane loop
// wait for this wave's LDS accesses only (a fence would wait for the ring's
// loads too), and the ring drains before the block ends (no load may write
// LDS after the block has released it).
template <int RPW, bool PTS, int RING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RPW <= 4 ? 3 : 1))) void k_modal(const ModalDesc* __restrict__ descs,
                                               const long long* __restrict__ starts, int ndesc, int par) {
  static_assert(RPW % 2 == 0, "rows per wave must be even");
  static_assert(RING == 0 || (RPW == 4 && RING == 2), "the ring holds 4 rows per wave, 2 planes (one per plane parity)");
  constexpr int TY = 4 * RPW;  // rows of the block tile
  constexpr int KB = 4;        // kz planes per batch of the x and y reductions
  constexpr int LD = 80;       // padded row of the mode tile: two rows 16 banks apart (x sums read 2 rows per 8 lanes)
  constexpr int NPT = ARD_MAX_SRC_PER_PART;
  constexpr int NRC = ARD_MAX_REC_PER_PART;
  static_assert(NPT == 4 && NRC == 4, "point tables hold 4 + 4 factors");
  __shared__ __attribute__((aligned(16))) float s_m[KB][TY][LD];     // new modes of a batch
  __shared__ __attribute__((aligned(16))) float s_ry[TY][4];         // y reduction weights of the tile's rows
  __shared__ __attribute__((aligned(16))) float s_rx[64 * 3];        // x reduction weights of the tile's kx
  __shared__ __attribute__((aligned(16))) float s_gx[KB][TY][8];     // x-face spectra of the batch's rows: even 0..2, odd 4..6
  __shared__ __attribute__((aligned(16))) float s_zt[KB][8];         // z factors of the batch's planes
  __shared__ __attribute__((aligned(16))) float s_pty[PTS ? TY : 1][8];  // point y factors of the tile's rows: sources 0..3, recorders 4..7
  __shared__ __attribute__((aligned(16))) float s_ptz[PTS ? KB : 1][8];  // point z factors of the batch's planes, likewise
  __shared__ __attribute__((aligned(16))) float s_rfx[PTS ? NRC : 1][64];     // recorder x factors
  __shared__ float s_rec[4][NRC];
  // the ring's two slots (even and odd planes) are separate arrays, so that
  // the compiler, which tracks direct-to-LDS loads per LDS variable, does not
  // make a read of one slot wait for the loads into the other
  __shared__ __attribute__((aligned(16))) float s_ring0[RING > 0 ? 4 : 1][2][RPW][64];
  __shared__ __attribute__((aligned(16))) float s_ring1[RING > 0 ? 4 : 1][2][RPW][64];
  const int bid = xcd_block(int(blockIdx.x), int(gridDim.x));
  const int di = find_desc_ll(starts, ndesc, bid);
  const ModalDesc& P = descs[di];
  const int lb = bid - int(starts[di]);
  const int bx = lb % P.nbx;
  const int by = (lb / P.nbx) % P.nby;
  const int bz = lb / (P.nbx * P.nby);
  const int W = P.W, H = P.H, D = P.D;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int kx = bx * 64 + lane;
  const bool xok = kx < W;
  const int kxc = xok ? kx : W - 1;
  const int px = kx & 1;
  const int kyb = by * TY;
  const int ky0 = kyb + wave * RPW;  // even
  const int kz0 = bz * P.kz_chunk;   // kz_chunk is even, so kz0 is even
  const int kz1 = min(D, kz0 + P.kz_chunk);
  const int HW = H * W, DH = D * H, DW = D * W;  // partitions hold < 2^30 cells
  const int Wp = P.Wp, HWp = H * Wp;            // mode arrays: rows padded to 128-byte lines
  const brsrc rMc = make_rsrc(P.M[par], unsigned(D * HWp) * 4u);
  const brsrc rMp = make_rsrc(P.M[par ^ 1], unsigned(D * HWp) * 4u);
  const brsrc rGy = make_rsrc(P.Gy, unsigned(6 * DW) * 4u);
  gfloat* __restrict__ Mp = ARD_G(P.M[par ^ 1]);
  gfloat* __restrict__ Txp = ARD_G(P.Tx) + (long long)bx * 6 * DH;
  gfloat* __restrict__ Typ = ARD_G(P.Ty) + (long long)by * 6 * DW;
  const float gsc = P.gscale;  // .999 * 4 * s_f * s_i * dt^2 / 4
  const unsigned vx = unsigned(kxc) * 4u;
  const int nsrc = PTS ? min(P.nsrc, NPT) : 0;
  const int nrec = PTS ? min(P.nrec, NRC) : 0;

  // x factors; the kx parity picks the folded spectra they multiply (even
  // modes the sums, odd modes the differences)
  float ex[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) ex[i] = ARD_CG(P.Ex)[kxc * 3 + i];
  if (wave == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) s_rx[lane * 3 + i] = xok ? ARD_CG(P.Rx)[kxc * 3 + i] : 0.f;
  }
  if (PTS && wave == 1) {
#pragma unroll
    for (int q = 0; q < NRC; ++q) s_rfx[q][lane] = q < nrec && xok ? P.rec[q].fx[kxc] : 0.f;
  }
  if (tid < TY * 4) {  // rows past H carry zero modes, any weight will do
    const int row = tid >> 2, j = tid & 3;
    s_ry[row][j] = j < 3 ? ARD_CC(P.yt)[min(kyb + row, H - 1) * 8 + 3 + j] : 0.f;
  }
  const float ux = ARD_CG(P.wx2)[kxc];
  // per-row factors of this wave's rows (wave-uniform)
  float ey[RPW][3], uy[RPW];
  bool rok[RPW];
  int rof[RPW], rom[RPW];  // row offsets ky * W of the face tables and ky * Wp of the modes (clamped)
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int kyc = min(ky0 + r, H - 1);
    rok[r] = ky0 + r < H;
    rof[r] = kyc * W;
    rom[r] = kyc * Wp;
    const cfloat* yt = ARD_CC(P.yt) + kyc * 8;
    uy[r] = yt[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) ey[r][j] = yt[j];
  }
  // z-face force spectrum of both kz parities, folded from the six slabs'
  // spectra (slab q at z = q, its mirror 5 - q at z = D - 1 - q, and
  // c_D(kz, D - 1 - z) = (-1)^kz c_D(kz, z)): q < 3 the sums (even kz), q >= 3
  // the differences (odd kz); and the z sums of this thread's rows
  float gz[RPW][6], tz[RPW][6];
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    float raw[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      raw[q] = ARD_CG(P.Gz)[q * HW + rof[r] + kxc];
      tz[r][q] = 0.f;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      gz[r][q] = raw[q] + raw[5 - q];
      gz[r][3 + q] = raw[q] - raw[5 - q];
    }
  }
  // point forces: x factor (times the source value) per lane
  float sfx[NPT];
#pragma unroll
  for (int q = 0; q < NPT; ++q) sfx[q] = PTS && q < nsrc ? P.src[q].fx[kxc] * *P.src[q].val : 0.f;
  float rec_acc[NRC];  // recorder sums of this thread's (plane, row) slots
#pragma unroll
  for (int q = 0; q < NRC; ++q) rec_acc[q] = 0.f;

  // factors of one batch staged in LDS: x-face spectra (6 of 8 per row),
  // z tables (8 per plane), point z factors (8 per plane) and point y
  // factors (8 per row: the same every batch). Loaded into registers one
  // batch ahead, stored after the batch's first barrier; no arithmetic on
  // them before then, so no wait for their loads.
  constexpr int NGX = KB * TY * 8, NZT = KB * 8, NPC = PTS ? KB * 8 : 0, NPY = PTS ? TY * 8 : 0;
  constexpr int NSTG = (NGX + NZT + NPC + NPY + 255) / 256;
  // the x-face spectra fill the first NSX slots of every thread: two loads
  // each (a slab's spectrum and its mirror's), folded when stored
  constexpr int NSX = NGX / 256;
  static_assert(NGX % 256 == 0 && NSX >= 1, "x-face staging must fill whole slots");
  float stg[NSTG], stg2[NSX];
  const gfloat* const gxt = pinned(P.Gx);
  const gfloat* const ztt = pinned(P.zt);
  const gfloat* const pzt = pinned(PTS ? P.ptz : P.zt);
  const gfloat* const pyt = pinned(PTS ? P.pty : P.zt);
  // one load per slot, every lane, with only its address selected: a load
  // under a branch would make the compiler wait for all loads in flight at
  // the branch's join
  auto stage_load = [&](const int kb) {
#pragma unroll
    for (int t = 0; t < NSTG; ++t) {
      const int e = tid + 256 * t;
      if (t < NSX) {
        // slots 0..2 the even-kx spectra (sum of slab i = sl and its mirror
        // 5 - i), 4..6 the odd ones (difference); slots 3 and 7 are never read
        const int sl = e & 7, row = (e >> 3) % TY, kz = min(kb + e / (8 * TY), kz1 - 1);
        const int ky = min(kyb + row, H - 1);
        const int i = min(sl & 3, 2), base = (kz * H + ky) * 6;
        stg[t] = gxt[base + i];
        stg2[t < NSX ? t : 0] = gxt[base + 5 - i];
        continue;
      }
      const gfloat* src = ztt;
      int idx = 0;  // past the staged tables: any valid address, value unused
      if (e < NGX + NZT) {
        const int f = e - NGX;
        idx = min(kb + (f >> 3), kz1 - 1) * 8 + (f & 7);
      } else if (PTS && e < NGX + NZT + NPC) {
        const int f = e - NGX - NZT;
        src = pzt;
        idx = min(kb + (f >> 3), kz1 - 1) * 8 + (f & 7);
      } else if (PTS && e < NGX + NZT + NPC + NPY) {
        const int f = e - NGX - NZT - NPC;
        src = pyt;
        idx = min(kyb + (f >> 3), H - 1) * 8 + (f & 7);
      }
      stg[t] = src[idx];
    }
  };
  auto stage_store = [&]() {
#pragma unroll
    for (int t = 0; t < NSTG; ++t) {
      const int e = tid + 256 * t;
      if (t < NSX) (&s_gx[0][0][0])[e] = stg[t] + ((tid & 4) ? -stg2[t < NSX ? t : 0] : stg2[t < NSX ? t : 0]);
      else if (e < NGX + NZT) (&s_zt[0][0])[e - NGX] = stg[t];
      else if (PTS && e < NGX + NZT + NPC) (&s_ptz[0][0])[e - NGX - NZT] = stg[t];
      else if (PTS && e < NGX + NZT + NPC + NPY) (&s_pty[0][0])[e - NGX - NZT - NPC] = stg[t];
    }
  };

  // streamed operands of one plane: modes of the rows (clamped rows; in the
  // LDS ring when RING > 0) and both y-parity spectra
  struct PlaneIn {
    float mc[RING > 0 ? 1 : RPW], mp[RING > 0 ? 1 : RPW], gy[6];
  };
  auto load = [&](PlaneIn& in, const int kz) {
    const unsigned po = unsigned(kz * HWp);
    if constexpr (RING == 0) {
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        in.mc[r] = bload_nt(rMc, vx, (po + unsigned(rom[r])) * 4u);
        in.mp[r] = bload_nt(rMp, vx, (po + unsigned(rom[r])) * 4u);
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) in.gy[q] = bload(rGy, vx, unsigned(q * DW + kz * W) * 4u);
  };
  // ring loads: lane l moves kx = 4 (l % 16) .. +3 of row l / 16 (clamped
  // rows; kx past W read padding or the next row, never stored)
  const int rsel = lane >> 4;
  const unsigned rvoff =
      unsigned((rsel == 0 ? rom[0] : rsel == 1 ? rom[1 % RPW] : rsel == 2 ? rom[2 % RPW] : rom[3 % RPW]) + bx * 64 +
               4 * (lane & 15)) * 4u;
  auto ring_load = [&](const int kz, auto slotc, const bool idle) {
    if constexpr (RING > 0) {
      constexpr int SL = decltype(slotc)::value;
      auto& ring = SL == 0 ? s_ring0 : s_ring1;
      asm volatile("" ::: "memory");
      const unsigned so = idle ? 0u : unsigned(kz * HWp) * 4u;
      const unsigned vo = idle ? 0u : rvoff;
#if defined(__HIP_DEVICE_COMPILE__)  // the host pass must not see the device-only builtin (it would drop the kernel's stub)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rMc, &ring[wave][0][0][0], 16, vo, so, 0, 2);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rMp, &ring[wave][1][0][0], 16, vo, so, 0, 2);
#endif
      asm volatile("" ::: "memory");
    }
  };
  // block barrier; with the ring, one that waits for this wave's LDS accesses
  // only (__syncthreads' fence would also wait for every direct-to-LDS load
  // in flight, none of which another wave reads)
  auto block_barrier = [&]() {
    if constexpr (RING > 0) asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else __syncthreads();
  };

  // one kz plane of this wave's rows; a plane past the batch (live false)
  // stores nothing and adds zeros
  auto plane = [&](const PlaneIn& in, const int kz, const int kl, auto pzc, const bool live) {
    constexpr int PZ = decltype(pzc)::value;
    float mc[RPW], mp[RPW];
    if constexpr (RING > 0) {
      // the slot of this plane's parity. Its loads were issued at the end of
      // the step two planes back; since then this wave issued at least 14
      // vector-memory operations (two steps' y-spectrum loads, 6 each, and the
      // other slot's 2 ring loads; 12 counted here), and such operations
      // complete in issue order, so waiting until at most 12 are outstanding
      // guarantees them. The compiler's own tracking of direct-to-LDS loads is
      // not relied on (it drops its wait when they are issued under a branch).
      auto& ring = PZ == 0 ? s_ring0 : s_ring1;
      asm volatile("s_waitcnt vmcnt(12)" ::: "memory");
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        mc[r] = ring[wave][0][r][lane];
        mp[r] = ring[wave][1][r][lane];
      }
    } else {
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        mc[r] = in.mc[r];
        mp[r] = in.mp[r];
      }
    }
    // y-face spectra of both ky parities from the six slabs' (as for z)
    float gy[6];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      gy[q] = in.gy[q] + in.gy[5 - q];
      gy[3 + q] = in.gy[q] - in.gy[5 - q];
    }
    const float4 za = *reinterpret_cast<const float4*>(&s_zt[kl][0]);  // Ez[kz][0..2], Rz[kz][0]
    const float4 zb = *reinterpret_cast<const float4*>(&s_zt[kl][4]);  // Rz[kz][1..2], u_z, 0
    const float uxz = ux + zb.z;
    const unsigned po = unsigned(kz * HWp);
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int PY = r & 1;
      const int row = wave * RPW + r;
      const float4 ga = *reinterpret_cast<const float4*>(&s_gx[kl][row][4 * px]);  // this kx parity's x spectra
      // modal coefficients (ARD/dct_partition.cpp:24-33,68) from u = (w dt / 2)^2,
      // separable in the 1-based wavenumbers:
      //   s^2 = sin^2(w dt / 2) = u S(u),  a = .999 * 2 cos(w dt) = 1.998 - 3.996 s^2,
      //   g = .999 * 4 s^2 / w^2 * s_f s_i = gscale * S(u)
      const float u = uxz + uy[r];
      const float S = sinc2_poly(u);
      const float a = fmaf(-3.996f, u * S, 1.998f);
      const float g = gsc * S;
      float fh = ex[0] * ga.x;
      fh = fmaf(ex[1], ga.y, fh);
      fh = fmaf(ex[2], ga.z, fh);
      fh = fmaf(ey[r][0], gy[PY * 3 + 0], fh);
      fh = fmaf(ey[r][1], gy[PY * 3 + 1], fh);
      fh = fmaf(ey[r][2], gy[PY * 3 + 2], fh);
      fh = fmaf(za.x, gz[r][PZ * 3 + 0], fh);
      fh = fmaf(za.y, gz[r][PZ * 3 + 1], fh);
      fh = fmaf(za.z, gz[r][PZ * 3 + 2], fh);
      if (PTS && nsrc > 0) {
        const float4 py = *reinterpret_cast<const float4*>(&s_pty[PTS ? row : 0][0]);
        const float4 pz = *reinterpret_cast<const float4*>(&s_ptz[PTS ? kl : 0][0]);
        fh = fmaf(sfx[0] * py.x, pz.x, fh);
        fh = fmaf(sfx[1] * py.y, pz.y, fh);
        fh = fmaf(sfx[2] * py.z, pz.z, fh);
        fh = fmaf(sfx[3] * py.w, pz.w, fh);
      }
      const float mv = fmaf(a, mc[r], fmaf(-0.999f, mp[r], g * fh));
      const bool ok = rok[r] && xok && live;
      if (ok) __builtin_nontemporal_store(mv, &Mp[po + unsigned(rom[r]) + unsigned(kx)]);  // streamed, not reused in this step
      const float mn = ok ? mv : 0.f;
      s_m[kl][row][lane] = mn;
      tz[r][PZ * 3 + 0] = fmaf(za.w, mn, tz[r][PZ * 3 + 0]);
      tz[r][PZ * 3 + 1] = fmaf(zb.x, mn, tz[r][PZ * 3 + 1]);
      tz[r][PZ * 3 + 2] = fmaf(zb.y, mn, tz[r][PZ * 3 + 2]);
    }
    // the slot just read takes the plane RING steps ahead; past the chunk's
    // end the step still issues its two loads (the same count every step, no
    // branch for the compiler's wait counts to merge), all lanes from the
    // array's first 16 bytes: one cached line instead of a plane re-read
    if (RING > 0) ring_load(kz + RING, pzc, kz + RING >= kz1);
  };

  // x and y sums of a finished batch (planes kb .. kb+n-1):
  // x (xreduce, before the batch barrier): each wave sums its own KB * RPW
  // (plane, row) slots of the mode tile, four lanes per slot; lane qtr of a
  // slot takes kx = 16 j + 4 qtr + t (j, t < 4) and the four meet through
  // quad permutes. The lanes of a slot share its trip count, so every quad is
  // whole. The recorders' x sums ride along, weighted by their y*z factors.
  // y (ysum, after it): the four waves' sums of each (plane, j, kx), four kx
  // per thread.
  auto xreduce = [&](const int kb, const int n) {
    const int qtr = lane & 3;
    for (int cc = lane >> 2; cc < KB * RPW; cc += 16) {
      const int kl = cc / RPW, row = wave * RPW + cc % RPW;
      float ev0 = 0.f, ev1 = 0.f, ev2 = 0.f, od0 = 0.f, od1 = 0.f, od2 = 0.f;
      float rs[NRC] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1  // one group of 4 kx at a time: few registers live beside the prefetched plane
      for (int j = 0; j < 4; ++j) {
        const int k0 = 16 * j + 4 * qtr;  // even
        const float4 m = *reinterpret_cast<const float4*>(&s_m[kl][row][k0]);
        const float4 ra = *reinterpret_cast<const float4*>(&s_rx[k0 * 3]);      // kx k0: x y z, k0+1: x
        const float4 rb = *reinterpret_cast<const float4*>(&s_rx[k0 * 3 + 4]);  // k0+1: y z, k0+2: x y
        const float4 rc = *reinterpret_cast<const float4*>(&s_rx[k0 * 3 + 8]);  // k0+2: z, k0+3: x y z
        ev0 = fmaf(ra.x, m.x, ev0); ev1 = fmaf(ra.y, m.x, ev1); ev2 = fmaf(ra.z, m.x, ev2);
        od0 = fmaf(ra.w, m.y, od0); od1 = fmaf(rb.x, m.y, od1); od2 = fmaf(rb.y, m.y, od2);
        ev0 = fmaf(rb.z, m.z, ev0); ev1 = fmaf(rb.w, m.z, ev1); ev2 = fmaf(rc.x, m.z, ev2);
        od0 = fmaf(rc.y, m.w, od0); od1 = fmaf(rc.z, m.w, od1); od2 = fmaf(rc.w, m.w, od2);
        if (PTS) {
#pragma unroll
          for (int q = 0; q < NRC; ++q) {
            if (q < nrec) {
              const float4 f = *reinterpret_cast<const float4*>(&s_rfx[PTS ? q : 0][k0]);
              rs[q] = fmaf(f.x, m.x, fmaf(f.y, m.y, fmaf(f.z, m.z, fmaf(f.w, m.w, rs[q]))));
            }
          }
        }
      }
#define ARD_QUAD_SUM(v)                                                                    \
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false)); \
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
      ARD_QUAD_SUM(ev0) ARD_QUAD_SUM(ev1) ARD_QUAD_SUM(ev2) ARD_QUAD_SUM(od0) ARD_QUAD_SUM(od1) ARD_QUAD_SUM(od2)
      const int ky = kyb + row;
      if (PTS && nrec > 0) {
#pragma unroll
        for (int q = 0; q < NRC; ++q) {
          if (q < nrec) {
            ARD_QUAD_SUM(rs[q])
            // rows past H and planes past the batch hold no modes of this step
            const float w =
                (kl < n && ky < H && qtr == 0) ? s_pty[PTS ? row : 0][4 + q] * s_ptz[PTS ? kl : 0][4 + q] : 0.f;
            rec_acc[q] = fmaf(w, rs[q], rec_acc[q]);
          }
        }
      }
#undef ARD_QUAD_SUM
      if (kl < n && ky < H && qtr < 2) {
        const int o = (kb + kl) * H + ky;
        if (qtr == 0) {
          Txp[0 * DH + o] = ev0; Txp[1 * DH + o] = ev1; Txp[2 * DH + o] = ev2;
        } else {
          Txp[3 * DH + o] = od0; Txp[4 * DH + o] = od1; Txp[5 * DH + o] = od2;
        }
      }
    }
  };
  // y sums of the batch (after the barrier: every wave's rows are in the
  // tile): thread (plane tid / 64, kx lane) sums the block's rows, even rows
  // into q < 3 and odd rows into q >= 3; rows past H hold zeros.
  auto ysum = [&](const int kb, const int n) {
    const int pl = tid >> 6;  // KB == 4 planes, one per wave
    if (pl >= n || !xok) return;
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, o0 = 0.f, o1 = 0.f, o2 = 0.f;
#pragma unroll
    for (int row = 0; row < TY; row += 2) {
      const float me = s_m[pl][row][lane], mo = s_m[pl][row + 1][lane];
      const float4 we = *reinterpret_cast<const float4*>(&s_ry[row][0]);
      const float4 wo = *reinterpret_cast<const float4*>(&s_ry[row + 1][0]);
      e0 = fmaf(we.x, me, e0); e1 = fmaf(we.y, me, e1); e2 = fmaf(we.z, me, e2);
      o0 = fmaf(wo.x, mo, o0); o1 = fmaf(wo.y, mo, o1); o2 = fmaf(wo.z, mo, o2);
    }
    gfloat* dst = Typ + (kb + pl) * W + kx;
    dst[0 * DW] = e0; dst[1 * DW] = e1; dst[2 * DW] = e2;
    dst[3 * DW] = o0; dst[4 * DW] = o1; dst[5 * DW] = o2;
  };

  PlaneIn cur, nxt;
  ring_load(kz0, std::integral_constant<int, 0>(), false);
  ring_load(kz0 + 1, std::integral_constant<int, 1>(), kz0 + 1 >= kz1);
  stage_load(kz0);
  load(cur, kz0);
  stage_store();
  __syncthreads();  // s_rx, s_rfx, first batch's factors
  for (int kb = kz0; kb < kz1; kb += KB) {
    const int n = min(KB, kz1 - kb);
    stage_load(kb + KB);  // next batch's factors (past the chunk: clamped, unused)
    // n >= 1: a loop that runs at least once, so that the compiler's count of
    // the loads issued after this batch's staging loads does not fall to zero
    // on a path that skips the planes (it would then wait for every load in
    // flight before storing the staged factors)
    int kl = 0;
    do {
      const int kz = kb + kl;
      // the next plane's loads go out before this plane's arithmetic (past
      // the chunk's end they re-read its last plane). Both planes of a pair
      // always run, the second one idle past an odd batch's end: no path
      // leaves a load in flight that the code after the loop does not
      // expect, so the next batch's first plane stays in flight through the
      // x and y sums and the barriers.
      load(nxt, min(kz + 1, kz1 - 1));
      plane(cur, kz, kl, std::integral_constant<int, 0>(), true);
      load(cur, min(kz + 2, kz1 - 1));
      plane(nxt, kz + 1, kl + 1, std::integral_constant<int, 1>(), kl + 1 < n);
      kl += 2;
    } while (kl < n);
    // the x sums read only this wave's rows of the mode tile: no block barrier
    __builtin_amdgcn_wave_barrier();
    xreduce(kb, n);
    block_barrier();  // y sums in LDS; the batch's factors no longer read
    ysum(kb, n);
    stage_store();
    block_barrier();  // LDS free for the next batch, its factors in place
  }
  // the ring's last loads (clamped planes) land before the block ends
  if (RING > 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  // z sums of this kz chunk
  gfloat* tzp = ARD_G(P.Tz) + (long long)bz * 6 * HW;
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    if (!rok[r] || !xok) continue;
#pragma unroll
    for (int q = 0; q < 6; ++q) tzp[q * HW + rof[r] + kx] = tz[r][q];
  }
  // recorder partials of this block
  if (PTS && nrec > 0) {
#pragma unroll
    for (int q = 0; q < NRC; ++q) {
      if (q < nrec) {
        float v = rec_acc[q];
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m);
        if (lane == 0) s_rec[wave][q] = v;
      }
    }
    __syncthreads();
    if (tid < nrec)
      P.rec_part[(long long)lb * NRC + tid] = s_rec[0][tid] + s_rec[1][tid] + s_rec[2][tid] + s_rec[3][tid];
  }
}

// ---------------------------------------------------------------------------
// Point sources and recorders of a partition past the modal pass's four of
// each (ARD/simulation.cpp:37-48 and ARD/recorder.cpp:26-38 set no limit).
//
// Sources (k_xsrc, before the modal pass): the update
//   M+ = a M - .999 M- + g Fh
// is linear in M-, so subtracting g dF / .999 from M- (which the pass then
// overwrites with M+) adds g dF to M+, and so to every face and recorder sum
// the pass forms from M+; dF = sum_s f_s fx_s(kx) fy_s(ky) fz_s(kz) is the
// extra sources' force spectrum, with their step samples from the value
// slots, and g is formed as the pass forms it.
// Recorders (k_xrec, after it): each one's sample is the sum over the new
// modes of M+ fx fy fz (the inverse transforms evaluated at its cell);
// per-block partials in FP32, summed over blocks in FP64 by k_brecord.
// A block walks kXRows mode rows (kz, ky), four per wave, lanes along kx.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_xsrc(const XsrcDesc* __restrict__ descs, const long long* __restrict__ starts,
                                              int ndesc, int par) {
  const int di = find_desc_ll(starts, ndesc, blockIdx.x);
  const XsrcDesc& X = descs[di];
  const int lb = int(blockIdx.x - starts[di]);
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = X.W, H = X.H, n = X.n;
  gfloat* __restrict__ Mp = ARD_G(X.M[par ^ 1]);
  const float rg = 1.f / 0.999f;
  for (int rr = 0; rr < kXRows / 4; ++rr) {
    const int row = lb * kXRows + wave * (kXRows / 4) + rr;  // kz * H + ky
    if (row >= X.D * H) break;
    const int kz = row / H, ky = row - kz * H;
    const float uz = ARD_CG(X.wz2)[kz], uy = ARD_CG(X.wy2)[ky];
    for (int kx0 = 0; kx0 < W; kx0 += 64) {
      const int kx = kx0 + lane, kxc = min(kx, W - 1);
      float f = 0.f;
      for (int s = 0; s < n; ++s) {
        const SrcTab& t = X.src[s];
        f = fmaf(*t.val * ARD_CG(t.fy)[ky] * ARD_CG(t.fz)[kz], ARD_CG(t.fx)[kxc], f);
      }
      // u summed in the modal pass's order: (ux + uz) + uy
      const float g = X.gscale * sinc2_poly((ARD_CG(X.wx2)[kxc] + uz) + uy);
      if (kx < W) {
        const long long o = (long long)row * X.Wp + kx;
        Mp[o] = fmaf(-g * f, rg, Mp[o]);
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_xrec(const XrecDesc* __restrict__ descs, const long long* __restrict__ starts,
                                              int ndesc, int par) {
  constexpr int G = 8;  // recorders per round over the block's rows
  __shared__ float s_part[4][G];
  const int di = find_desc_ll(starts, ndesc, blockIdx.x);
  const XrecDesc& X = descs[di];
  const int lb = int(blockIdx.x - starts[di]);
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = X.W, H = X.H, n = X.n;
  const gfloat* __restrict__ Mn = ARD_CG(X.M[par ^ 1]);
  for (int r0 = 0; r0 < n; r0 += G) {
    float acc[G];
#pragma unroll
    for (int q = 0; q < G; ++q) acc[q] = 0.f;
    for (int rr = 0; rr < kXRows / 4; ++rr) {
      const int row = lb * kXRows + wave * (kXRows / 4) + rr;
      if (row >= X.D * H) break;
      const int kz = row / H, ky = row - kz * H;
      float wyz[G];
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const RecTab& t = X.rec[min(r0 + q, n - 1)];
        wyz[q] = r0 + q < n ? ARD_CG(t.fy)[ky] * ARD_CG(t.fz)[kz] : 0.f;
      }
      for (int kx0 = 0; kx0 < W; kx0 += 64) {
        const int kx = kx0 + lane, kxc = min(kx, W - 1);
        const float m = kx < W ? Mn[(long long)row * X.Wp + kx] : 0.f;
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q] = fmaf(wyz[q] * ARD_CG(X.rec[min(r0 + q, n - 1)].fx)[kxc], m, acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      float v = acc[q];
#pragma unroll
      for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m);
      if (lane == 0) s_part[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < G && r0 + int(threadIdx.x) < n)
      X.part[(long long)lb * n + r0 + threadIdx.x] =
          s_part[0][threadIdx.x] + s_part[1][threadIdx.x] + s_part[2][threadIdx.x] + s_part[3][threadIdx.x];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Recorder samples from the block partials, then the step counter, then the
// next step's source samples: they overwrite their cells after this step's
// interface forcing and before the next update, where the reference sets
// them (ARD/simulation.cpp:306-312), one launch fewer per step.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_brecord(const BRecDesc* __restrict__ r, int n, float* __restrict__ trace,
                                                  int cap, int* __restrict__ d_step, const BSourceDesc* __restrict__ src,
                                                  int nsrc, float courant) {
  // one wave per recorder: its lanes sum the block partials in FP64, eight
  // loads in flight per lane
  const int t = *d_step;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = wave; i < n; i += 16) {
    const int nb = r[i].nblocks;
    const float* __restrict__ part = r[i].part;
    double s = 0.0;
    for (int b0 = 0; b0 < nb; b0 += 512) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int b = b0 + 64 * j + lane;
        v[j] = b < nb ? part[(long long)b * r[i].stride] : 0.f;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) s += double(v[j]);
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) s += __shfl_xor(s, m);
    if (lane == 0 && t < cap) trace[(long long)t * n + i] = float(s);
  }
  __syncthreads();
  if (threadIdx.x == 0) *d_step = t + 1;
  for (int i = int(threadIdx.x); i < nsrc; i += 1024) *src[i].dst = bsource_value(src[i], t + 1, courant);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int grid_of(long long total) { return int((total + 255) / 256); }

int launch_unfold(const UnfoldDesc* d, const long long* starts, int n, long long total, hipStream_t s) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_unfold, dim3(grid_of(total)), dim3(256), 0, s, d, starts, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_bsources(const BSourceDesc* d, int n, const int* d_step, float courant, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_bsources, dim3((n + 63) / 64), dim3(64), 0, s, d, n, d_step, courant);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_modal(int rpw, bool pts, int ring, const ModalDesc* d, const long long* starts, int n, long long blocks,
                 int par, hipStream_t s) {
  if (blocks <= 0) return 0;
  const dim3 g{unsigned(blocks)}, b{256};
  if (ring != 0 && rpw != 4) return -1;
#define ARD_MODAL_CASE(R, G)                                                                   \
  if (rpw == R && ring == G) {                                                                 \
    if (pts) hipLaunchKernelGGL((k_modal<R, true, G>), g, b, 0, s, d, starts, n, par);         \
    else hipLaunchKernelGGL((k_modal<R, false, G>), g, b, 0, s, d, starts, n, par);            \
    return hipGetLastError() == hipSuccess ? 0 : -2;                                           \
  }
  ARD_MODAL_CASE(2, 0)
  ARD_MODAL_CASE(4, 0)
  ARD_MODAL_CASE(4, 2)
  ARD_MODAL_CASE(6, 0)
  ARD_MODAL_CASE(8, 0)
#undef ARD_MODAL_CASE
  return -1;
}

int launch_xsrc(const XsrcDesc* d, const long long* starts, int n, long long blocks, int par, hipStream_t s) {
  if (blocks <= 0) return 0;
  hipLaunchKernelGGL(k_xsrc, dim3(unsigned(blocks)), dim3(256), 0, s, d, starts, n, par);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_xrec(const XrecDesc* d, const long long* starts, int n, long long blocks, int par, hipStream_t s) {
  if (blocks <= 0) return 0;
  hipLaunchKernelGGL(k_xrec, dim3(unsigned(blocks)), dim3(256), 0, s, d, starts, n, par);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_brecord(const BRecDesc* d, int n, float* trace, int cap, int* d_step, const BSourceDesc* src, int nsrc,
                   float courant, hipStream_t s) {
  hipLaunchKernelGGL(k_brecord, dim3(1), dim3(1024), 0, s, d, n, trace, cap, d_step, src, nsrc, courant);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace ard
