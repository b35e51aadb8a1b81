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

// ---------------------------------------------------------------------------
// Fold: face pair (x, n-1-x) -> sum / difference planes, x = 0..2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fold(const FoldDesc* __restrict__ d, const long long* __restrict__ starts,
                                              int n) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const int i = find_desc_ll(starts, n, e);
  const FoldDesc& F = d[i];
  const long long le = e - starts[i];
  if (le >= 3 * F.plane) return;
  const unsigned pl = unsigned(F.plane);  // planes hold < 2^31 cells
  const int q = int(unsigned(le) / pl);
  const long long o = unsigned(le) - unsigned(q) * pl;
  const gfloat* src = ARD_CG(F.src);
  gfloat* dst = ARD_G(F.dst);
  const float a = src[q * F.plane + o];
  const float b = src[(5 - q) * F.plane + o];
  dst[q * F.plane + o] = a + b;
  dst[(3 + q) * F.plane + o] = a - b;
}

// ---------------------------------------------------------------------------
// Unfold: sum the per-chunk partial reductions and turn the (even, odd) pair
// back into the face pair: x -> even + odd, n-1-x -> even - odd.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_unfold(const UnfoldDesc* __restrict__ d,
                                                const long long* __restrict__ starts, int n) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const int i = find_desc_ll(starts, n, e);
  const UnfoldDesc& U = d[i];
  const long long le = e - starts[i];
  if (le >= 3 * U.plane) return;
  const unsigned pl = unsigned(U.plane);  // planes hold < 2^31 cells
  const int q = int(unsigned(le) / pl);
  const long long o = unsigned(le) - unsigned(q) * pl;
  float ev = 0.f, od = 0.f;
  for (int c = 0; c < U.nchunk; ++c) {
    const gfloat* p = ARD_CG(U.part) + (long long)c * 6 * U.plane;
    ev += p[q * U.plane + o];
    od += p[(3 + q) * U.plane + o];
  }
  ARD_G(U.dst)[q * U.plane + o] = ev + od;
  ARD_G(U.dst)[(5 - q) * U.plane + o] = ev - od;
}

// ---------------------------------------------------------------------------
// Source samples: GaussianSource (ARD/gaussian_source.cpp:16-20), a waveform,
// or a constant (a force set through set_force at an interior cell). Shell
// cells receive the value in the face slab; interior cells in a value slot
// read by k_modal.
// ---------------------------------------------------------------------------
__global__ void k_bsources(const BSourceDesc* __restrict__ s, int n, const int* __restrict__ d_step,
                           float courant) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = *d_step;
  float v;
  if (s[i].kind == 0) {
    const float x = 3.14159265358979f * ((2.f * courant * float(t)) / 6.f - 2.f);
    v = 1e9f * expf(-(x * x));
  } else if (s[i].kind == 1) {
    v = t < s[i].n ? s[i].wave[t] : 0.f;
  } else {
    v = s[i].value;
  }
  *s[i].dst = v;
}

// ---------------------------------------------------------------------------
// The modal pass. Block = 4 waves; lane <-> kx (64 per block), wave w owns
// rows ky = ky0 + w + 4 r (r < RPW), the block walks kz over its chunk.
// Per mode:   Fh = sum_i Ex[kx][i] Gx[par kx][i][kz][ky] + (y, z alike) + sources
//             M+ = a M - b M- + g Fh           (stored over M-)
// and reduces M+ onto the face planes and recorders:
//   Tx[par kx][i][kz][ky] += Rx[kx][i] M+   (across lanes)
//   Ty[par ky][j][kz][kx] += Ry[ky][j] M+   (across rows and waves)
//   Tz[par kz][l][ky][kx] += Rz[kz][l] M+   (along the kz walk)
// Partial sums go to per-chunk slots, summed by k_unfold.
// ---------------------------------------------------------------------------
// sum over the 8 lanes of one parity in each 16-lane row (DPP, no LDS):
// xor 2 inside the quad, then rotations by 4 and 8 inside the row
__device__ __forceinline__ float row_parity_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x124, 0xF, 0xF, false)); // row_ror:4
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xF, 0xF, false)); // row_ror:8
  return v;
}

// sin(x) for 0 <= x <= 1.3 (half the largest mode phase c dt pi sqrt(3) / h at
// Courant 0.43): odd Taylor series to x^11, truncation error < x^13/13! < 6e-9.
__device__ __forceinline__ float sin_small(float x) {
  const float x2 = x * x;
  float p = -2.5052108e-8f;           // -1/11!
  p = fmaf(p, x2, 2.7557319e-6f);     //  1/9!
  p = fmaf(p, x2, -1.9841270e-4f);    // -1/7!
  p = fmaf(p, x2, 8.3333333e-3f);     //  1/5!
  p = fmaf(p, x2, -1.6666667e-1f);    // -1/3!
  p = fmaf(p, x2, 1.0f);
  return x * p;
}

template <int RPW, bool PTS>
__global__ __launch_bounds__(256) void k_modal(const ModalDesc* __restrict__ descs,
                                               const long long* __restrict__ starts, int ndesc, int par) {
  constexpr int TY = 4 * RPW;  // rows of the block tile
  constexpr int KB = 8;        // kz planes accumulated in LDS between flushes
  __shared__ float s_ty[KB][2][3][64];    // [plane][ky parity][j][lane]
  __shared__ float s_tx[KB][TY][2][3];    // [plane][row][kx parity][i]
  __shared__ float s_rec[4][ARD_MAX_REC_PER_PART];
  __shared__ float s_gz[RPW][6][256];     // thread-private: z-face force spectrum of this thread's rows
  const int bid = blockIdx.x;
  const int di = find_desc_ll(starts, ndesc, bid);
  const ModalDesc& P = descs[di];  // fields used in the loops are copied to locals below
  const int lb = bid - int(starts[di]);
  const int bx = lb % P.nbx;
  const int by = (lb / P.nbx) % P.nby;
  const int bz = lb / (P.nbx * P.nby);
  const int W = P.W, H = P.H, D = P.D;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int kx = bx * 64 + lane;
  const bool xok = kx < W;
  const int kxc = xok ? kx : W - 1;
  const int px = kx & 1;
  const int kyb = by * TY;           // first row of the block
  const int ky0 = kyb + wave;        // rows of this wave: ky0 + 4 r
  const int py = ky0 & 1;            // shared by every row of this wave
  const int kz0 = bz * P.kz_chunk;   // kz_chunk is even, so kz0 is even
  const int kz1 = min(D, kz0 + P.kz_chunk);
  const long long HW = (long long)H * W;
  const long long DH = (long long)D * H, DW = (long long)D * W;
  const gfloat* __restrict__ Gx = ARD_CG(P.Gx);
  const gfloat* __restrict__ Gy = ARD_CG(P.Gy);
  const gfloat* __restrict__ Ez = ARD_CG(P.Ez);
  const gfloat* __restrict__ Rz = ARD_CG(P.Rz);
  const gfloat* __restrict__ Ey = ARD_CG(P.Ey);
  const gfloat* __restrict__ Ry = ARD_CG(P.Ry);
  const gfloat* __restrict__ Wz2 = ARD_CG(P.wz2);
  const gfloat* __restrict__ Wy2 = ARD_CG(P.wy2);
  const gfloat* __restrict__ Mc = ARD_CG(P.M[par]);
  gfloat* __restrict__ Mp = ARD_G(P.M[par ^ 1]);
  const float hdt = 0.5f * P.dt;
  const float gsc = P.gscale;  // .999 * 4 * s_f * s_i

  float ex[3], rx[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ex[i] = ARD_CG(P.Ex)[kxc * 3 + i];
    rx[i] = xok ? ARD_CG(P.Rx)[kxc * 3 + i] : 0.f;
  }
  const float wx2 = ARD_CG(P.wx2)[kxc];
  float tz[RPW][6];
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int kyc = min(ky0 + 4 * r, H - 1);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      s_gz[r][q][tid] = ARD_CG(P.Gz)[q * HW + (long long)kyc * W + kxc];
      tz[r][q] = 0.f;
    }
  }
  const int nsrc = PTS ? min(P.nsrc, ARD_MAX_SRC_PER_PART) : 0;
  const int nrec = PTS ? P.nrec : 0;
  float sfx[ARD_MAX_SRC_PER_PART], rfx[ARD_MAX_REC_PER_PART], rec_acc[ARD_MAX_REC_PER_PART];
#pragma unroll
  for (int q = 0; q < ARD_MAX_SRC_PER_PART; ++q) sfx[q] = PTS && q < nsrc ? P.src[q].fx[kxc] * *P.src[q].val : 0.f;
#pragma unroll
  for (int q = 0; q < ARD_MAX_REC_PER_PART; ++q) {
    rfx[q] = PTS && q < nrec && xok ? P.rec[q].fx[kxc] : 0.f;
    rec_acc[q] = 0.f;
  }
  for (int e = tid; e < KB * 2 * 3 * 64; e += 256) (&s_ty[0][0][0][0])[e] = 0.f;
  for (int e = tid; e < KB * TY * 6; e += 256) (&s_tx[0][0][0][0])[e] = 0.f;
  __syncthreads();

  auto body = [&](const int kz, const int kl, auto pzc) {
    constexpr int PZ = decltype(pzc)::value;
    const float ez0 = Ez[kz * 3 + 0], ez1 = Ez[kz * 3 + 1], ez2 = Ez[kz * 3 + 2];
    const float rz0 = Rz[kz * 3 + 0], rz1 = Rz[kz * 3 + 1], rz2 = Rz[kz * 3 + 2];
    const float wxz2 = wx2 + Wz2[kz];
    float gy[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) gy[j] = Gy[(py * 3 + j) * DW + (long long)kz * W + kxc];
    float sz[ARD_MAX_SRC_PER_PART], rzp[ARD_MAX_REC_PER_PART];
#pragma unroll
    for (int q = 0; q < ARD_MAX_SRC_PER_PART; ++q) sz[q] = PTS && q < nsrc ? P.src[q].fz[kz] : 0.f;
#pragma unroll
    for (int q = 0; q < ARD_MAX_REC_PER_PART; ++q) rzp[q] = PTS && q < nrec ? P.rec[q].fz[kz] : 0.f;
    float ty[3] = {0.f, 0.f, 0.f};
    // loads of all rows first, at clamped (always valid) addresses, so that
    // they are in flight together; rows and lanes outside the partition are
    // masked afterwards
    float mcv[RPW], mpv[RPW], gxs[RPW][3], gxd[RPW][3], eyv[RPW][3], ryv[RPW][3], wyv[RPW];
    long long ov[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int kyc = min(ky0 + 4 * r, H - 1);
      const long long gxo = (long long)kz * H + kyc;
      ov[r] = (long long)kz * HW + (long long)kyc * W + kxc;
      mcv[r] = Mc[ov[r]];
      mpv[r] = Mp[ov[r]];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        gxs[r][i] = Gx[i * DH + gxo];
        gxd[r][i] = Gx[(3 + i) * DH + gxo];
        eyv[r][i] = Ey[kyc * 3 + i];
        ryv[r][i] = Ry[kyc * 3 + i];
      }
      wyv[r] = Wy2[kyc];
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int ky = ky0 + 4 * r;
      const bool rok = ky < H;  // uniform across the wave
      const bool ok = rok && xok;
      // modal coefficients (ARD/dct_partition.cpp:24-33,68) from the separable
      // w^2 = wx2 + wy2 + wz2 (1-based wavenumbers):
      //   a = .999 * 2 cos(w dt) = .999 * 2 (1 - 2 s^2),  g = .999 * 4 s^2 / w^2,  s = sin(w dt / 2)
      const float w2 = wxz2 + wyv[r];
      const float sn = sin_small(hdt * __builtin_amdgcn_sqrtf(w2));
      const float s2 = sn * sn;
      const float a = fmaf(-3.996f, s2, 1.998f);
      const float g = gsc * s2 * __builtin_amdgcn_rcpf(w2);
      float fh = 0.f;
#pragma unroll
      for (int i = 0; i < 3; ++i) fh = fmaf(ex[i], px ? gxd[r][i] : gxs[r][i], fh);
      fh = fmaf(eyv[r][0], gy[0], fh);
      fh = fmaf(eyv[r][1], gy[1], fh);
      fh = fmaf(eyv[r][2], gy[2], fh);
      fh = fmaf(ez0, s_gz[r][PZ * 3 + 0][tid], fh);
      fh = fmaf(ez1, s_gz[r][PZ * 3 + 1][tid], fh);
      fh = fmaf(ez2, s_gz[r][PZ * 3 + 2][tid], fh);
      if (PTS) {
        const int kyc = min(ky, H - 1);
#pragma unroll
        for (int q = 0; q < ARD_MAX_SRC_PER_PART; ++q)
          if (q < nsrc) fh = fmaf(sfx[q], P.src[q].fy[kyc] * sz[q], fh);
      }
      const float mn = ok ? (a * mcv[r] - 0.999f * mpv[r] + g * fh) : 0.f;
      if (ok) Mp[ov[r]] = mn;
      tz[r][PZ * 3 + 0] = fmaf(rz0, mn, tz[r][PZ * 3 + 0]);
      tz[r][PZ * 3 + 1] = fmaf(rz1, mn, tz[r][PZ * 3 + 1]);
      tz[r][PZ * 3 + 2] = fmaf(rz2, mn, tz[r][PZ * 3 + 2]);
      ty[0] = fmaf(ryv[r][0], mn, ty[0]);
      ty[1] = fmaf(ryv[r][1], mn, ty[1]);
      ty[2] = fmaf(ryv[r][2], mn, ty[2]);
      if (PTS) {
        const int kyc = min(ky, H - 1);
#pragma unroll
        for (int q = 0; q < ARD_MAX_REC_PER_PART; ++q)
          if (q < nrec) rec_acc[q] = fmaf(rfx[q] * P.rec[q].fy[kyc] * rzp[q], mn, rec_acc[q]);
      }
      // x reduction: parity sums inside each 16-lane row, then the four rows
      // meet in LDS
      const float v0 = row_parity_sum(rx[0] * mn);
      const float v1 = row_parity_sum(rx[1] * mn);
      const float v2 = row_parity_sum(rx[2] * mn);
      if (rok && (lane & 14) == 0) {
        float* t = s_tx[kl][wave + 4 * r][lane & 1];
        atomicAdd(t + 0, v0);
        atomicAdd(t + 1, v1);
        atomicAdd(t + 2, v2);
      }
    }
    // y reduction: waves 0,2 (even ky) and 1,3 (odd ky) meet in LDS
#pragma unroll
    for (int j = 0; j < 3; ++j) atomicAdd(&s_ty[kl][py][j][lane], ty[j]);
  };
  // write the LDS accumulators of planes kb .. kb+n-1 to the partial
  // reductions and clear them
  auto flush = [&](const int kb, const int n) {
    __syncthreads();
    gfloat* __restrict__ Txp = ARD_G(P.Tx) + (long long)bx * 6 * DH;
    gfloat* __restrict__ Typ = ARD_G(P.Ty) + (long long)by * 6 * DW;
    for (int e = tid; e < n * TY * 6; e += 256) {
      const int q = e % 6, row = (e / 6) % TY, kl = e / (6 * TY);  // q = parity * 3 + i
      const int ky = kyb + row;
      float* src = &s_tx[kl][row][q / 3][q % 3];
      if (ky < H) Txp[q * DH + (long long)(kb + kl) * H + ky] = *src;
      *src = 0.f;
    }
    for (int e = tid; e < n * 6 * 64; e += 256) {
      const int l = e & 63, q = (e >> 6) % 6, kl = e / 384;
      const int kxx = bx * 64 + l;
      float* src = &s_ty[kl][q / 3][q % 3][l];
      if (kxx < W) Typ[q * DW + (long long)(kb + kl) * W + kxx] = *src;
      *src = 0.f;
    }
    __syncthreads();
  };
  int kb = kz0;
  for (int kz = kz0; kz < kz1; kz += 2) {
    body(kz, kz - kb, std::integral_constant<int, 0>());
    if (kz + 1 < kz1) body(kz + 1, kz + 1 - kb, std::integral_constant<int, 1>());
    const int next = kz + 2;
    if (next - kb >= KB || next >= kz1) {
      flush(kb, min(next, kz1) - kb);
      kb = next;
    }
  }
  // z reduction partials of this kz chunk
  gfloat* tzp = ARD_G(P.Tz) + (long long)bz * 6 * HW;
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int ky = ky0 + 4 * r;
    if (ky >= H || !xok) continue;
#pragma unroll
    for (int q = 0; q < 6; ++q) tzp[q * HW + (long long)ky * W + kx] = tz[r][q];
  }
  // recorder partials of this block
  if (PTS && nrec > 0) {
#pragma unroll
    for (int q = 0; q < ARD_MAX_REC_PER_PART; ++q) {
      if (q < nrec) {
        float v = rec_acc[q];
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m);
        if (lane == 0) s_rec[wave][q] = v;
      }
    }
    __syncthreads();
    if (tid < nrec)
      P.rec_part[(long long)lb * ARD_MAX_REC_PER_PART + tid] =
          s_rec[0][tid] + s_rec[1][tid] + s_rec[2][tid] + s_rec[3][tid];
  }
}

// ---------------------------------------------------------------------------
// Recorder samples from the block partials, then the step counter.
// ---------------------------------------------------------------------------
__global__ void k_brecord(const BRecDesc* __restrict__ r, int n, float* __restrict__ trace, int cap,
                          int* __restrict__ d_step) {
  const int t = *d_step;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < r[i].nblocks; ++b) s += r[i].part[(long long)b * ARD_MAX_REC_PER_PART];
    if (t < cap) trace[(long long)t * n + i] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) *d_step = t + 1;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int grid_of(long long total) { return int((total + 255) / 256); }

int launch_fold(const FoldDesc* d, const long long* starts, int n, long long total, hipStream_t s) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_fold, dim3(grid_of(total)), dim3(256), 0, s, d, starts, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

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

int launch_modal(int rpw, bool pts, const ModalDesc* d, const long long* starts, int n, long long blocks, int par,
                 hipStream_t s) {
  if (blocks <= 0) return 0;
  const dim3 g{unsigned(blocks)}, b{256};
#define ARD_MODAL_CASE(R)                                                                      \
  case R:                                                                                      \
    if (pts) hipLaunchKernelGGL((k_modal<R, true>), g, b, 0, s, d, starts, n, par);            \
    else hipLaunchKernelGGL((k_modal<R, false>), g, b, 0, s, d, starts, n, par);               \
    break;
  switch (rpw) {
    ARD_MODAL_CASE(4)
    ARD_MODAL_CASE(5)
    ARD_MODAL_CASE(8)
    default:
      return -1;
  }
#undef ARD_MODAL_CASE
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_brecord(const BRecDesc* d, int n, float* trace, int cap, int* d_step, hipStream_t s) {
  hipLaunchKernelGGL(k_brecord, dim3(1), dim3(256), 0, s, d, n, trace, cap, d_step);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace ard
