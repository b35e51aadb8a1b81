// FFT-style per-axis DCT passes (transform variant 4): the O(L log L)
// algorithm class of the reference's FFTW plans (ARD/dct_volume.cpp:17,19:
// fftwf_plan_r2r_3d with FFTW_REDFT10 / FFTW_REDFT01), written for gfx950
// from the published methods, not from FFTW's or VkFFT's code:
//   * Makhoul's reordering (IEEE Trans. ASSP 28(1), 1980): a line x of length
//     L becomes v[j] = x[2j] (j < ceil(L/2)), v[L-1-j] = x[2j+1]; with
//     V = DFT_L(v), REDFT10(x)_k = 2 Re(e^{-i pi k / 2L} V_k). REDFT01(X) is
//     the real part of the unnormalised inverse DFT of
//     e^{i pi k / 2L} (X_k - i X_{L-k}) (X_L = 0), read back through the same
//     reordering.
//   * two real lines per complex line: A in the real part, B in the
//     imaginary part; A_k = (Z_k + conj Z_{L-k}) / 2, B_k = (Z_k - conj Z_{L-k}) / 2i
//     (forward); the inverse's outputs are real, so A and B come back as the
//     real and imaginary parts.
//   * the DFT by iterative in-place mixed-radix Cooley-Tukey (decimation in
//     time) in LDS: the input sits in digit-reversed order (composed with
//     Makhoul's reordering into one load table), each stage's butterflies
//     read and rewrite their own R slots, a barrier between stages; radix 4,
//     2, 3 and 5 butterflies, other primes up to 13 as a direct DFT.
// The pass is a drop-in for the dense GEMM pass: it reads the same GemmDesc
// (lines of length L = K with element stride bK, line offsets b1/b2, output
// element stride cM, line offsets c1/c2, alpha, and the plain form's modal
// epilogue C = a cur - b C + g alpha X).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ard_ctx_impl.h"
#include "ard_kernels.h"

namespace ard {

namespace {

struct cf {
  float x, y;
};
__device__ __forceinline__ cf cmul(cf a, cf b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }

// a / d and a % d for 0 <= a < 2^22 and 0 < d, by a float reciprocal and
// one correction step
__device__ __forceinline__ int qdiv(int a, int d, float rd, int* rem) {
  int q = int(float(a) * rd);
  int r = a - q * d;
  if (r < 0) { --q; r += d; }
  else if (r >= d) { ++q; r -= d; }
  *rem = r;
  return q;
}

// One stage of one wave: radix R butterflies over spans Ns R, in place in
// the wave's lines buf [lw][L] (no other wave touches them).
template <int R>
__device__ __forceinline__ void stage(cf* __restrict__ buf, const cf* __restrict__ tw, int L, int lw, int Ns) {
  const int T = L / R;
  const float rN = 1.f / float(Ns);
  const int step = L / (Ns * R);  // twiddle W_{Ns R}^{k r} = tw[k r step]
  for (int c = 0; c < lw; ++c)
  for (int bb = threadIdx.x & 63; bb < T; bb += 64) {
    int k;
    const int grp = qdiv(bb, Ns, rN, &k);
    cf* x = buf + c * L + grp * Ns * R + k;
    cf u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = x[r * Ns];
#pragma unroll
    for (int r = 1; r < R; ++r) u[r] = cmul(u[r], tw[k * r * step]);
    if (R == 2) {
      x[0] = cadd(u[0], u[1]);
      x[Ns] = csub(u[0], u[1]);
    } else if (R == 4) {
      const cf s02 = cadd(u[0], u[2]), d02 = csub(u[0], u[2]);
      const cf s13 = cadd(u[1], u[3]), d13 = csub(u[1], u[3]);
      // W_4 = -i: X1 = d02 - i d13, X3 = d02 + i d13
      x[0] = cadd(s02, s13);
      x[2 * Ns] = csub(s02, s13);
      x[Ns] = {d02.x + d13.y, d02.y - d13.x};
      x[3 * Ns] = {d02.x - d13.y, d02.y + d13.x};
    } else if (R == 3) {
      // W_3 = -1/2 - i sqrt(3)/2
      const float h3 = 0.86602540378443864676f;
      const cf t = cadd(u[1], u[2]);
      const cf d = {(u[1].x - u[2].x) * h3, (u[1].y - u[2].y) * h3};
      const cf m = {u[0].x - 0.5f * t.x, u[0].y - 0.5f * t.y};
      x[0] = cadd(u[0], t);
      x[Ns] = {m.x + d.y, m.y - d.x};
      x[2 * Ns] = {m.x - d.y, m.y + d.x};
    } else if (R == 5) {
      const float k1 = 0.30901699437494742410f, k2 = -0.80901699437494742410f;  // cos(2 pi/5), cos(4 pi/5)
      const float q1 = 0.95105651629515357212f, q2 = 0.58778525229247312917f;   // sin(2 pi/5), sin(4 pi/5)
      const cf t1 = cadd(u[1], u[4]), t2 = cadd(u[2], u[3]), t3 = csub(u[1], u[4]), t4 = csub(u[2], u[3]);
      const cf b1 = {u[0].x + k1 * t1.x + k2 * t2.x, u[0].y + k1 * t1.y + k2 * t2.y};
      const cf b2 = {u[0].x + k2 * t1.x + k1 * t2.x, u[0].y + k2 * t1.y + k1 * t2.y};
      const cf d1 = {q1 * t3.x + q2 * t4.x, q1 * t3.y + q2 * t4.y};
      const cf d2 = {q2 * t3.x - q1 * t4.x, q2 * t3.y - q1 * t4.y};
      x[0] = {u[0].x + t1.x + t2.x, u[0].y + t1.y + t2.y};
      // X1 = b1 - i d1, X4 = b1 + i d1, X2 = b2 - i d2, X3 = b2 + i d2
      x[Ns] = {b1.x + d1.y, b1.y - d1.x};
      x[4 * Ns] = {b1.x - d1.y, b1.y + d1.x};
      x[2 * Ns] = {b2.x + d2.y, b2.y - d2.x};
      x[3 * Ns] = {b2.x - d2.y, b2.y + d2.x};
    } else {
      // direct DFT of a prime length: X_q = sum_r u_r W_R^{r q}, W_R^{e} = tw[e (L / R)]
      const int sR = L / R;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        cf a = u[0];
#pragma unroll
        for (int r = 1; r < R; ++r) a = cadd(a, cmul(u[r], tw[((r * q) % R) * sR]));
        x[q * Ns] = a;
      }
    }
  }
}

}  // namespace

// Plan of one length L: radices, digit-reversed positions, tables. Layout of
// the device table (4-byte words): tw[L] complex (W_L^e, e < L), mk[L]
// complex (e^{-i pi k / 2L}), ld[L] (LDS slot of memory element i of a line
// for the forward pass: the digit-reversed position of v's index perminv(i)),
// pos[L] (digit-reversed slot of natural index k), pinv[L] (v index j of
// memory element i: j = i/2 for even i, L-1-(i-1)/2 for odd i).
bool fft_factor(int L, std::vector<int>& rad) {
  rad.clear();
  int m = L;
  while (m % 4 == 0) { rad.push_back(4); m /= 4; }
  while (m % 2 == 0) { rad.push_back(2); m /= 2; }
  for (int p : {3, 5, 7, 11, 13})
    while (m % p == 0) { rad.push_back(p); m /= p; }
  return m == 1 && L >= 2 && L <= 1024 && rad.size() <= sizeof(GemmDesc::rad);
}

// Complex lines per tile: 4 waves, each transforming lw whole lines alone;
// lw is the largest power of two with lw L <= 1024 (at most 16), so the
// tile's 2 tlc = 8 lw real lines divide the block's 256 threads.
int fft_lines(int L) {
  int lw = 1;
  while (lw < 16 && 2 * lw * L <= 1024) lw *= 2;
  return 4 * lw;
}

std::vector<float> fft_plan_table(int L, const std::vector<int>& rad) {
  const double pi = 3.14159265358979323846;
  std::vector<float> t(size_t(7) * L);
  float* tw = t.data();
  float* mk = tw + 2 * L;
  int* ld = reinterpret_cast<int*>(mk + 2 * L);
  int* pos = ld + L;
  int* pinv = pos + L;
  for (int e = 0; e < L; ++e) {
    tw[2 * e] = float(std::cos(2 * pi * e / L));
    tw[2 * e + 1] = float(-std::sin(2 * pi * e / L));
    mk[2 * e] = float(std::cos(pi * e / (2.0 * L)));
    mk[2 * e + 1] = float(-std::sin(pi * e / (2.0 * L)));
  }
  for (int n = 0; n < L; ++n) {  // digits of n, least significant in base rad[0]
    int q = n, span = L, p = 0;
    for (int r : rad) {
      span /= r;
      p += (q % r) * span;
      q /= r;
    }
    pos[n] = p;
  }
  for (int i = 0; i < L; ++i) {
    pinv[i] = (i % 2 == 0) ? i / 2 : L - 1 - (i - 1) / 2;
    ld[i] = pos[pinv[i]];
  }
  return t;
}

// Dynamic LDS of a descriptor: [tlc][L] complex, the twiddles and Makhoul
// factors (complex), the digit-reversed slots (int, padded to 8 bytes) and the
// tile's line offsets (2 x 2 tlc 64-bit): at most about 50 KB.
size_t fft_lds_bytes(const GemmDesc& g) {
  return size_t(g.tlc + 2) * g.K * 8 + size_t((g.K + 1) / 2) * 8 + size_t(4 * g.tlc) * 8;
}

template <bool B_K_CONTIG, bool C_M_CONTIG, int EP>
__global__ __launch_bounds__(256) void k_dct_fft(const GemmDesc* __restrict__ descs, int ndesc) {
  extern __shared__ cf sm[];
  const int tile = xcd_block(int(blockIdx.x), int(gridDim.x));
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmDesc& g = descs[lo];
  const int L = g.K, N = g.N, N1 = g.N1, tlc = g.tlc, tl = 2 * tlc;
  const int n0 = (tile - g.tile0) * tl;
  const int tid = threadIdx.x;
  cf* __restrict__ buf = sm;
  cf* __restrict__ tw = sm + tlc * L;
  cf* __restrict__ mk = tw + L;
  int* __restrict__ pos = reinterpret_cast<int*>(mk + L);
  const cf* __restrict__ ptw = reinterpret_cast<const cf*>(g.plan);
  const int* __restrict__ ppos = reinterpret_cast<const int*>(ptw + 2 * L) + L;
  const gfloat* __restrict__ B = ARD_CG(g.B);
  for (int e = tid; e < L; e += 256) {
    tw[e] = ptw[e];
    mk[e] = ptw[L + e];
    pos[e] = ppos[e];
  }
  // per-line input and output offsets of the tile (lines past N: -1)
  long long* __restrict__ lin = reinterpret_cast<long long*>(pos + 2 * ((L + 1) / 2));  // 8-byte aligned
  long long* __restrict__ lout = lin + tl;
  if (tid < tl) {
    const int n = n0 + tid;
    lin[tid] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : -1;
    lout[tid] = n < N ? (long long)(n % N1) * g.c1 + (long long)(n / N1) * g.c2 : -1;
  }
  // v index of memory element i under Makhoul's reordering
  auto pinv = [&](int i) { return (i & 1) ? L - 1 - (i >> 1) : (i >> 1); };
  __syncthreads();  // tables in LDS
  // the tile's input, U elements per thread in flight: along each line when it
  // is contiguous (line ll, elements tid, tid + 256, ...), else across the
  // tile's lines (line tid % tl, elements tid / tl + (256 / tl) u)
  constexpr int U = 8;
  const int gpl = 256 / tl;  // B_K_CONTIG == false: elements per sweep
  const int myl = tid % tl, myi = tid / tl;
  // contiguous lines shorter than a sweep go flat over the tile (e = 256 sw + tid)
  const bool flat = L < 256;
  const float rL = 1.f / float(L);
  const int nsweep = B_K_CONTIG ? (flat ? (tl * L + 255) / 256 : tl * ((L + 255) / 256)) : (L + gpl - 1) / gpl;
  for (int s0 = 0; s0 < nsweep; s0 += U) {
    float v[U];
    int slot[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int sw = s0 + u;
      int ll, i;
      if (B_K_CONTIG && flat) {
        ll = qdiv(min(sw * 256 + tid, tl * L - 1), L, rL, &i);
      } else if (B_K_CONTIG) {
        const int per = (L + 255) / 256;
        ll = sw / per;
        i = (sw - ll * per) * 256 + tid;
      } else {
        ll = myl;
        i = myi + gpl * sw;
      }
      const bool ok = sw < nsweep && i < L && (!(B_K_CONTIG && flat) || sw * 256 + tid < tl * L);
      const long long b0 = ok ? lin[ll] : -1;
      v[u] = b0 >= 0 ? B[b0 + (long long)i * g.bK] : 0.f;
      // forward: the digit-reversed slot of v index pinv(i); inverse: natural order
      slot[u] = ok ? (ll % tlc) * L * 2 + 2 * (g.inv ? i : pos[pinv(i)]) + ll / tlc : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (slot[u] >= 0) reinterpret_cast<float*>(buf)[slot[u]] = v[u];
  }
  if (g.inv) {
    // Z_k = V_A[k] + i V_B[k], V[k] = e^{i pi k/2L} (X_k - i X_{L-k}),
    // conjugated, to the digit-reversed slots
    __syncthreads();
    constexpr int MAXP = 9;  // pairs per thread: tlc (L/2 + 1) <= 2048 + 64 < 2304
    const int half = L / 2 + 1, npair = tlc * half;
    const float rh = 1.f / float(half);
    cf z0[MAXP], z1[MAXP];
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
      const int pr = tid + 256 * p;
      if (pr < npair) {
        int k;
        const int c = qdiv(pr, half, rh, &k);
        const int kk = (L - k) % L;  // partner
        const cf xk = buf[c * L + k];
        const cf xm = k ? buf[c * L + kk] : cf{0.f, 0.f};  // X_{L-k}; X_L = 0
        const cf w = mk[k];                                  // (cos, -sin) of pi k / 2L
        const float ec = w.x, es = -w.y;
        // line A: a1 = X_A[k] = xk.x, a2 = X_A[L-k] = xm.x; likewise B from .y
        const cf va = {ec * xk.x + es * xm.x, es * xk.x - ec * xm.x};
        const cf vb = {ec * xk.y + es * xm.y, es * xk.y - ec * xm.y};
        z0[p] = {va.x - vb.y, -(va.y + vb.x)};  // conj(Z_k)
        if (k && kk != k) {                      // Z_{L-k}: roles of X_k and X_{L-k} swap
          const cf w2 = mk[kk];
          const float fc = w2.x, fs = -w2.y;
          const cf ua = {fc * xm.x + fs * xk.x, fs * xm.x - fc * xk.x};
          const cf ub = {fc * xm.y + fs * xk.y, fs * xm.y - fc * xk.y};
          z1[p] = {ua.x - ub.y, -(ua.y + ub.x)};
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
      const int pr = tid + 256 * p;
      if (pr < npair) {
        int k;
        const int c = qdiv(pr, half, rh, &k);
        const int kk = (L - k) % L;
        buf[c * L + pos[k]] = z0[p];
        if (k && kk != k) buf[c * L + pos[kk]] = z1[p];
      }
    }
  }
  __syncthreads();
  // stages, innermost radix (last in the list) first; wave w owns lines
  // [w lw, (w + 1) lw): the LDS operations of one wave execute in order, so a
  // stage needs only a compiler fence before the next, no block barrier
  {
    const int lw = tlc >> 2;
    cf* __restrict__ wb = buf + (tid >> 6) * lw * L;
    int Ns = 1;
    for (int s = g.nrad - 1; s >= 0; --s) {
      const int R = g.rad[s];
      switch (R) {
        case 2: stage<2>(wb, tw, L, lw, Ns); break;
        case 3: stage<3>(wb, tw, L, lw, Ns); break;
        case 4: stage<4>(wb, tw, L, lw, Ns); break;
        case 5: stage<5>(wb, tw, L, lw, Ns); break;
        case 7: stage<7>(wb, tw, L, lw, Ns); break;
        case 11: stage<11>(wb, tw, L, lw, Ns); break;
        default: stage<13>(wb, tw, L, lw, Ns); break;
      }
      Ns *= R;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  // output in coalesced order: along k when the output is contiguous along
  // it, across lines otherwise; U outputs per thread in flight
  gfloat* __restrict__ C = ARD_G(g.C);
  const int osweep = C_M_CONTIG ? (flat ? (tl * L + 255) / 256 : tl * ((L + 255) / 256)) : (L + gpl - 1) / gpl;
  for (int s0 = 0; s0 < osweep; s0 += U) {
    float v[U];
    long long o[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int sw = s0 + u;
      int ll, k;
      if (C_M_CONTIG && flat) {
        ll = qdiv(min(sw * 256 + tid, tl * L - 1), L, rL, &k);
      } else if (C_M_CONTIG) {
        const int per = (L + 255) / 256;
        ll = sw / per;
        k = (sw - ll * per) * 256 + tid;
      } else {
        ll = myl;
        k = myi + gpl * sw;
      }
      const bool ok = sw < osweep && k < L && (!(C_M_CONTIG && flat) || sw * 256 + tid < tl * L);
      const long long c0 = ok ? lout[ll] : -1;
      o[u] = c0 >= 0 ? (long long)k * g.cM + c0 : -1;
      const int c = ok ? ll % tlc : 0, part = ll / tlc;
      const int kk = ok ? k : 0;
      if (!g.inv) {
        const cf zk = buf[c * L + kk], zm = buf[c * L + (kk ? L - kk : 0)];
        // A_k = (Z_k + conj Z_{L-k}) / 2, B_k = (Z_k - conj Z_{L-k}) / 2i
        const cf q = part == 0 ? cf{0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)}
                               : cf{0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x)};
        const cf w = mk[kk];
        v[u] = 2.f * (q.x * w.x - q.y * w.y);
      } else {
        // memory element k is v index pinv(k); z = conj(DFT(conj Z)): A = Re, B = -Im
        const cf y = buf[c * L + pinv(kk)];
        v[u] = part == 0 ? y.x : -y.y;
      }
    }
    if (EP == 0) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (o[u] >= 0) C[o[u]] = g.alpha * v[u];
    } else {
      float a[U], cu[U], pr[U], gg[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long q = o[u] >= 0 ? o[u] : 0;
        a[u] = ARD_CG(g.ep_a)[q];
        cu[u] = ARD_CG(g.ep_cur)[q];
        pr[u] = C[q];
        gg[u] = ARD_CG(g.ep_g)[q];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (o[u] >= 0) C[o[u]] = a[u] * cu[u] - g.ep_b * pr[u] + gg[u] * (g.alpha * v[u]);
    }
  }
}

int launch_fft(const GemmDesc* d_descs, int ndesc, int total_tiles, bool b_k_contig, bool c_m_contig, int ep,
               size_t lds_bytes, hipStream_t s) {
  if (total_tiles <= 0) return 0;
  const dim3 grid(total_tiles), block(256);
#define ARD_FFT_CASE(BK, CM, EPV)                                                                     \
  if (b_k_contig == BK && c_m_contig == CM && ep == EPV) {                                           \
    hipLaunchKernelGGL((k_dct_fft<BK, CM, EPV>), grid, block, lds_bytes, s, d_descs, ndesc);        \
    return hipGetLastError() == hipSuccess ? 0 : -2;                                                  \
  }
  ARD_FFT_CASE(true, true, 0)
  ARD_FFT_CASE(false, false, 0)
  ARD_FFT_CASE(false, false, 1)
  ARD_FFT_CASE(true, false, 0)
  ARD_FFT_CASE(false, true, 0)
#undef ARD_FFT_CASE
  return -1;
}

}  // namespace ard

namespace ard {

int fft_setup(Ctx& c, GemmDesc& g, bool inverse) {
  const int L = g.K;
  std::vector<int> rad;
  if (g.M != L || !fft_factor(L, rad))
    return set_error(ARD_ERR_UNSUPPORTED, "axis length with a prime factor above 13 (FFT-style transform)");
  auto it = c.fft_plans.find(L);
  if (it == c.fft_plans.end()) {
    const std::vector<float> t = fft_plan_table(L, rad);
    float* d;
    int r;
    if ((r = c.dalloc((void**)&d, t.size() * 4)) || (r = c.upload(d, t.data(), t.size() * 4))) return r;
    it = c.fft_plans.emplace(L, d).first;
  }
  g.plan = it->second;
  g.tlc = fft_lines(L);
  g.inv = inverse ? 1 : 0;
  g.nrad = int(rad.size());
  for (size_t i = 0; i < rad.size(); ++i) g.rad[i] = (unsigned char)rad[i];
  return ARD_OK;
}

int fft_tiles(const GemmDesc& g) { return (g.N + 2 * g.tlc - 1) / (2 * g.tlc); }

int fft_selftest(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  Ctx c;
  c.device = device;
  if ((e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess)
    return hip_fail(e, "hipStreamCreate");
  const size_t n = size_t(L) * N;
  float *dX, *dY;
  int r;
  if ((r = c.dalloc((void**)&dX, n * 4)) || (r = c.dalloc((void**)&dY, n * 4)) || (r = c.upload(dX, X, n * 4)))
    return r;
  GemmDesc g{};
  g.B = dX; g.C = dY;
  g.M = L; g.K = L; g.N = N; g.N1 = N;
  if (b_layout == 0) { g.bK = N; g.b1 = 1; g.cM = N; g.c1 = 1; }  // X, Y [L][N]
  else { g.bK = 1; g.b1 = L; g.cM = 1; g.c1 = L; }                // X, Y [N][L]
  g.alpha = 1.f;
  if ((r = fft_setup(c, g, inverse != 0))) return r;
  GemmDesc* dd;
  if ((r = c.dalloc((void**)&dd, sizeof g)) || (r = c.upload(dd, &g, sizeof g))) return r;
  if (launch_fft(dd, 1, fft_tiles(g), b_layout == 1, b_layout == 1, 0, fft_lds_bytes(g), c.stream))
    return set_error(ARD_ERR_HIP, "FFT transform launch");
  if ((e = hipMemcpyAsync(Y, dY, n * 4, hipMemcpyDeviceToHost, c.stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c.stream)) != hipSuccess)
    return hip_fail(e, "FFT transform result");
  return ARD_OK;
}

}  // namespace ard
