// ARD time-stepping oracle: a plain CPU restatement of the reference's hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP product path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// the library built from it, and never as the thing measured or shipped.
//
// Every function below cites the reference file:line it restates
// (paths relative to the reference root, ARD = ARD-simulator-190113/).
// It is templated on the arithmetic type: R=float reproduces the reference's
// FP32 operation order (real_t = float, ARD/types.h:2); R=double is the FP64
// oracle used to state the error of the FP32 forms.
//
// Pins (tests/test_oracle_*.py): the DCT against scipy.fft.dctn type 2/3
// (FFTW REDFT10/REDFT01 definitions), the Gaussian pulse against closed-form
// values, the modal recurrence against its single-mode closed form, the PML
// operator against scipy.ndimage stencils with zero padding, the interface
// forcing against "global 6th-order Laplacian minus the mirrored (Neumann)
// one", the topology against hand-derived partition/boundary lists, and the
// coupled solver against physical invariants (linearity, mirror symmetry,
// split-vs-merged room). The reference itself cannot be built here: it needs
// Vulkan headers and the FFTW and SDL libraries, none of which the image has.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <algorithm>
#include <limits>

namespace {

constexpr double kPi = 3.14159265358979323846;
// Loops over fewer elements than this run on one thread: on small scenes a
// parallel region per loop costs more than it saves.
constexpr long kParallelMin = 1L << 16;

enum PmlType { P_LEFT = 0, P_RIGHT, P_TOP, P_BOTTOM, P_FRONT, P_BACK };
enum BType { X_BOUNDARY = 0, Y_BOUNDARY, Z_BOUNDARY };

// Per-axis DCT tables in double. REDFT10: Y_k = 2 sum_j x_j cos(pi k (j+1/2)/n);
// REDFT01: Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2)/n)  (FFTW's definitions,
// used by ARD/dct_volume.cpp:17,19).
struct DctTables {
  int n;
  std::vector<double> fwd;  // [k][j]
  std::vector<double> inv;  // [k][j]
  explicit DctTables(int n_) : n(n_), fwd(size_t(n_) * n_), inv(size_t(n_) * n_) {
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        fwd[size_t(k) * n + j] = 2.0 * std::cos(kPi * k * (j + 0.5) / n);
        inv[size_t(k) * n + j] = (j == 0) ? 1.0 : 2.0 * std::cos(kPi * j * (k + 0.5) / n);
      }
  }
};

// Unnormalised separable 3D transform along x, y, z (data z-major, x fastest:
// ARD/dct_volume.cpp:70-88), accumulated in double.
void transform3d(int W, int H, int D, const DctTables& tx, const DctTables& ty, const DctTables& tz,
                 bool inverse, std::vector<double>& a, std::vector<double>& tmp) {
  const std::vector<double>& mx = inverse ? tx.inv : tx.fwd;
  const std::vector<double>& my = inverse ? ty.inv : ty.fwd;
  const std::vector<double>& mz = inverse ? tz.inv : tz.fwd;
  const size_t plane = size_t(W) * H;
  tmp.assign(a.size(), 0.0);
  const bool par = long(a.size()) > kParallelMin;
  // x
#pragma omp parallel for schedule(static) if (par)
  for (long r = 0; r < long(H) * D; ++r) {
    const double* in = &a[size_t(r) * W];
    double* out = &tmp[size_t(r) * W];
    for (int k = 0; k < W; ++k) {
      double s = 0.0;
      const double* c = &mx[size_t(k) * W];
      for (int j = 0; j < W; ++j) s += c[j] * in[j];
      out[k] = s;
    }
  }
  // y
  std::fill(a.begin(), a.end(), 0.0);
#pragma omp parallel for collapse(2) schedule(static) if (par)
  for (int z = 0; z < D; ++z)
    for (int k = 0; k < H; ++k) {
      double* out = &a[z * plane + size_t(k) * W];
      for (int j = 0; j < H; ++j) {
        const double c = my[size_t(k) * H + j];
        const double* in = &tmp[z * plane + size_t(j) * W];
        for (int x = 0; x < W; ++x) out[x] += c * in[x];
      }
    }
  // z
  std::fill(tmp.begin(), tmp.end(), 0.0);
#pragma omp parallel for schedule(static) if (par)
  for (int k = 0; k < D; ++k) {
    double* out = &tmp[size_t(k) * plane];
    for (int j = 0; j < D; ++j) {
      const double c = mz[size_t(k) * D + j];
      const double* in = &a[size_t(j) * plane];
      for (size_t i = 0; i < plane; ++i) out[i] += c * in[i];
    }
  }
  a.swap(tmp);
}

// ---------------------------------------------------------------------------
// O(n log n) REDFT10 / REDFT01, the algorithm class of the reference's FFTW
// plans (ARD/dct_volume.cpp:17,19 plan fftwf_plan_r2r_3d with FFTW_REDFT10 /
// FFTW_REDFT01; FFTW 3.3.5 computes them through FFTs). Restated here from the
// published method, not from FFTW's code:
//   * Makhoul's reordering (IEEE Trans. ASSP 28(1), 1980): for a line x of
//     length n, v[j] = x[2j] (j < ceil(n/2)), v[n-1-j] = x[2j+1]
//     (j < floor(n/2)); with V = DFT_n(v), REDFT10(x)_k = 2 Re(e^{-i pi k/2n} V_k).
//     Inversely REDFT01(X) = the real part of the unnormalised inverse DFT of
//     e^{i pi k/2n} (X_k - i X_{n-k}) (X_n = 0), read back through the same
//     reordering.
//   * the DFT by recursive mixed-radix Cooley-Tukey decimation in time (any n:
//     radix 4, 2, 3, 5 butterflies, other primes as a direct DFT of that
//     length), in double, applied to a batch of lines at once: element j of
//     every line of the batch is one contiguous vector, so the inner loops run
//     over the batch.
// Pinned against the dense sums (transform3d) and scipy.fft.dct type 2/3.

struct FftPlan {
  int n = 0;
  std::vector<int> radix;         // factorisation of n, radices of the recursion levels
  std::vector<double> wr, wi;     // e^{-2 pi i e / n}, e in [0, n)
  std::vector<double> pr, pi;     // Makhoul post/pre twiddles e^{-i pi k / 2n}
  std::vector<int> perm;          // v[j] = x[perm[j]]
  explicit FftPlan(int n_) : n(n_), wr(n_), wi(n_), pr(n_), pi(n_), perm(n_) {
    int m = n;
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    for (int p = 3; m > 1; p += 2)
      while (m % p == 0) { radix.push_back(p); m /= p; }
    for (int e = 0; e < n; ++e) {
      wr[e] = std::cos(2.0 * kPi * e / n);
      wi[e] = -std::sin(2.0 * kPi * e / n);
      pr[e] = std::cos(kPi * e / (2.0 * n));
      pi[e] = -std::sin(kPi * e / (2.0 * n));
    }
    for (int j = 0; 2 * j < n; ++j) perm[j] = 2 * j;
    for (int j = 0; 2 * j + 1 < n; ++j) perm[n - 1 - j] = 2 * j + 1;
  }
};

// out[k] (k < len) = DFT of the len inputs in[r * istride], r < len, every
// element a batch vector of B doubles (re and im arrays); level = index into
// P.radix; the twiddle W_len^e is P.w[e * (P.n / len)].
void fft_rec(const FftPlan& P, int level, int len, const double* in_r, const double* in_i, long istride,
             double* out_r, double* out_i, int B) {
  if (len == 1) {
    std::memcpy(out_r, in_r, sizeof(double) * B);
    std::memcpy(out_i, in_i, sizeof(double) * B);
    return;
  }
  const int p = P.radix[level];
  const int m = len / p;
  for (int r = 0; r < p; ++r)  // sub-DFT r of x[r + p j] -> out[r m + k]
    fft_rec(P, level + 1, m, in_r + r * istride * B, in_i + r * istride * B, istride * p, out_r + long(r) * m * B,
            out_i + long(r) * m * B, B);
  const long ts = P.n / len;  // twiddle stride for W_len
  std::vector<double> tr(p), ti(p), gr, gi;
  if (p > 5) { gr.resize(size_t(p) * B); gi.resize(size_t(p) * B); }
  // Twiddled input r of column k, element b: t_r = W_len^{r k} Y_r[k].
  for (int k = 0; k < m; ++k) {
    for (int r = 1; r < p; ++r) {
      const long e = (long(r) * k) * ts;
      tr[r] = P.wr[e];
      ti[r] = P.wi[e];
    }
    // X[k + m q] = sum_r t_r W_p^{r q}, written over the inputs (the same slots)
    double* __restrict y0r = out_r + long(k) * B;
    double* __restrict y0i = out_i + long(k) * B;
    const long sB = long(m) * B;
    if (p == 2) {
      double* __restrict y1r = y0r + sB;
      double* __restrict y1i = y0i + sB;
      const double c1 = tr[1], s1 = ti[1];
      for (int b = 0; b < B; ++b) {
        const double ar = y0r[b], ai = y0i[b];
        const double br = y1r[b] * c1 - y1i[b] * s1, bi = y1r[b] * s1 + y1i[b] * c1;
        y0r[b] = ar + br; y0i[b] = ai + bi;
        y1r[b] = ar - br; y1i[b] = ai - bi;
      }
    } else if (p == 4) {
      double* __restrict y1r = y0r + sB;
      double* __restrict y1i = y0i + sB;
      double* __restrict y2r = y0r + 2 * sB;
      double* __restrict y2i = y0i + 2 * sB;
      double* __restrict y3r = y0r + 3 * sB;
      double* __restrict y3i = y0i + 3 * sB;
      const double c1 = tr[1], s1 = ti[1], c2 = tr[2], s2 = ti[2], c3 = tr[3], s3 = ti[3];
      for (int b = 0; b < B; ++b) {
        const double a0r = y0r[b], a0i = y0i[b];
        const double a1r = y1r[b] * c1 - y1i[b] * s1, a1i = y1r[b] * s1 + y1i[b] * c1;
        const double a2r = y2r[b] * c2 - y2i[b] * s2, a2i = y2r[b] * s2 + y2i[b] * c2;
        const double a3r = y3r[b] * c3 - y3i[b] * s3, a3i = y3r[b] * s3 + y3i[b] * c3;
        const double s02r = a0r + a2r, s02i = a0i + a2i, d02r = a0r - a2r, d02i = a0i - a2i;
        const double s13r = a1r + a3r, s13i = a1i + a3i, d13r = a1r - a3r, d13i = a1i - a3i;
        // W_4 = -i: X1 = d02 - i d13, X3 = d02 + i d13
        y0r[b] = s02r + s13r; y0i[b] = s02i + s13i;
        y2r[b] = s02r - s13r; y2i[b] = s02i - s13i;
        y1r[b] = d02r + d13i; y1i[b] = d02i - d13r;
        y3r[b] = d02r - d13i; y3i[b] = d02i + d13r;
      }
    } else if (p == 3) {
      // W_3 = -1/2 - i sqrt(3)/2: X0 = a0 + t, X1,2 = a0 - t/2 -/+ i (sqrt(3)/2) (a1 - a2)
      double* __restrict y1r = y0r + sB;
      double* __restrict y1i = y0i + sB;
      double* __restrict y2r = y0r + 2 * sB;
      double* __restrict y2i = y0i + 2 * sB;
      const double c1 = tr[1], s1 = ti[1], c2 = tr[2], s2 = ti[2];
      const double h3 = std::sqrt(3.0) / 2.0;
      for (int b = 0; b < B; ++b) {
        const double a0r = y0r[b], a0i = y0i[b];
        const double a1r = y1r[b] * c1 - y1i[b] * s1, a1i = y1r[b] * s1 + y1i[b] * c1;
        const double a2r = y2r[b] * c2 - y2i[b] * s2, a2i = y2r[b] * s2 + y2i[b] * c2;
        const double tr_ = a1r + a2r, ti_ = a1i + a2i, ur = (a1r - a2r) * h3, ui = (a1i - a2i) * h3;
        const double mr = a0r - 0.5 * tr_, mi = a0i - 0.5 * ti_;
        y0r[b] = a0r + tr_; y0i[b] = a0i + ti_;
        y1r[b] = mr + ui; y1i[b] = mi - ur;
        y2r[b] = mr - ui; y2i[b] = mi + ur;
      }
    } else if (p == 5) {
      // W_5^r = cos(2 pi r/5) - i sin(2 pi r/5); pairs (1,4) and (2,3)
      double* __restrict y1r = y0r + sB;
      double* __restrict y1i = y0i + sB;
      double* __restrict y2r = y0r + 2 * sB;
      double* __restrict y2i = y0i + 2 * sB;
      double* __restrict y3r = y0r + 3 * sB;
      double* __restrict y3i = y0i + 3 * sB;
      double* __restrict y4r = y0r + 4 * sB;
      double* __restrict y4i = y0i + 4 * sB;
      const double c1 = tr[1], s1 = ti[1], c2 = tr[2], s2 = ti[2], c3 = tr[3], s3 = ti[3], c4 = tr[4], s4 = ti[4];
      const double k1 = std::cos(2 * kPi / 5), k2 = std::cos(4 * kPi / 5);
      const double q1 = std::sin(2 * kPi / 5), q2 = std::sin(4 * kPi / 5);
      for (int b = 0; b < B; ++b) {
        const double a0r = y0r[b], a0i = y0i[b];
        const double a1r = y1r[b] * c1 - y1i[b] * s1, a1i = y1r[b] * s1 + y1i[b] * c1;
        const double a2r = y2r[b] * c2 - y2i[b] * s2, a2i = y2r[b] * s2 + y2i[b] * c2;
        const double a3r = y3r[b] * c3 - y3i[b] * s3, a3i = y3r[b] * s3 + y3i[b] * c3;
        const double a4r = y4r[b] * c4 - y4i[b] * s4, a4i = y4r[b] * s4 + y4i[b] * c4;
        const double t1r = a1r + a4r, t1i = a1i + a4i, t2r = a2r + a3r, t2i = a2i + a3i;
        const double t3r = a1r - a4r, t3i = a1i - a4i, t4r = a2r - a3r, t4i = a2i - a3i;
        const double b1r = a0r + k1 * t1r + k2 * t2r, b1i = a0i + k1 * t1i + k2 * t2i;
        const double b2r = a0r + k2 * t1r + k1 * t2r, b2i = a0i + k2 * t1i + k1 * t2i;
        const double d1r = q1 * t3r + q2 * t4r, d1i = q1 * t3i + q2 * t4i;
        const double d2r = q2 * t3r - q1 * t4r, d2i = q2 * t3i - q1 * t4i;
        y0r[b] = a0r + t1r + t2r; y0i[b] = a0i + t1i + t2i;
        // X1 = b1 - i d1, X4 = b1 + i d1, X2 = b2 - i d2, X3 = b2 + i d2
        y1r[b] = b1r + d1i; y1i[b] = b1i - d1r;
        y4r[b] = b1r - d1i; y4i[b] = b1i + d1r;
        y2r[b] = b2r + d2i; y2i[b] = b2i - d2r;
        y3r[b] = b2r - d2i; y3i[b] = b2i + d2r;
      }
    } else {
      // any other prime: direct DFT of length p over the twiddled inputs
      for (int r = 0; r < p; ++r) {
        const double* __restrict sr = y0r + r * sB;
        const double* __restrict si = y0i + r * sB;
        double* __restrict g_r = gr.data() + size_t(r) * B;
        double* __restrict g_i = gi.data() + size_t(r) * B;
        const double c = r ? tr[r] : 1.0, sn = r ? ti[r] : 0.0;
        for (int b = 0; b < B; ++b) {
          g_r[b] = sr[b] * c - si[b] * sn;
          g_i[b] = sr[b] * sn + si[b] * c;
        }
      }
      const long ps = P.n / p;
      for (int q = 0; q < p; ++q) {
        double* __restrict orr = y0r + q * sB;
        double* __restrict oi = y0i + q * sB;
        std::memcpy(orr, gr.data(), sizeof(double) * B);
        std::memcpy(oi, gi.data(), sizeof(double) * B);
        for (int r = 1; r < p; ++r) {
          const long e = (long(r) * q) % p * ps;
          const double cr = P.wr[e], ci = P.wi[e];
          const double* __restrict sr = gr.data() + size_t(r) * B;
          const double* __restrict si = gi.data() + size_t(r) * B;
          for (int b = 0; b < B; ++b) {
            orr[b] += sr[b] * cr - si[b] * ci;
            oi[b] += sr[b] * ci + si[b] * cr;
          }
        }
      }
    }
  }
}

// REDFT10 (inverse = false) or REDFT01 (inverse = true) of `nb` lines of
// length P.n: element j of line b at base[b * bstride + j * lstride].
// Two real lines share one complex FFT: line b in the real part and line
// b + Bh in the imaginary part (Bh = ceil(nb / 2)); the forward transform
// separates their spectra as A_k = (Z_k + conj Z_{n-k}) / 2,
// B_k = (Z_k - conj Z_{n-k}) / 2i; the inverse's outputs are real, so the two
// lines come back as the real and imaginary parts.
void dct_lines(const FftPlan& P, double* base, long lstride, long bstride, int nb, bool inverse,
               std::vector<double>& scratch) {
  const int n = P.n;
  const int B = (nb + 1) / 2;
  const int nhi = nb - B;  // lines in the imaginary part
  scratch.resize(size_t(4) * n * B);
  double* vr = scratch.data();
  double* vi = vr + size_t(n) * B;
  double* xr = vi + size_t(n) * B;
  double* xi = xr + size_t(n) * B;
  if (!inverse) {
    for (int j = 0; j < n; ++j) {
      const double* src = base + long(P.perm[j]) * lstride;
      double* dr = vr + size_t(j) * B;
      double* di = vi + size_t(j) * B;
      for (int b = 0; b < B; ++b) dr[b] = src[b * bstride];
      for (int b = 0; b < nhi; ++b) di[b] = src[(B + b) * bstride];
      if (nhi < B) di[B - 1] = 0.0;
    }
    fft_rec(P, 0, n, vr, vi, 1, xr, xi, B);
    for (int k = 0; k < n; ++k) {
      const double c = P.pr[k], s = P.pi[k];
      const int nk = k ? n - k : 0;
      const double* zr = xr + size_t(k) * B;
      const double* zi = xi + size_t(k) * B;
      const double* wr = xr + size_t(nk) * B;
      const double* wi = xi + size_t(nk) * B;
      double* d = base + long(k) * lstride;
      for (int b = 0; b < B; ++b) {
        const double ar = 0.5 * (zr[b] + wr[b]), ai = 0.5 * (zi[b] - wi[b]);
        d[b * bstride] = 2.0 * (ar * c - ai * s);
      }
      for (int b = 0; b < nhi; ++b) {
        const double br = 0.5 * (zi[b] + wi[b]), bi = -0.5 * (zr[b] - wr[b]);
        d[(B + b) * bstride] = 2.0 * (br * c - bi * s);
      }
    }
  } else {
    // V_k = e^{i pi k/2n} (X_k - i X_{n-k}) per line; Z = V_lo + i V_hi;
    // inverse DFT as conj(DFT(conj Z))
    for (int k = 0; k < n; ++k) {
      const double c = P.pr[k], s = -P.pi[k];  // e^{+i pi k/2n} = c + i s
      const double* a = base + long(k) * lstride;
      const double* bq = k ? base + long(n - k) * lstride : nullptr;
      double* dr = vr + size_t(k) * B;
      double* di = vi + size_t(k) * B;
      for (int b = 0; b < B; ++b) {
        const double re = a[b * bstride];
        const double im = bq ? -bq[b * bstride] : 0.0;
        double zr = re * c - im * s, zi = re * s + im * c;
        if (b < nhi) {
          const double re2 = a[(B + b) * bstride];
          const double im2 = bq ? -bq[(B + b) * bstride] : 0.0;
          zr -= re2 * s + im2 * c;  // + i (re2 c - im2 s + i (re2 s + im2 c))
          zi += re2 * c - im2 * s;
        }
        dr[b] = zr;
        di[b] = -zi;  // conjugated
      }
    }
    fft_rec(P, 0, n, vr, vi, 1, xr, xi, B);
    for (int j = 0; j < n; ++j) {
      const double* ar = xr + size_t(j) * B;
      const double* ai = xi + size_t(j) * B;
      double* d = base + long(P.perm[j]) * lstride;
      for (int b = 0; b < B; ++b) d[b * bstride] = ar[b];
      for (int b = 0; b < nhi; ++b) d[(B + b) * bstride] = -ai[b];  // conj of the DFT
    }
  }
}

// The separable 3D transform of transform3d by FFT, in place on a (z-major,
// x fastest). Lines along y and z are batched over x (and y), which are
// contiguous; lines along x are batched over rows.
void transform3d_fft(int W, int H, int D, const FftPlan& px, const FftPlan& py, const FftPlan& pz, bool inverse,
                     std::vector<double>& a) {
  const long plane = long(W) * H;
  auto chunk = [](int n) { return std::max(8, std::min(512, (1 << 15) / std::max(n, 1))); };
  {  // x: lines contiguous, batch over rows
    const int B = chunk(W);
    const long rows = long(H) * D;
#pragma omp parallel if (long(W) * H * D > kParallelMin)
    {
      std::vector<double> s;
#pragma omp for schedule(dynamic)
      for (long r0 = 0; r0 < rows; r0 += B)
        dct_lines(px, a.data() + r0 * W, 1, W, int(std::min<long>(B, rows - r0)), inverse, s);
    }
  }
  {  // y: for each z plane, lines of stride W, batch over x
    const int B = chunk(H);
    const long nchunks = (W + B - 1) / B;
#pragma omp parallel if (long(W) * H * D > kParallelMin)
    {
      std::vector<double> s;
#pragma omp for schedule(dynamic) collapse(2)
      for (int z = 0; z < D; ++z)
        for (long c = 0; c < nchunks; ++c)
          dct_lines(py, a.data() + z * plane + c * B, W, 1, int(std::min<long>(B, W - c * B)), inverse, s);
    }
  }
  {  // z: lines of stride W*H, batch over the plane
    const int B = chunk(D);
    const long nchunks = (plane + B - 1) / B;
#pragma omp parallel if (long(W) * H * D > kParallelMin)
    {
      std::vector<double> s;
#pragma omp for schedule(dynamic)
      for (long c = 0; c < nchunks; ++c)
        dct_lines(pz, a.data() + c * B, plane, 1, int(std::min<long>(B, plane - c * B)), inverse, s);
    }
  }
}

template <typename R>
struct Part {
  int id = 0;
  bool is_pml = false;
  int pml_type = -1;
  int parent = -1;
  int xs = 0, ys = 0, zs = 0, w = 0, h = 0, d = 0;
  int xe() const { return xs + w; }
  int ye() const { return ys + h; }
  int ze() const { return zs + d; }
  bool include_self = true;  // ARD/partition.h:38; false for PML (ARD/pml_partition.cpp:20)
  // free-border masks, [depth][height] or [depth][width] (ARD/partition.cpp:22-51)
  std::vector<std::vector<int>> left_free, right_free, top_free, bottom_free;
  std::vector<int> sources;  // indices into Sim::sources

  // DCT partition state (ARD/dct_partition.h:4-15, ARD/dct_volume.h)
  std::vector<R> p_values, p_modes, f_values, f_modes, prev_modes, next_modes, cwt, w2;
  // PML state (ARD/pml_partition.h:8-25); arrays hold w*h*d+1 entries, the last
  // one being the zero sentinel read for out-of-range neighbours.
  std::vector<R> p_old, p, p_new, phx, phx_n, phy, phy_n, phz, phz_n, force, zx, zy, zz;

  size_t ncell() const { return size_t(w) * h * d; }
};

struct Bnd {
  int type;
  double absorption;
  int a, b;
  int xs, xe, ys, ye, zs, ze;
};

// Face elements of a boundary (the loop count of its forcing).
long bnd_area(const Bnd& b) {
  return b.type == X_BOUNDARY ? long(b.ye - b.ys) * (b.ze - b.zs)
         : b.type == Y_BOUNDARY ? long(b.xe - b.xs) * (b.ze - b.zs) : long(b.xe - b.xs) * (b.ye - b.ys);
}

struct Source {
  int x, y, z;
  std::vector<double> wave;  // empty -> Gaussian pulse
};

struct Recorder {
  int x, y, z;  // global cell coordinates
  int part = -1;
};

template <typename R>
struct Sim {
  // parameters (ARD/main.cpp:29-45 statics)
  R dh, dt, c0, absorption;
  int pml_layers;
  double pml_scale[6] = {1, 1, 1, 1, 1, 1};  // per-face damping scale (1 = reference)
  std::vector<std::unique_ptr<Part<R>>> parts;
  std::vector<Bnd> bnds;
  std::vector<Source> sources;
  std::vector<Recorder> recs;
  int num_dct = 0;
  int time_step = 0;
  int rec_local = 0;  // 0: reference addressing (global coords index the array), 1: local cell
  bool finalized = false;
  int fft_dct = 1;  // 1: O(n log n) transforms (transform3d_fft), 0: dense sums (transform3d)
  std::vector<std::unique_ptr<DctTables>> tables;  // cached by size
  std::vector<std::unique_ptr<FftPlan>> plans;     // cached by size
  std::vector<double> work, tmp;

  const DctTables& table(int n) {
    for (auto& t : tables)
      if (t->n == n) return *t;
    tables.emplace_back(new DctTables(n));
    return *tables.back();
  }
  const FftPlan& plan(int n) {
    for (auto& t : plans)
      if (t->n == n) return *t;
    plans.emplace_back(new FftPlan(n));
    return *plans.back();
  }
  void transform(Part<R>& p, bool inverse) {
    if (fft_dct) transform3d_fft(p.w, p.h, p.d, plan(p.w), plan(p.h), plan(p.d), inverse, work);
    else transform3d(p.w, p.h, p.d, table(p.w), table(p.h), table(p.d), inverse, work, tmp);
  }

  // Partition ctor + DctPartition ctor tables: ARD/partition.cpp:10-52, ARD/dct_partition.cpp:7-37.
  Part<R>* new_part(int xs, int ys, int zs, int w, int h, int d) {
    auto* p = new Part<R>();
    p->id = int(parts.size());
    p->xs = xs; p->ys = ys; p->zs = zs; p->w = w; p->h = h; p->d = d;
    p->left_free.assign(d, std::vector<int>(h, 1));
    p->right_free.assign(d, std::vector<int>(h, 1));
    p->top_free.assign(d, std::vector<int>(w, 1));
    p->bottom_free.assign(d, std::vector<int>(w, 1));
    parts.emplace_back(p);
    return p;
  }

  int add_dct(int xs, int ys, int zs, int w, int h, int d) {
    Part<R>* p = new_part(xs, ys, zs, w, h, d);
    const size_t n = p->ncell();
    p->p_values.assign(n, R(0)); p->p_modes.assign(n, R(0));
    p->f_values.assign(n, R(0)); p->f_modes.assign(n, R(0));
    p->prev_modes.assign(n, R(0)); p->next_modes.assign(n, R(0));
    p->cwt.assign(n, R(0)); p->w2.assign(n, R(0));
    // ARD/dct_partition.cpp:20-36: 1-based wavenumbers, FP32 in the reference.
    const R lx2 = R(w * w) * dh * dh;
    const R ly2 = R(h * h) * dh * dh;
    const R lz2 = R(d * d) * dh * dh;
    for (int i = 1; i <= d; ++i)
      for (int j = 1; j <= h; ++j)
        for (int k = 1; k <= w; ++k) {
          const size_t idx = size_t(i - 1) * h * w + size_t(j - 1) * w + (k - 1);
          const R wv = c0 * R(kPi) * std::sqrt(R(i * i) / lz2 + R(j * j) / ly2 + R(k * k) / lx2);
          p->w2[idx] = wv * wv;
          p->cwt[idx] = std::cos(wv * dt);
        }
    num_dct++;
    return p->id;
  }

  // PmlPartition ctor: ARD/pml_partition.cpp:17-93.
  Part<R>* add_pml(int parent, int type, int xs, int ys, int zs, int w, int h, int d) {
    Part<R>* p = new_part(xs, ys, zs, w, h, d);
    p->is_pml = true;
    p->pml_type = type;
    p->parent = parent;
    p->include_self = false;
    const size_t n = p->ncell() + 1;
    for (auto* v : {&p->p_old, &p->p, &p->p_new, &p->phx, &p->phx_n, &p->phy, &p->phy_n, &p->phz,
                    &p->phz_n, &p->force, &p->zx, &p->zy, &p->zz})
      v->assign(n, R(0));
    const R thickness = R(pml_layers) * dh;
    const R R_ = R(0.00001f);
    const R zeta = c0 / thickness * std::log10(R(1) / R_);
    const R pi = R(kPi);
    const R sc = R(pml_scale[type]);
    auto prof = [&](int q) {  // q = distance index 1..layers into the layer
      return zeta * (R(q) / thickness * dh - std::sin(R(2) * pi * R(q) * dh / thickness) / R(2) / pi);
    };
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
          const size_t idx = size_t(k) * h * w + size_t(j) * w + i;
          switch (type) {
            case P_BACK: p->zz[idx] = sc * prof(k + 1); break;
            case P_FRONT: p->zz[idx] = sc * prof(d - k); break;
            case P_BOTTOM: p->zy[idx] = sc * prof(j + 1); break;
            case P_TOP: p->zy[idx] = sc * prof(h - j); break;
            case P_RIGHT: p->zx[idx] = sc * prof(i + 1); break;
            case P_LEFT: p->zx[idx] = sc * prof(w - i); break;
          }
        }
    return p;
  }

  // Boundary::FindBoundary: ARD/boundary.cpp:289-334.
  bool find_boundary(int ia, int ib, double absorp, Bnd* out) {
    const Part<R>& a = *parts[ia];
    const Part<R>& b = *parts[ib];
    const int xa_min = a.xs, xa_max = a.xs + a.w, xb_min = b.xs, xb_max = b.xs + b.w;
    const int x_ov = std::min(xa_max, xb_max) - std::max(xa_min, xb_min);
    const int ya_min = a.ys, ya_max = a.ys + a.h, yb_min = b.ys, yb_max = b.ys + b.h;
    const int y_ov = std::min(ya_max, yb_max) - std::max(ya_min, yb_min);
    const int zs = a.zs, ze = a.zs + a.d;
    if (x_ov == 0 && y_ov > 0) {
      const bool right = (xa_max == xb_min);
      const int xs = right ? xa_max - 3 : xb_max - 3;
      const int ys = ya_min + std::max(0, yb_min - ya_min);
      *out = Bnd{X_BOUNDARY, absorp, ia, ib, xs, xs + 6, ys, ys + y_ov, zs, ze};
      return true;
    } else if (y_ov == 0 && x_ov > 0) {
      const bool bottom = (ya_max == yb_min);
      const int xs = xa_min + std::max(0, xb_min - xa_min);
      const int ys = bottom ? ya_max - 3 : yb_max - 3;
      *out = Bnd{Y_BOUNDARY, absorp, ia, ib, xs, xs + x_ov, ys, ys + 6, zs, ze};
      return true;
    }
    return false;
  }

  // Partition::AddBoundary: ARD/partition.cpp:96-143.
  void add_boundary_mask(int ip, const Bnd& bd) {
    Part<R>& p = *parts[ip];
    if (bd.type == X_BOUNDARY) {
      auto& m = (bd.xs < p.xs && bd.xe > p.xs) ? p.left_free : p.right_free;
      for (int i = 0; i < p.d; ++i)
        for (int j = bd.ys; j < bd.ye; ++j) m[i][j - p.ys] = 0;
    } else if (bd.type == Y_BOUNDARY) {
      auto& m = (bd.ys < p.ys && bd.ye > p.ys) ? p.top_free : p.bottom_free;
      for (int i = 0; i < p.d; ++i)
        for (int j = bd.xs; j < bd.xe; ++j) m[i][j - p.xs] = 0;
    }
  }

  // One face's PML run detection: ARD/simulation.cpp:62-212 (four copies of one loop).
  template <typename F>
  void runs(const std::vector<int>& freev, int n, F make) {
    int start = 0, end = 0;
    bool started = false;
    for (int i = 0; i < n; ++i) {
      if (started) {
        end++;
        if (freev[i] && i != n - 1) continue;
        else if (start != end) {
          if (i != n - 1) end--;
          make(start, end - start + 1);
          started = false;
        }
      } else if (freev[i]) {
        start = i;
        end = i;
        started = true;
      }
    }
  }

  // Simulation ctor: ARD/simulation.cpp:14-294 (topology part).
  int finalize() {
    if (finalized) return -1;
    const int ndct = int(parts.size());
    for (int i = 0; i < ndct; ++i)
      for (int j = i + 1; j < ndct; ++j) {
        Bnd bd;
        if (find_boundary(i, j, 1.0, &bd)) {
          bnds.push_back(bd);
          add_boundary_mask(i, bd);
          add_boundary_mask(j, bd);
        }
      }
    // sources -> partitions (ARD/simulation.cpp:37-48), DCT partitions only.
    for (int i = 0; i < ndct; ++i) {
      Part<R>& p = *parts[i];
      for (int s = 0; s < int(sources.size()); ++s) {
        const Source& so = sources[s];
        if (so.x >= p.xs && so.x < p.xe() && so.y >= p.ys && so.y < p.ye() && so.z >= p.zs && so.z < p.ze())
          p.sources.push_back(s);
      }
    }
    const double ab = double(absorption);
    for (int c = 0; c < ndct; ++c) {
      const int L = pml_layers;
      auto mk = [&](int type, int xs, int ys, int zs, int w, int h, int d) {
        Part<R>* pml = add_pml(c, type, xs, ys, zs, w, h, d);
        return pml->id;
      };
      {
        const Part<R> P = copy_geom(c);
        runs(P.left_free[0], P.h, [&](int s, int len) {
          int id = mk(P_LEFT, P.xs - L, P.ys + s, P.zs, L, len, P.d);
          Bnd bd; find_boundary(id, c, ab, &bd); bnds.push_back(bd);
        });
        runs(P.right_free[0], P.h, [&](int s, int len) {
          int id = mk(P_RIGHT, P.xe(), P.ys + s, P.zs, L, len, P.d);
          Bnd bd; find_boundary(id, c, ab, &bd); bnds.push_back(bd);
        });
        runs(P.top_free[0], P.w, [&](int s, int len) {
          int id = mk(P_TOP, P.xs + s, P.ys - L, P.zs, len, L, P.d);
          Bnd bd; find_boundary(id, c, ab, &bd); bnds.push_back(bd);
        });
        runs(P.bottom_free[0], P.w, [&](int s, int len) {
          int id = mk(P_BOTTOM, P.xs + s, P.ye(), P.zs, len, L, P.d);
          Bnd bd; find_boundary(id, c, ab, &bd); bnds.push_back(bd);
        });
        int idf = mk(P_FRONT, P.xs, P.ys, P.zs - L, P.w, P.h, L);
        bnds.push_back(Bnd{Z_BOUNDARY, ab, idf, c, P.xs, P.xe(), P.ys, P.ye(), P.zs - 3, P.zs + 3});
        int idb = mk(P_BACK, P.xs, P.ys, P.ze(), P.w, P.h, L);
        bnds.push_back(Bnd{Z_BOUNDARY, ab, idb, c, P.xs, P.xe(), P.ys, P.ye(), P.ze() - 3, P.ze() + 3});
      }
    }
    // recorders: Recorder::FindPartition over the DCT partitions (ARD/recorder.cpp:26-38).
    for (auto& r : recs) {
      r.part = -1;
      for (int i = 0; i < ndct; ++i) {
        const Part<R>& p = *parts[i];
        if (p.xs < r.x - 5 && p.xe() > r.x + 4 && p.ys < r.y - 5 && p.ye() > r.y + 4 && p.zs < r.z - 5 &&
            p.ze() > r.z + 4) {
          r.part = i;
          break;
        }
      }
      if (r.part < 0) return -2;
      if (rec_flat(r) < 0) return -3;
    }
    finalized = true;
    return 0;
  }

  // Flat index of a recorder's sample in its partition's pressure array.
  long long rec_flat(const Recorder& r) const {
    const Part<R>& p = *parts[r.part];
    long long flat;
    if (rec_local)
      flat = (long long)(r.z - p.zs) * p.h * p.w + (long long)(r.y - p.ys) * p.w + (r.x - p.xs);
    else  // ARD/recorder.cpp:55 -> ARD/dct_volume.cpp:70-73 with global coordinates
      flat = (long long)r.z * p.h * p.w + (long long)r.y * p.w + r.x;
    return (flat < 0 || flat >= (long long)p.ncell()) ? -1 : flat;
  }

  Part<R> copy_geom(int c) {
    Part<R> P;
    const Part<R>& s = *parts[c];
    P.xs = s.xs; P.ys = s.ys; P.zs = s.zs; P.w = s.w; P.h = s.h; P.d = s.d;
    P.left_free = s.left_free; P.right_free = s.right_free;
    P.top_free = s.top_free; P.bottom_free = s.bottom_free;
    return P;
  }

  // GaussianSource::SampleValue: ARD/gaussian_source.cpp:16-20.
  R sample(const Source& s, int step) const {
    if (!s.wave.empty()) return (step >= 0 && step < int(s.wave.size())) ? R(s.wave[step]) : R(0);
    const R t = R(step);
    const R arg = std::pow(R(kPi) * ((R(2) * (c0 * dt / dh) * t) / R(6) - R(2)), R(2));
    return R(1e9f) * std::exp(-arg);
  }

  // DctVolume::ExecuteDct / ExecuteIdct with scaling: ARD/dct_volume.cpp:38-68.
  void dct_forward(Part<R>& p) {
    const long n = long(p.ncell());
    work.resize(n);
#pragma omp parallel for schedule(static) if (n > kParallelMin)
    for (long i = 0; i < n; ++i) work[i] = double(p.f_values[i]);
    transform(p, false);
    const R scale = R(1) / (R(2) * std::sqrt(R(2) * R(p.d) * R(p.w) * R(p.h)));
#pragma omp parallel for schedule(static) if (n > kParallelMin)
    for (long i = 0; i < n; ++i) p.f_modes[i] = R(work[i]) * scale;
  }
  void dct_inverse(Part<R>& p) {
    const long n = long(p.ncell());
    work.resize(n);
#pragma omp parallel for schedule(static) if (n > kParallelMin)
    for (long i = 0; i < n; ++i) work[i] = double(p.p_modes[i]);
    transform(p, true);
    const R scale = R(1) / std::sqrt(R(2) * R(p.d) * R(p.w) * R(p.h));
#pragma omp parallel for schedule(static) if (n > kParallelMin)
    for (long i = 0; i < n; ++i) p.p_values[i] = R(work[i]) * scale;
  }

  // DctPartition::Update: ARD/dct_partition.cpp:48-73 (the #else branch).
  void dct_update(Part<R>& p) {
    dct_forward(p);
    const long n = long(p.ncell());
#pragma omp parallel for schedule(static) if (n > kParallelMin)
    for (long i = 0; i < n; ++i)
      p.next_modes[i] = R(0.999f) * (R(2) * p.p_modes[i] * p.cwt[i] - p.prev_modes[i] +
                                     (R(2) * p.f_modes[i] / p.w2[i]) * (R(1) - p.cwt[i]));
    p.prev_modes.swap(p.p_modes);  // prev <- current (memcpy at :69)
    p.p_modes.swap(p.next_modes);  // current <- next (memcpy at :70; next is rewritten before any read)
    dct_inverse(p);
  }

  // PmlPartition::GetIndex: ARD/pml_partition.cpp:6-15.
  static size_t gidx(const Part<R>& p, int x, int y, int z) {
    if (x < 0 || x >= p.w || y < 0 || y >= p.h || z < 0 || z >= p.d) return p.ncell();
    return size_t(z) * p.h * p.w + size_t(y) * p.w + x;
  }

  // PmlPartition::Update: ARD/pml_partition.cpp:113-200.
  void pml_update(Part<R>& P) {
    static const R c6[7] = {2, -27, 270, -490, 270, -27, 2};
    static const R c4[5] = {1, -8, 0, 8, -1};
#pragma omp parallel for schedule(static) if (long(P.ncell()) > kParallelMin)
    for (int k = 0; k < P.d; ++k)
      for (int j = 0; j < P.h; ++j)
        for (int i = 0; i < P.w; ++i) {
          R dxx = 0, dyy = 0, dzz = 0;
          for (int m = 0; m < 7; ++m) {
            dxx += c6[m] * P.p[gidx(P, i + m - 3, j, k)];
            dyy += c6[m] * P.p[gidx(P, i, j + m - 3, k)];
            dzz += c6[m] * P.p[gidx(P, i, j, k + m - 3)];
          }
          dxx /= (R(180) * dh * dh);
          dyy /= (R(180) * dh * dh);
          dzz /= (R(180) * dh * dh);
          const size_t c = gidx(P, i, j, k);
          const R zx = P.zx[c], zy = P.zy[c], zz = P.zz[c];
          const R term1 = R(2) * P.p[c];
          const R term2 = -P.p_old[c];
          const R term3 = c0 * c0 * (dxx + dyy + dzz);
          const R term4 = -(zx + zy + zz) * (P.p[c] - P.p_old[c]) / dt;
          const R term5 = -(zx * zy + zy * zz + zx * zz) * P.p[c];
          R dpx = 0, dpy = 0, dpz = 0;
          for (int m = 0; m < 5; ++m) {
            dpx += c4[m] * P.phx[gidx(P, i + m - 2, j, k)];
            dpy += c4[m] * P.phy[gidx(P, i, j + m - 2, k)];
            dpz += c4[m] * P.phz[gidx(P, i, j, k + m - 2)];
          }
          dpx /= (R(12) * dh);
          dpy /= (R(12) * dh);
          dpz /= (R(12) * dh);
          const R term6 = dpx + dpy + dpz;
          P.p_new[c] = term1 + term2 + dt * dt * (term3 + term4 + term5 + term6 + P.force[c]);
          R dux = 0, duy = 0, duz = 0;
          for (int m = 0; m < 5; ++m) {
            dux += c4[m] * P.p[gidx(P, i + m - 2, j, k)];
            duy += c4[m] * P.p[gidx(P, i, j + m - 2, k)];
            duz += c4[m] * P.p[gidx(P, i, j, k + m - 2)];
          }
          dux /= (R(12) * dh);
          duy /= (R(12) * dh);
          duz /= (R(12) * dh);
          P.phx_n[c] = P.phx[c] - dt * zx * P.phx[c] + dt * (zy + zz - zx) * dux;
          P.phy_n[c] = P.phy[c] - dt * zy * P.phy[c] + dt * (zx + zz - zy) * duy;
          P.phz_n[c] = P.phz[c] - dt * zz * P.phz[c] + dt * (zx + zy - zz) * duz;
        }
    P.phx.swap(P.phx_n);
    P.phy.swap(P.phy_n);
    P.phz.swap(P.phz_n);
    // p_old <- p, p <- p_new, p_new <- old p_old (:194-197)
    std::swap(P.p_old, P.p);
    std::swap(P.p, P.p_new);
    std::fill(P.force.begin(), P.force.begin() + P.ncell(), R(0));  // :199 (sentinel untouched)
  }

  R get_p(const Part<R>& P, int x, int y, int z) const {
    if (P.is_pml) return P.p[gidx(P, x, y, z)];
    return P.p_values[size_t(z) * P.h * P.w + size_t(y) * P.w + x];  // ARD/dct_volume.cpp:70-73
  }
  void set_f(Part<R>& P, int x, int y, int z, R f) {
    if (P.is_pml) P.force[gidx(P, x, y, z)] = f;
    else P.f_values[size_t(z) * P.h * P.w + size_t(y) * P.w + x] = f;  // ARD/dct_volume.cpp:80-83
  }

  // Boundary::ComputeForcingTerms: ARD/boundary.cpp:21-173.
  void forcing(const Bnd& bd) {
    static const R C[6][6] = {{0, 0, -2, 2, 0, 0},       {0, -2, 27, -27, 2, 0},
                              {-2, 27, -270, 270, -27, 2}, {2, -27, 270, -270, 27, -2},
                              {0, 2, -27, 27, -2, 0},      {0, 0, 2, -2, 0, 0}};
    Part<R>& A = *parts[bd.a];
    Part<R>& B = *parts[bd.b];
    const R scale_den = R(180) * dh * dh;
    const R c02 = c0 * c0;
    const R ab = R(bd.absorption);
    if (bd.type == X_BOUNDARY) {
      const bool a_left = (bd.xs <= A.xe() && bd.xe >= A.xe());
      Part<R>& Lp = a_left ? A : B;
      Part<R>& Rp = a_left ? B : A;
#pragma omp parallel for collapse(2) schedule(static) if (bnd_area(bd) > kParallelMin / 64)
      for (int i = bd.ys; i < bd.ye; ++i)
        for (int j = bd.zs; j < bd.ze; ++j) {
          const int ly = i - Lp.ys, ry = i - Rp.ys, lz = j - Lp.zs, rz = j - Rp.zs;
          for (int m = -3; m < 3; ++m) {
            R s1 = 0, s2 = 0;
            for (int n = 0; n < 3; ++n) s1 += C[m + 3][n] * get_p(Lp, Lp.w - 3 + n, ly, lz);
            for (int n = 3; n < 6; ++n) s2 += C[m + 3][n] * get_p(Rp, n - 3, ry, rz);
            if (m < 0) {
              const R sip = Lp.include_self ? (s1 + s2) : s2;
              const R fi = sip * c02 / scale_den;
              set_f(Lp, Lp.w + m, ly, lz, ab * fi);
            } else {
              const R sip = Rp.include_self ? (s1 + s2) : s1;
              const R fi = sip * c02 / scale_den;
              set_f(Rp, m, ry, rz, ab * fi);
            }
          }
        }
    } else if (bd.type == Y_BOUNDARY) {
      const bool a_top = (bd.ys <= A.ye() && bd.ye >= A.ye());
      Part<R>& T = a_top ? A : B;
      Part<R>& Bo = a_top ? B : A;
#pragma omp parallel for collapse(2) schedule(static) if (bnd_area(bd) > kParallelMin / 64)
      for (int i = bd.xs; i < bd.xe; ++i)
        for (int j = bd.zs; j < bd.ze; ++j) {
          const int tx = i - T.xs, bx = i - Bo.xs, tz = j - T.zs, bz = j - Bo.zs;
          for (int m = -3; m < 3; ++m) {
            R s1 = 0, s2 = 0;
            for (int n = 0; n < 3; ++n) s1 += C[m + 3][n] * get_p(T, tx, T.h - 3 + n, tz);
            for (int n = 3; n < 6; ++n) s2 += C[m + 3][n] * get_p(Bo, bx, n - 3, bz);
            if (m < 0) {
              const R sip = T.include_self ? (s1 + s2) : s2;
              set_f(T, tx, T.h + m, tz, ab * (sip * c02 / scale_den));
            } else {
              const R sip = Bo.include_self ? (s1 + s2) : s1;
              set_f(Bo, bx, m, bz, ab * (sip * c02 / scale_den));
            }
          }
        }
    } else {
      const bool a_front = (bd.zs <= A.ze() && bd.ze >= A.ze());
      Part<R>& Fr = a_front ? A : B;
      Part<R>& Bk = a_front ? B : A;
#pragma omp parallel for collapse(2) schedule(static) if (bnd_area(bd) > kParallelMin / 64)
      for (int i = bd.xs; i < bd.xe; ++i)
        for (int j = bd.ys; j < bd.ye; ++j) {
          const int fx = i - Fr.xs, bx = i - Bk.xs, fy = j - Fr.ys, by = j - Bk.ys;
          for (int m = -3; m < 3; ++m) {
            R s1 = 0, s2 = 0;
            for (int n = 0; n < 3; ++n) s1 += C[m + 3][n] * get_p(Fr, fx, fy, Fr.d - 3 + n);
            for (int n = 3; n < 6; ++n) s2 += C[m + 3][n] * get_p(Bk, bx, by, n - 3);
            if (m < 0) {
              const R sip = Fr.include_self ? (s1 + s2) : s2;
              set_f(Fr, fx, fy, Fr.d + m, ab * (sip * c02 / scale_den));
            } else {
              const R sip = Bk.include_self ? (s1 + s2) : s1;
              set_f(Bk, bx, by, m, ab * (sip * c02 / scale_den));
            }
          }
        }
    }
  }

  // Simulation::Update: ARD/simulation.cpp:300-317, serial order (the golden
  // order for the last-writer-wins edge cells), then Recorder::RecordField's
  // point sample (ARD/recorder.cpp:40-57) with the reference's addressing:
  // the recorder's global cell coordinates index the partition's array.
  int step(double* rec_out) {
    const int t = time_step++;
    for (auto& up : parts) {
      Part<R>& P = *up;
      for (int s : P.sources) {
        const Source& so = sources[s];
        set_f(P, so.x - P.xs, so.y - P.ys, so.z - P.zs, sample(so, t));
      }
      if (P.is_pml) pml_update(P);
      else dct_update(P);
    }
    for (const Bnd& bd : bnds) forcing(bd);
    if (rec_out)
      for (size_t r = 0; r < recs.size(); ++r) {
        const Part<R>& P = *parts[recs[r].part];
        rec_out[r] = double(P.p_values[size_t(rec_flat(recs[r]))]);
      }
    return t;
  }
};

struct Handle {
  bool dbl;
  std::unique_ptr<Sim<float>> f;
  std::unique_ptr<Sim<double>> d;
};

template <typename Fn>
auto dispatch(Handle* h, Fn fn) {
  if (h->dbl) return fn(*h->d);
  return fn(*h->f);
}

}  // namespace

extern "C" {

void* orc_create(int use_double, double h, double dt, double c0, int pml_layers, double absorption) {
  auto* hd = new Handle();
  hd->dbl = use_double != 0;
  auto init = [&](auto& s) {
    using R = std::remove_reference_t<decltype(s.dh)>;
    s.dh = R(h); s.dt = R(dt); s.c0 = R(c0); s.absorption = R(absorption);
    s.pml_layers = pml_layers;
  };
  if (hd->dbl) { hd->d.reset(new Sim<double>()); init(*hd->d); }
  else { hd->f.reset(new Sim<float>()); init(*hd->f); }
  return hd;
}

void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

int orc_set_pml_scale(void* h, const double* scale6) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (s.finalized) return -1;
    for (int i = 0; i < 6; ++i) s.pml_scale[i] = scale6[i];
    return 0;
  });
}

// 1: O(n log n) transforms (default), 0: the dense per-axis sums.
int orc_set_fft_dct(void* h, int on) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    s.fft_dct = on ? 1 : 0;
    return 0;
  });
}

int orc_set_recorder_mode(void* h, int local) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (s.finalized) return -1;
    s.rec_local = local ? 1 : 0;
    return 0;
  });
}

int orc_add_dct(void* h, int xs, int ys, int zs, int w, int hh, int d) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (s.finalized || w <= 0 || hh <= 0 || d <= 0) return -1;
    return s.add_dct(xs, ys, zs, w, hh, d);
  });
}

int orc_add_source(void* h, int x, int y, int z, const double* wave, int n) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (s.finalized) return -1;
    Source so{x, y, z, {}};
    if (wave && n > 0) so.wave.assign(wave, wave + n);
    s.sources.push_back(so);
    return int(s.sources.size()) - 1;
  });
}

int orc_add_recorder(void* h, int x, int y, int z) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (s.finalized) return -1;
    s.recs.push_back(Recorder{x, y, z, -1});
    return int(s.recs.size()) - 1;
  });
}

int orc_finalize(void* h) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) { return s.finalize(); });
}

int orc_num_partitions(void* h) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) { return int(s.parts.size()); });
}
int orc_num_boundaries(void* h) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) { return int(s.bnds.size()); });
}

// out[9] = is_pml, pml_type, parent, xs, ys, zs, w, h, d
int orc_partition_info(void* h, int i, int* out) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (i < 0 || i >= int(s.parts.size())) return -1;
    const auto& p = *s.parts[i];
    int v[9] = {p.is_pml ? 1 : 0, p.pml_type, p.parent, p.xs, p.ys, p.zs, p.w, p.h, p.d};
    std::memcpy(out, v, sizeof v);
    return 0;
  });
}

// out[9] = type, a, b, xs, xe, ys, ye, zs, ze
int orc_boundary_info(void* h, int i, int* out) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (i < 0 || i >= int(s.bnds.size())) return -1;
    const Bnd& b = s.bnds[i];
    int v[9] = {b.type, b.a, b.b, b.xs, b.xe, b.ys, b.ye, b.zs, b.ze};
    std::memcpy(out, v, sizeof v);
    return 0;
  });
}

int orc_recorder_partition(void* h, int r) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (r < 0 || r >= int(s.recs.size())) return -1;
    return s.recs[r].part;
  });
}

// Runs n steps; rec_out (may be null) receives [n][num_recorders] samples.
int orc_step(void* h, int n, double* rec_out) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (!s.finalized) return -1;
    const size_t nr = s.recs.size();
    for (int i = 0; i < n; ++i) s.step(rec_out ? rec_out + size_t(i) * nr : nullptr);
    return 0;
  });
}

// which: 0 = pressure (DCT values / PML p), 1 = force, 2 = DCT current modes,
// 3 = DCT previous modes, 4..6 = PML phi x/y/z, 7 = PML p_old.
int orc_get_field(void* h, int part, int which, double* out) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (part < 0 || part >= int(s.parts.size())) return -1;
    auto& P = *s.parts[part];
    const size_t n = P.ncell();
    const std::vector<std::remove_reference_t<decltype(s.dh)>>* v = nullptr;
    if (P.is_pml) {
      switch (which) {
        case 0: v = &P.p; break;
        case 1: v = &P.force; break;
        case 4: v = &P.phx; break;
        case 5: v = &P.phy; break;
        case 6: v = &P.phz; break;
        case 7: v = &P.p_old; break;
        default: return -2;
      }
    } else {
      switch (which) {
        case 0: v = &P.p_values; break;
        case 1: v = &P.f_values; break;
        case 2: v = &P.p_modes; break;
        case 3: v = &P.prev_modes; break;
        default: return -2;
      }
    }
    for (size_t i = 0; i < n; ++i) out[i] = double((*v)[i]);
    return 0;
  });
}

int orc_set_field(void* h, int part, int which, const double* in) {
  return dispatch(static_cast<Handle*>(h), [&](auto& s) {
    if (part < 0 || part >= int(s.parts.size())) return -1;
    auto& P = *s.parts[part];
    using R = std::remove_reference_t<decltype(s.dh)>;
    std::vector<R>* v = nullptr;
    if (P.is_pml) {
      switch (which) {
        case 0: v = &P.p; break;
        case 1: v = &P.force; break;
        case 4: v = &P.phx; break;
        case 5: v = &P.phy; break;
        case 6: v = &P.phz; break;
        case 7: v = &P.p_old; break;
        default: return -2;
      }
    } else {
      switch (which) {
        case 0: v = &P.p_values; break;
        case 1: v = &P.f_values; break;
        case 2: v = &P.p_modes; break;
        case 3: v = &P.prev_modes; break;
        default: return -2;
      }
    }
    for (size_t i = 0; i < P.ncell(); ++i) (*v)[i] = R(in[i]);
    return 0;
  });
}

// Unit-level entries used by the oracle's own pin tests.

// Scaled 3D DCT-II (inverse=0) or DCT-III (inverse=1) exactly as
// DctVolume::ExecuteDct/ExecuteIdct (ARD/dct_volume.cpp:38-68), in R.
// fft: 1 computes the transforms by FFT (transform3d_fft), 0 by the dense sums.
void orc_dct3d(int use_double, int w, int h, int d, const double* in, double* out, int inverse, int fft) {
  std::vector<double> a(in, in + size_t(w) * h * d), tmp;
  if (fft) {
    FftPlan px(w), py(h), pz(d);
    transform3d_fft(w, h, d, px, py, pz, inverse != 0, a);
  } else {
    DctTables tx(w), ty(h), tz(d);
    transform3d(w, h, d, tx, ty, tz, inverse != 0, a, tmp);
  }
  if (use_double) {
    const double sc = inverse ? 1.0 / std::sqrt(2.0 * d * w * h) : 1.0 / (2.0 * std::sqrt(2.0 * d * w * h));
    for (size_t i = 0; i < a.size(); ++i) out[i] = a[i] * sc;
  } else {
    const float sc = inverse ? 1.0f / std::sqrt(2.0f * float(d) * float(w) * float(h))
                             : 1.0f / (2.0f * std::sqrt(2.0f * float(d) * float(w) * float(h)));
    for (size_t i = 0; i < a.size(); ++i) out[i] = double(float(a[i]) * sc);
  }
}

// Mode tables of DctPartition (ARD/dct_partition.cpp:20-36): out_w2, out_cwt.
void orc_mode_tables(int use_double, double dh, double dt, double c0, int w, int h, int d, double* out_w2,
                     double* out_cwt) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    Sim<R> s;
    s.dh = R(dh); s.dt = R(dt); s.c0 = R(c0); s.absorption = R(1); s.pml_layers = 5;
    int id = s.add_dct(0, 0, 0, w, h, d);
    for (size_t i = 0; i < s.parts[id]->ncell(); ++i) {
      out_w2[i] = double(s.parts[id]->w2[i]);
      out_cwt[i] = double(s.parts[id]->cwt[i]);
    }
  };
  if (use_double) run(double(0)); else run(float(0));
}

// Gaussian pulse samples (ARD/gaussian_source.cpp:16-20) for t = 0..n-1.
void orc_gaussian(int use_double, double dh, double dt, double c0, int n, double* out) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    Sim<R> s;
    s.dh = R(dh); s.dt = R(dt); s.c0 = R(c0);
    Source so{0, 0, 0, {}};
    for (int t = 0; t < n; ++t) out[t] = double(s.sample(so, t));
  };
  if (use_double) run(double(0)); else run(float(0));
}

// PML zeta profile for one slab (ARD/pml_partition.cpp:28-92): out[3][w*h*d].
void orc_pml_profile(int use_double, double dh, double c0, int layers, int type, int w, int h, int d, double* out) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    Sim<R> s;
    s.dh = R(dh); s.dt = R(1); s.c0 = R(c0); s.pml_layers = layers;
    Part<R>* p = s.add_pml(0, type, 0, 0, 0, w, h, d);
    const size_t n = p->ncell();
    for (size_t i = 0; i < n; ++i) {
      out[i] = double(p->zx[i]);
      out[n + i] = double(p->zy[i]);
      out[2 * n + i] = double(p->zz[i]);
    }
  };
  if (use_double) run(double(0)); else run(float(0));
}

}  // extern "C"
