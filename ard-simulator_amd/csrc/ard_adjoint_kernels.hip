// Kernels of the discrete adjoint of the plain time step (ard_adjoint.cpp).
// Every operator of the step (ARD/simulation.cpp:300-317) is linear in the
// state, so the reverse sweep applies the transposes in reverse order:
//   recorder gather^T   adds dL/dy into the pressure adjoints
//   interface^T         scatters the adjoints of the written forces into the
//                       adjoints of the pressures the 6th-order stencil read
//   PML update^T        the Grote-Sim update with the stencils transposed, plus
//                       the derivative with respect to the per-face damping scale
//   modal update^T      elementwise (the transforms use the GEMM kernels with
//                       transposed cosine matrices)
//   sources^T           the adjoint of an overwritten force sample
#include "ard_adjoint.h"

namespace ard {

// lambda_p[cell] += dL/dy for each recorder (several recorders may share a cell)
__global__ void k_adj_record(const RecDesc* __restrict__ r, int n, const float* __restrict__ g) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(const_cast<float*>(r[i].p) + r[i].cell, g[i]);
}

int launch_adj_record(const RecDesc* d_rec, int n, const float* d_g, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_adj_record, dim3(1), dim3(256), 0, s, d_rec, n, d_g);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Transpose of k_interface (ard_kernels.hip) for the final writer of every
// force cell (mask bits). Part tables: p[0] = adjoint of the pressure read
// (accumulated with atomics: cells near an edge are read by several
// boundaries), f = adjoint of the force written (read, then cleared: an
// overwritten force does not depend on its previous value).
__device__ __forceinline__ float* adj_cell(const PartAccess& A, bool force, int x, int y, int z) {
  float* base = force ? A.f : A.p[0];
  return base + ((long long)z * A.H + y) * A.W + x;
}

__global__ __launch_bounds__(256) void k_adj_interface(const BndDesc* __restrict__ descs, int ndesc,
                                                       const PartAccess* __restrict__ parts, long long total) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].elem0 <= e) lo = mid;
    else hi = mid - 1;
  }
  const BndDesc& B = descs[lo];
  const unsigned le = unsigned(e - B.elem0);
  const unsigned lv = le / unsigned(B.t0n);
  const int u = B.t0s + int(le - lv * unsigned(B.t0n));
  const int v = B.t1s + int(lv);
  const PartAccess& S0 = parts[B.side[0]];
  const PartAccess& S1 = parts[B.side[1]];
  const int u0 = u - B.org[0][0], v0 = v - B.org[0][1];
  const int u1 = u - B.org[1][0], v1 = v - B.org[1][1];
  const int n0 = B.type == 0 ? S0.W : (B.type == 1 ? S0.H : S0.D);
  auto cell = [&](const PartAccess& A, bool force, int nrm, int uu, int vv) {
    if (B.type == 0) return adj_cell(A, force, nrm, uu, vv);
    if (B.type == 1) return adj_cell(A, force, uu, nrm, vv);
    return adj_cell(A, force, uu, vv, nrm);
  };
  const float C[6][6] = {{0, 0, -2, 2, 0, 0},       {0, -2, 27, -27, 2, 0},  {-2, 27, -270, 270, -27, 2},
                         {2, -27, 270, -270, 27, -2}, {0, 2, -27, 27, -2, 0}, {0, 0, 2, -2, 0, 0}};
  const unsigned mask = B.mask[le];
  float gl[3] = {0.f, 0.f, 0.f}, gr[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    if (!(mask & (1u << r))) continue;
    const int m = r - 3;
    float* fc = m < 0 ? cell(S0, true, n0 + m, u0, v0) : cell(S1, true, m, u1, v1);
    const float lam = *fc;
    *fc = 0.f;
    if (m < 0) {
      // val = k0 * (self0 ? s1 + s2 : s2)
      const float w = lam * B.k[0];
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        gr[n] += C[r][n + 3] * w;
        if (B.self[0]) gl[n] += C[r][n] * w;
      }
    } else {
      // val = k1 * (self1 ? s1 + s2 : s1)
      const float w = lam * B.k[1];
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        gl[n] += C[r][n] * w;
        if (B.self[1]) gr[n] += C[r][n + 3] * w;
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    if (gl[n] != 0.f) atomicAdd(cell(S0, false, n0 - 3 + n, u0, v0), gl[n]);
    if (gr[n] != 0.f) atomicAdd(cell(S1, false, n, u1, v1), gr[n]);
  }
}

int launch_adj_interface(const BndDesc* d_descs, int ndesc, const PartAccess* d_parts, long long total,
                         hipStream_t s) {
  if (total <= 0) return 0;
  hipLaunchKernelGGL(k_adj_interface, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, d_descs, ndesc, d_parts,
                     total);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Transpose of k_pml. With one damping axis A per slab (zeta_B = 0 for B != A):
//   p+    = 2p - q + dt^2 (c0^2 L6(p) - zeta (p - q) / dt + sum_a D4_a(phi_a) + f)
//   phiA+ = phiA (1 - dt zeta) - dt zeta D4_A(p)
//   phiB+ = phiB + dt zeta D4_B(p)
// L6 (6th-order Laplacian, symmetric) and D4 (4th-order first derivative,
// antisymmetric: D4^T = -D4) read zero outside the slab. Given the adjoints
// P of p+, Phi_a of phi_a+ and Q of the carried p (the next step's q):
//   lp   = (2 - dt zeta) P + dt^2 c0^2 L6(P) - dt sum_a D4_a(kappa_a Phi_a) + Q
//   lq   = (dt zeta - 1) P
//   lphi_a = Phi_a (1 - dt zeta [a = A]) - dt^2 D4_a(P)
//   lf   = dt^2 P
// with kappa_A = -zeta, kappa_B = zeta. The damping scale theta of the slab's
// face (zeta = theta * profile) gets
//   dL/dtheta += sum P (-dt prof (p - q)) + Phi_A (-dt prof (phiA + D4_A(p))) + sum_B Phi_B dt prof D4_B(p)
// evaluated on the forward state (p, q, phi) before the step.
__global__ __launch_bounds__(256) void k_adj_pml(const AdjPmlDesc* __restrict__ descs, int ndesc, float dh, float dt,
                                                 float c0, double* __restrict__ dtheta) {
  const long long b = blockIdx.x;
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const AdjPmlDesc& P = descs[lo];
  const int W = P.W, H = P.H, D = P.D;
  const unsigned cell = unsigned(b - P.block0) * 256u + threadIdx.x;
  float g = 0.f;
  if (cell < unsigned(W * H * D)) {
    const unsigned row = cell / unsigned(W);
    const int i = int(cell - row * unsigned(W));
    const int j = int(row % unsigned(H));
    const int k = int(row / unsigned(H));
    const int A = P.axis;
    auto idx = [&](int x, int y, int z) { return ((long long)z * H + y) * W + x; };
    auto at = [&](const float* a, int x, int y, int z) -> float {
      if (x < 0 || x >= W || y < 0 || y >= H || z < 0 || z >= D) return 0.f;
      return a[idx(x, y, z)];
    };
    auto zeta_at = [&](int x, int y, int z) -> float {
      const int q = A == 0 ? x : (A == 1 ? y : z);
      const int n = A == 0 ? W : (A == 1 ? H : D);
      return (q >= 0 && q < n) ? P.zeta[q] : 0.f;
    };
    const float c6[7] = {2.f, -27.f, 270.f, -490.f, 270.f, -27.f, 2.f};
    const float c4[5] = {1.f, -8.f, 0.f, 8.f, -1.f};
    const int ax[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const float lP = P.lP[cell];
    const float zc = zeta_at(i, j, k);
    float l6 = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int m = 0; m < 7; ++m)
        l6 += c6[m] * at(P.lP, i + (m - 3) * ax[a][0], j + (m - 3) * ax[a][1], k + (m - 3) * ax[a][2]);
    l6 /= 180.f * dh * dh;
    float dkp = 0.f;  // sum_a D4_a(kappa_a Phi_a)
    float d4P[3] = {0.f, 0.f, 0.f}, d4p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        if (m == 2) continue;
        const int x = i + (m - 2) * ax[a][0], y = j + (m - 2) * ax[a][1], z = k + (m - 2) * ax[a][2];
        const float kap = (a == A ? -1.f : 1.f) * zeta_at(x, y, z);
        dkp += c4[m] * kap * at(P.lPhi[a], x, y, z);
        d4P[a] += c4[m] * at(P.lP, x, y, z);
        d4p[a] += c4[m] * at(P.p, x, y, z);
      }
    }
    const float d12 = 12.f * dh;
    dkp /= d12;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      d4P[a] /= d12;
      d4p[a] /= d12;
    }
    P.op[cell] = (2.f - dt * zc) * lP + dt * dt * c0 * c0 * l6 - dt * dkp + P.lQ[cell];
    P.oq[cell] = (dt * zc - 1.f) * lP;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      P.ophi[a][cell] = P.lPhi[a][cell] * (a == A ? 1.f - dt * zc : 1.f) - dt * dt * d4P[a];
    P.of[cell] = dt * dt * lP;
    // derivative with respect to the face's damping scale
    const float pr = P.prof[A == 0 ? i : (A == 1 ? j : k)];
    g = lP * (-dt * pr * (P.p[cell] - P.q[cell]));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == A) g += P.lPhi[a][cell] * (-dt * pr * (P.phi[a][cell] + d4p[a]));
      else g += P.lPhi[a][cell] * (dt * pr * d4p[a]);
    }
  }
  // block sum, one FP64 atomic per block
  __shared__ float s_g[4];
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) g += __shfl_xor(g, m);
  if ((threadIdx.x & 63) == 0) s_g[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(dtheta + P.face, double(s_g[0]) + double(s_g[1]) + double(s_g[2]) + double(s_g[3]));
}

int launch_adj_pml(const AdjPmlDesc* d_descs, int ndesc, long long total_blocks, float dh, float dt, float c0,
                   double* d_dtheta, hipStream_t s) {
  if (total_blocks <= 0) return 0;
  hipLaunchKernelGGL(k_adj_pml, dim3(unsigned(total_blocks)), dim3(256), 0, s, d_descs, ndesc, dh, dt, c0, d_dtheta);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Modal update transpose, elementwise over one DCT partition:
//   forward  M+ = a M - b M- + g F^,  (M, M-) <- (M+, M)
//   reverse  lM+ = t + lM;  lM <- a lM+ + lM-;  lM- <- -b lM+;  t <- g lM+
// t holds IDCT^T(lambda_p) on entry and g lM+ (the adjoint of F^) on exit.
__global__ void k_adj_modal(float* __restrict__ t, float* __restrict__ lM, float* __restrict__ lMp,
                            const float* __restrict__ a, const float* __restrict__ g, float b, long long n) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const float lnew = t[c] + lM[c];
  lM[c] = a[c] * lnew + lMp[c];
  lMp[c] = -b * lnew;
  t[c] = g[c] * lnew;
}

int launch_adj_modal(float* t, float* lM, float* lMp, const float* a, const float* g, float b, long long n,
                     hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_adj_modal, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, t, lM, lMp, a, g, b, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// lF += t (the adjoint of the force before this step's interface writes)
__global__ void k_adj_add(float* __restrict__ lF, const float* __restrict__ t, long long n) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c < n) lF[c] += t[c];
}

int launch_adj_add(float* lF, const float* t, long long n, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_adj_add, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, lF, t, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Sources^T: the sample overwrote its cell, so dL/du = lambda_F there and the
// cell's adjoint is cleared. One thread walks the sources in reverse
// registration order, so of two sources on one cell the last writer gets it.
__global__ void k_adj_sources(const AdjSource* __restrict__ s, int n, float* __restrict__ du, int step, int nsteps) {
  if (threadIdx.x != 0) return;
  for (int i = n - 1; i >= 0; --i) {
    float* f = s[i].lF + s[i].cell;
    du[(long long)s[i].index * nsteps + step] = *f;
    *f = 0.f;
  }
}

int launch_adj_sources(const AdjSource* d_src, int n, float* d_du, int step, int nsteps, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_adj_sources, dim3(1), dim3(64), 0, s, d_src, n, d_du, step, nsteps);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace ard
