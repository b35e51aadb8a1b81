// HIP kernels of the ARD time step for gfx950 (MI355X).
//
// Plain form (one partition per launch, the reference's structure):
//   k_gemm         per-axis DCT-II / DCT-III passes as strided GEMMs against the
//                  cosine matrix of the axis, LDS-tiled, 64x64 tiles per 256 threads;
//                  the last forward pass fuses the modal update
//                  (ARD/dct_volume.cpp:38-68, ARD/dct_partition.cpp:48-73).
//   k_pml_tiled    Grote-Sim FDTD update of PML slabs (ARD/pml_partition.cpp:113-200),
//                  LDS-tiled through the slab's thickness (3 to 8 cells);
//   k_pml          the same update one thread per cell, for thicker slabs.
//   k_interface    6th-order interface forcing (ARD/boundary.cpp:21-173).
//   k_sources      source forcing (ARD/partition.cpp:187-197, ARD/gaussian_source.cpp:16-20).
//   k_record       recorder point samples (ARD/recorder.cpp:40-57) into a device trace.
// All kernels accept a table of problems so that one launch can also cover
// every partition of a phase (the batched form).
#include <type_traits>

#include "ard_kernels.h"

namespace ard {

// ---------------------------------------------------------------------------
// Grouped strided GEMM: 64x64 output tile per 256-thread block, K in chunks
// of 16 staged in LDS (two buffers; the next chunk is loaded into registers
// during the current one), 4x4 outputs per thread.
// ---------------------------------------------------------------------------
template <bool B_K_CONTIG, bool C_M_CONTIG, int EP>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc* __restrict__ descs, int ndesc) {
  __shared__ float As[2][16][68];
  __shared__ float Bs[2][16][68];
  const int tile = xcd_block(int(blockIdx.x), int(gridDim.x));  // contiguous tile ranges per XCD
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmDesc& g = descs[lo];
  const int M = g.M, N = g.N, K = g.K, N1 = g.N1;
  const int t = tile - g.tile0;
  // m tiles fastest: the tiles of one B panel run together and share it in L2
  const int tiles_m = (M + 63) >> 6;
  const int m0 = (t % tiles_m) * 64;
  const int n0 = (t / tiles_m) * 64;
  const int tid = threadIdx.x;

  // B element offsets of the columns this thread loads (fixed over k)
  long long boff[4];
  int bn[4];
  if (B_K_CONTIG) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = n0 + tid / 16 + 16 * r;
      bn[r] = n;
      boff[r] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : 0;
    }
  } else {
    const int n = n0 + (tid & 63);
    bn[0] = n;
    boff[0] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : 0;
  }
  const int tm = C_M_CONTIG ? (tid & 15) : (tid >> 4);
  const int tn = C_M_CONTIG ? (tid >> 4) : (tid & 15);

  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;

  const gfloat* __restrict__ A = ARD_CG(g.A);
  const gfloat* __restrict__ B = ARD_CG(g.B);
  // the next K chunk is loaded into registers while the current one is
  // multiplied from LDS (two LDS buffers, one barrier per chunk)
  float ra[4], rb[4];
  auto gload = [&](const int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = tid & 15, m = (tid >> 4) + 16 * r;
      ra[r] = A[(long long)min(m0 + m, M - 1) * K + min(k0 + k, K - 1)];  // clamped: loads stay unconditional
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gk = B_K_CONTIG ? k0 + (tid & 15) : k0 + (tid >> 6) + 4 * r;
      rb[r] = B[(long long)min(gk, K - 1) * g.bK + boff[B_K_CONTIG ? r : 0]];
    }
  };
  auto sstore = [&](const int k0, const int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = tid & 15, m = (tid >> 4) + 16 * r;
      As[buf][k][m] = (m0 + m < M && k0 + k < K) ? ra[r] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (B_K_CONTIG) {
        const int k = tid & 15, n = (tid >> 4) + 16 * r;
        Bs[buf][k][n] = (bn[r] < N && k0 + k < K) ? rb[r] : 0.f;
      } else {
        const int k = (tid >> 6) + 4 * r, n = tid & 63;
        Bs[buf][k][n] = (bn[0] < N && k0 + k < K) ? rb[r] : 0.f;
      }
    }
  };
  gload(0);
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += 16, buf ^= 1) {
    sstore(k0, buf);
    __syncthreads();
    if (k0 + 16 < K) gload(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const float4 a = *reinterpret_cast<const float4*>(&As[buf][kk][4 * tm]);
      const float4 b = *reinterpret_cast<const float4*>(&Bs[buf][kk][4 * tn]);
      const float av[4] = {a.x, a.y, a.z, a.w};
      const float bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
    }
  }

  long long coff[4];
  bool nok[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + 4 * tn + j;
    nok[j] = n < N;
    coff[j] = nok[j] ? (long long)(n % N1) * g.c1 + (long long)(n / N1) * g.c2 : 0;
  }
  gfloat* __restrict__ C = ARD_G(g.C);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * tm + i;
    if (m >= M) continue;
    const long long mo = (long long)m * g.cM;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!nok[j]) continue;
      const long long o = mo + coff[j];
      const float v = g.alpha * acc[i][j];
      if (EP == 0) {
        C[o] = v;
      } else {
        C[o] = ARD_CG(g.ep_a)[o] * ARD_CG(g.ep_cur)[o] - g.ep_b * C[o] + ARD_CG(g.ep_g)[o] * v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The same grouped GEMM on the matrix cores: each of the 4 waves owns a 32x32
// quarter of the 64x64 tile and accumulates it with v_mfma_f32_32x32x2_f32
// (FP32 in, FP32 accumulate). Operand layout of the instruction: lane l holds
// A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator register r of lane l is
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <bool B_K_CONTIG, int EP>
__global__ __launch_bounds__(256) void k_gemm_mfma(const GemmDesc* __restrict__ descs, int ndesc) {
  __shared__ float As[2][16][68];
  __shared__ float Bs[2][16][68];
  const int tile = xcd_block(int(blockIdx.x), int(gridDim.x));  // contiguous tile ranges per XCD
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmDesc& g = descs[lo];
  const int M = g.M, N = g.N, K = g.K, N1 = g.N1;
  const int t = tile - g.tile0;
  // m tiles fastest: the tiles of one B panel run together and share it in L2
  const int tiles_m = (M + 63) >> 6;
  const int m0 = (t % tiles_m) * 64;
  const int n0 = (t / tiles_m) * 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  long long boff[4];
  int bn[4];
  if (B_K_CONTIG) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = n0 + tid / 16 + 16 * r;
      bn[r] = n;
      boff[r] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : 0;
    }
  } else {
    const int n = n0 + (tid & 63);
    bn[0] = n;
    boff[0] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : 0;
  }
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;

  const gfloat* __restrict__ A = ARD_CG(g.A);
  const gfloat* __restrict__ B = ARD_CG(g.B);
  float ra[4], rb[4];
  auto gload = [&](const int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = tid & 15, m = (tid >> 4) + 16 * r;
      ra[r] = A[(long long)min(m0 + m, M - 1) * K + min(k0 + k, K - 1)];  // clamped: loads stay unconditional
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gk = B_K_CONTIG ? k0 + (tid & 15) : k0 + (tid >> 6) + 4 * r;
      rb[r] = B[(long long)min(gk, K - 1) * g.bK + boff[B_K_CONTIG ? r : 0]];
    }
  };
  auto sstore = [&](const int k0, const int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = tid & 15, m = (tid >> 4) + 16 * r;
      As[buf][k][m] = (m0 + m < M && k0 + k < K) ? ra[r] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (B_K_CONTIG) {
        const int k = tid & 15, n = (tid >> 4) + 16 * r;
        Bs[buf][k][n] = (bn[r] < N && k0 + k < K) ? rb[r] : 0.f;
      } else {
        const int k = (tid >> 6) + 4 * r, n = tid & 63;
        Bs[buf][k][n] = (bn[0] < N && k0 + k < K) ? rb[r] : 0.f;
      }
    }
  };
  gload(0);
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += 16, buf ^= 1) {
    sstore(k0, buf);
    __syncthreads();
    if (k0 + 16 < K) gload(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; kk += 2) {
      const float a = As[buf][kk + (lane >> 5)][wm + (lane & 31)];
      const float b = Bs[buf][kk + (lane >> 5)][wn + (lane & 31)];
      // K-contiguous B: in the transform passes C is then contiguous along
      // m, so the instruction's columns (lane % 32) take m: C^T = B^T A^T
      acc = B_K_CONTIG ? __builtin_amdgcn_mfma_f32_32x32x2f32(b, a, acc, 0, 0, 0)
                       : __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
  }
  // accumulator r of this lane: (m, n) = (wm + lane % 32, wn + row(r)) when
  // B is K-contiguous, else (wm + row(r), wn + lane % 32), with
  // row(r) = 8 (r / 4) + 4 (lane / 32) + r % 4
  const int jl = (B_K_CONTIG ? m0 + wm : n0 + wn) + (lane & 31);
  if (jl >= (B_K_CONTIG ? M : N)) return;
  long long cn = 0;
  if (!B_K_CONTIG) cn = (long long)(jl % N1) * g.c1 + (long long)(jl / N1) * g.c2;
  gfloat* __restrict__ C = ARD_G(g.C);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int il = (B_K_CONTIG ? n0 + wn : m0 + wm) + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (il >= (B_K_CONTIG ? N : M)) continue;
    const int i = B_K_CONTIG ? jl : il;
    if (B_K_CONTIG) cn = (long long)(il % N1) * g.c1 + (long long)(il / N1) * g.c2;
    const long long o = (long long)i * g.cM + cn;
    const float v = g.alpha * acc[r];
    if (EP == 0) C[o] = v;
    else C[o] = ARD_CG(g.ep_a)[o] * ARD_CG(g.ep_cur)[o] - g.ep_b * C[o] + ARD_CG(g.ep_g)[o] * v;
  }
}

// ---------------------------------------------------------------------------
// The same grouped GEMM in BF16 split precision on the matrix cores: every
// FP32 operand is split into x = hi + lo (hi = bf16(x), lo = bf16(x - hi)) and
//   A B ~ A_hi B_hi + A_hi B_lo + A_lo B_hi
// with v_mfma_f32_32x32x8_bf16 and FP32 accumulation (relative error of a
// product about 2^-16). Operand layout of the instruction: lane l holds
// A[l % 32][4 (l / 32) + i] and B[4 (l / 32) + i][l % 32], i < 4; the
// accumulator layout is that of v_mfma_f32_32x32x2_f32.
// ---------------------------------------------------------------------------
typedef short s16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rn(float x) {  // round to nearest even (finite inputs)
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return short(u >> 16);
}
__device__ __forceinline__ float bf16_f(short h) { return __uint_as_float(unsigned(static_cast<unsigned short>(h)) << 16); }

template <bool B_K_CONTIG, int EP>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(const GemmDesc* __restrict__ descs, int ndesc) {
  __shared__ float As[2][16][68];
  __shared__ float Bs[2][16][68];
  const int tile = xcd_block(int(blockIdx.x), int(gridDim.x));  // contiguous tile ranges per XCD
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  const GemmDesc& g = descs[lo];
  const int M = g.M, N = g.N, K = g.K, N1 = g.N1;
  const int t = tile - g.tile0;
  // m tiles fastest: the tiles of one B panel run together and share it in L2
  const int tiles_m = (M + 63) >> 6;
  const int m0 = (t % tiles_m) * 64;
  const int n0 = (t / tiles_m) * 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  long long boff[4];
  int bn[4];
  if (B_K_CONTIG) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = n0 + tid / 16 + 16 * r;
      bn[r] = n;
      boff[r] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : 0;
    }
  } else {
    const int n = n0 + (tid & 63);
    bn[0] = n;
    boff[0] = n < N ? (long long)(n % N1) * g.b1 + (long long)(n / N1) * g.b2 : 0;
  }
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;

  const gfloat* __restrict__ A = ARD_CG(g.A);
  const gfloat* __restrict__ B = ARD_CG(g.B);
  float ra[4], rb[4];
  auto gload = [&](const int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = tid & 15, m = (tid >> 4) + 16 * r;
      ra[r] = A[(long long)min(m0 + m, M - 1) * K + min(k0 + k, K - 1)];  // clamped: loads stay unconditional
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gk = B_K_CONTIG ? k0 + (tid & 15) : k0 + (tid >> 6) + 4 * r;
      rb[r] = B[(long long)min(gk, K - 1) * g.bK + boff[B_K_CONTIG ? r : 0]];
    }
  };
  auto sstore = [&](const int k0, const int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = tid & 15, m = (tid >> 4) + 16 * r;
      As[buf][k][m] = (m0 + m < M && k0 + k < K) ? ra[r] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (B_K_CONTIG) {
        const int k = tid & 15, n = (tid >> 4) + 16 * r;
        Bs[buf][k][n] = (bn[r] < N && k0 + k < K) ? rb[r] : 0.f;
      } else {
        const int k = (tid >> 6) + 4 * r, n = tid & 63;
        Bs[buf][k][n] = (bn[0] < N && k0 + k < K) ? rb[r] : 0.f;
      }
    }
  };
  gload(0);
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += 16, buf ^= 1) {
    sstore(k0, buf);
    __syncthreads();
    if (k0 + 16 < K) gload(k0 + 16);
#pragma unroll
    for (int ks = 0; ks < 16; ks += 8) {
      const int kb = ks + 4 * (lane >> 5);
      s16x4 ah, al, bh, bl;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float a = As[buf][kb + i][wm + (lane & 31)];
        const float b = Bs[buf][kb + i][wn + (lane & 31)];
        ah[i] = bf16_rn(a);
        al[i] = bf16_rn(a - bf16_f(ah[i]));
        bh[i] = bf16_rn(b);
        bl[i] = bf16_rn(b - bf16_f(bh[i]));
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ah, bh, acc, 0, 0, 0);
    }
  }
  const int j = n0 + wn + (lane & 31);
  if (j >= N) return;
  const long long cn = (long long)(j % N1) * g.c1 + (long long)(j / N1) * g.c2;
  gfloat* __restrict__ C = ARD_G(g.C);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (i >= M) continue;
    const long long o = (long long)i * g.cM + cn;
    const float v = g.alpha * acc[r];
    if (EP == 0) C[o] = v;
    else C[o] = ARD_CG(g.ep_a)[o] * ARD_CG(g.ep_cur)[o] - g.ep_b * C[o] + ARD_CG(g.ep_g)[o] * v;
  }
}

// ---------------------------------------------------------------------------
// Even/odd folded per-axis transforms on the matrix cores (FP32). With
// c_L(k, L-1-j) = (-1)^k c_L(k, j):
//   forward (REDFT10)  X[2m]   = sum_{j<ceil(L/2)} C[2m][j]   (x[j] + x[L-1-j])
//                      X[2m+1] = sum_{j<floor(L/2)} C[2m+1][j] (x[j] - x[L-1-j])
//   inverse (REDFT01)  E[j] = sum_k' Ci[j][2k'] X[2k'],  O[j] = sum_k' Ci[j][2k'+1] X[2k'+1]
//                      x[j] = E + O,  x[L-1-j] = E - O
// (the middle element of an odd L counts once). Half the multiply-adds of the
// dense pass. Forward: rows [0, Mh) of the folded matrix are the even outputs
// (Mh a multiple of 32, so each wave's 32 rows are all even or all odd); B is
// staged twice, for the tile's first and second 32 rows, each as the sum (even
// rows) or the difference (odd rows). Inverse: a tile covers 32 outputs j; its
// first 32 A rows are A_E and the next 32 A_O rows of those j (waves wm = 0
// accumulate E, wm = 32 O), so M counts 64 rows per 32 j; E and O meet
// through LDS and the E waves write x[j], the O waves x[L-1-j].
// ---------------------------------------------------------------------------
template <bool B_K_CONTIG, int FOLD, int EP>
__global__ __launch_bounds__(256) void k_gemm_mfma_fold(const GemmDesc* __restrict__ descs,
                                                        const int2* __restrict__ tiles) {
  // rows of 66: the staging stores (four k rows per quad of lanes) fall in
  // distinct banks, as do the operand reads (one row per half-wave)
  __shared__ float As[2][16][66];
  __shared__ float Bs[2][2][16][66];  // inverse: even, odd rows of X; forward: sums, differences
  // the launch's tile table: (descriptor, tile within it), longest K first,
  // so the long tiles start early and the short ones fill in at the end
  const int2 tm = tiles[blockIdx.x];
  const GemmDesc& g = descs[tm.x];
  const int M = g.M, N = g.N, K = g.K, N1 = g.N1, L = g.L, Kp = g.Kp;
  const int t = tm.y;
  // m tiles fastest: the tiles of one B panel run together and share it in L2
  const int tiles_m = (M + 63) >> 6;
  const int m0 = (t % tiles_m) * 64;
  const int n0 = (t / tiles_m) * 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // B set h feeds the tile's rows m0 + 32 h .. +31 (waves wm = 32 h). Forward:
  // x[j] + sg_h x[L-1-j], sg_h = -1 for odd rows (uniform per block); inverse:
  // set 0 the even rows of X (for E), set 1 the odd ones (for O)
  const float sg0 = FOLD == 1 && m0 >= g.Mh ? -1.f : 1.f;
  const float sg1 = FOLD == 1 && m0 + 32 >= g.Mh ? -1.f : 1.f;
  const int bsel = wm >> 5;

  // Operands through buffer resources: A is zero-padded to whole 64 x 16
  // tiles, so its loads need no clamps; B loads past the operand array read
  // 0 (range-checked), and B elements past K or N meet zero A columns or
  // are never stored, so no selects either.
  const brsrc rA = make_rsrc(g.A, unsigned((FOLD == 2 ? 2 : 1) * g.Ap * Kp) * 4u);
  // B's 16-byte loads that hold needed elements end at most 3 floats past a
  // row, so past the array's last row too: B arrays are followed by at least
  // 4 floats of the same allocation (the face buffers are carved from one block)
  const brsrc rB = make_rsrc(g.B, (g.b_elems + 4u) * 4u);
  const int bK = int(g.bK);
  // A: thread (tile row am = tid / 4, k = 4 (tid % 4) .. +3), one 16-byte load
  // per chunk (byte offsets at k0 = 0; 16-byte loads need only 4-byte
  // alignment). Inverse: tile row am is row j0 + am % 32 of A_E (am < 32) or
  // A_O, the tables [2][Ap][Kp]
  const int am = tid >> 2;
  const int arow = FOLD == 2 ? (am >> 5) * g.Ap + (m0 >> 1) + (am & 31) : m0 + am;
  const unsigned aoff = unsigned(arow * Kp + 4 * (tid & 3)) * 4u;
  // B(k, n) = B[brow(k) * bK + col(n)]: brow(k) = k, L-1-k (forward), 2k + w (inverse)
  // K-contiguous B: thread (n = tid / 4, k = 4 (tid % 4) .. +3), 16-byte loads
  int bcol[4];
  if (B_K_CONTIG) {
    const int n = min(n0 + (tid >> 2), N - 1);
    bcol[0] = int((n % N1) * g.b1 + (n / N1) * g.b2);
  } else {
    const int n = min(n0 + lane, N - 1);
    bcol[0] = int((n % N1) * g.b1 + (n / N1) * g.b2);
  }
  const int kt = B_K_CONTIG ? (tid & 15) : (tid >> 6);  // this thread's k within a chunk (+ 4 r if !B_K_CONTIG)
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;

  v4f ra;
  float rb[2][4];  // K-contiguous B: [w][i] for k = 4 (tid % 4) + i; else [w][r] for k = tid / 64 + 4 r
  auto gload = [&](const int k0) {
    ra = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rA, int(aoff), int(k0 * 4), 0));
    if (B_K_CONTIG) {
      const int k = k0 + 4 * (tid & 3);
      if (FOLD == 1) {
        // x[k .. k+3] and x[L-4-k .. L-1-k] (reversed); a 16-byte load at a
        // negative offset reads 0 as a whole, which loses no needed element
        // (k < ceil(L/2)) when L >= 6, the smallest partition side
        const v4f d = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol[0] + k) * 4u), 0, 0));
        const v4f m = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol[0] + L - 4 - k) * 4u), 0, 0));
        rb[0][0] = d.x; rb[0][1] = d.y; rb[0][2] = d.z; rb[0][3] = d.w;
        rb[1][0] = m.w; rb[1][1] = m.z; rb[1][2] = m.y; rb[1][3] = m.x;
      } else {
        // rows 2k .. 2k+7: (k, E), (k, O), (k+1, E), ...
        const v4f f0 = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol[0] + 2 * k) * 4u), 0, 0));
        const v4f f1 = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol[0] + 2 * k + 4) * 4u), 0, 0));
        rb[0][0] = f0.x; rb[1][0] = f0.y; rb[0][1] = f0.z; rb[1][1] = f0.w;
        rb[0][2] = f1.x; rb[1][2] = f1.y; rb[0][3] = f1.z; rb[1][3] = f1.w;
      }
    } else {
#pragma unroll
      for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = k0 + kt + 4 * r;
          const int row = FOLD == 1 ? (w == 0 ? k : L - 1 - k) : 2 * k + w;  // negative: offset past the end, reads 0
          rb[w][r] = bload(rB, unsigned(row * bK + bcol[0]) * 4u, 0u);
        }
    }
  };
  auto sstore = [&](const int buf) {
    const int ka = 4 * (tid & 3), ma = tid >> 2;
    As[buf][ka + 0][ma] = ra.x;
    As[buf][ka + 1][ma] = ra.y;
    As[buf][ka + 2][ma] = ra.z;
    As[buf][ka + 3][ma] = ra.w;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = B_K_CONTIG ? ka + r : (tid >> 6) + 4 * r;
      const int n = B_K_CONTIG ? ma : (tid & 63);
      if (FOLD == 1) {  // x[j] +- x[L-1-j], rounded once
        Bs[buf][0][k][n] = rb[0][r] + sg0 * rb[1][r];
        Bs[buf][1][k][n] = rb[0][r] + sg1 * rb[1][r];
      } else {
        Bs[buf][0][k][n] = rb[0][r];
        Bs[buf][1][k][n] = rb[1][r];
      }
    }
  };
  auto mma = [&](const int buf, auto steps) {
#pragma unroll
    for (int kk = 0; kk < 2 * decltype(steps)::value; kk += 2) {  // past K: zero-padded A columns
      const float a = As[buf][kk + (lane >> 5)][wm + (lane & 31)];
      const float b = Bs[buf][bsel][kk + (lane >> 5)][wn + (lane & 31)];
      // K-contiguous B: the output is contiguous along m, so the
      // instruction's columns (lane % 32) take m: C^T = B^T A^T
      acc = B_K_CONTIG ? __builtin_amdgcn_mfma_f32_32x32x2f32(b, a, acc, 0, 0, 0)
                       : __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
  };
  gload(0);
  int buf = 0, k0 = 0;
  for (; k0 + 8 < K; k0 += 16, buf ^= 1) {  // chunks with more than 8 columns of K left
    sstore(buf);
    __syncthreads();
    if (k0 + 16 < K) gload(k0 + 16);
    mma(buf, std::integral_constant<int, 8>());
  }
  if (k0 < K) {  // a last chunk of at most 8 columns: its second half meets zero A columns only
    sstore(buf);
    __syncthreads();
    mma(buf, std::integral_constant<int, 4>());
  }
  // inverse: the E and O waves of the same outputs exchange accumulators
  // ([set][wn / 32][r][lane] in the B tiles, free once every wave is done)
  f32x16 oth;
  if (FOLD == 2) {
    __syncthreads();
    float* __restrict__ xs = &Bs[0][0][0][0];  // 4096 of its 4224 floats
#pragma unroll
    for (int r = 0; r < 16; ++r) xs[((bsel * 2 + (wn >> 5)) * 16 + r) * 64 + lane] = acc[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) oth[r] = xs[(((bsel ^ 1) * 2 + (wn >> 5)) * 16 + r) * 64 + lane];
  }
  gfloat* __restrict__ C = ARD_G(g.C);
  // accumulator r of this lane: (m, n) = (mb + lane % 32, wn + row(r)) when
  // B is K-contiguous, else (mb + row(r), wn + lane % 32), with
  // row(r) = 8 (r / 4) + 4 (lane / 32) + r % 4 and the wave's first output
  // row mb = m0 + wm (forward) or j0 = m0 / 2 (inverse, both waves)
  const int mb = FOLD == 2 ? (m0 >> 1) : m0 + wm;
  const int mlim = FOLD == 2 ? (L + 1) / 2 : M;
  const int jl = (B_K_CONTIG ? mb : n0 + wn) + (lane & 31);
  if (jl >= (B_K_CONTIG ? mlim : N)) return;
  long long cn = 0;
  if (!B_K_CONTIG) cn = (long long)(jl % N1) * g.c1 + (long long)(jl / N1) * g.c2;
  // K-contiguous B: the lane's 16 columns are nb + 8 (r / 4) + r % 4; one
  // quotient and remainder by N1 for nb (a reciprocal estimate, corrected),
  // then each column's by stepping the remainder past N1
  const int nb = n0 + wn + 4 * (lane >> 5);
  int nq = 0, nr = 0;
  if (B_K_CONTIG) {
    nq = int(float(nb) * (1.f / float(N1)));
    nr = nb - nq * N1;
    while (nr < 0) { nr += N1; --nq; }
    while (nr >= N1) { nr -= N1; ++nq; }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int il = (B_K_CONTIG ? n0 + wn : mb) + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (il >= (B_K_CONTIG ? N : mlim)) continue;
    const int i = B_K_CONTIG ? jl : il;
    if (B_K_CONTIG) {
      int q = nq, rr = nr + 8 * (r >> 2) + (r & 3);  // rr < N1 + 28
      if (N1 >= 28) {  // uniform: one step at most
        const bool w = rr >= N1;
        rr -= w ? N1 : 0;
        q += w ? 1 : 0;
      } else {
        while (rr >= N1) { rr -= N1; ++q; }
      }
      cn = (long long)rr * g.c1 + (long long)q * g.c2;
    }
    if (FOLD == 1) {
      const int k = i < g.Mh ? 2 * i : 2 * (i - g.Mh) + 1;
      if (k < L) {
        const long long o = (long long)k * g.cM + cn;
        float v = g.alpha * acc[r];
        // modal epilogue (the plain form's last forward pass): C holds the
        // previous modes and receives the next, as in k_gemm_mfma
        if (EP == 1) v = ARD_CG(g.ep_a)[o] * ARD_CG(g.ep_cur)[o] - g.ep_b * C[o] + ARD_CG(g.ep_g)[o] * v;
        C[o] = v;
      }
    } else {
      const float e = bsel == 0 ? acc[r] : oth[r], o = bsel == 0 ? oth[r] : acc[r];
      if (bsel == 0) C[(long long)i * g.cM + cn] = g.alpha * (e + o);
      else if (L - 1 - i != i) C[(long long)(L - 1 - i) * g.cM + cn] = g.alpha * (e - o);
    }
  }
}

// ---------------------------------------------------------------------------
// The even/odd folded transforms in BF16 split precision on gfx950's
// v_mfma_f32_32x32x16_bf16: x = hi + lo in BF16, A B ~ A_hi B_hi + A_hi B_lo
// + A_lo B_hi, FP32 accumulation (a product's relative error about 2^-16),
// three instructions per 16 columns of the folded K against eight FP32
// v_mfma_f32_32x32x2_f32. Operand layout of the instruction: lane l holds
// A[l % 32][8 (l / 32) + i] and B[8 (l / 32) + i][l % 32], i < 8; the
// accumulator layout is that of the other 32 x 32 instructions.
// A (the folded cosine tables, g.A / g.Alo: BF16 hi and lo parts split on the
// host, [Ap][Kp] forward or [2][Ap][Kp] inverse) goes from memory straight
// into the operand registers (16 bytes per lane and part); B is folded and
// split once while staged in LDS as [n][k] rows (24 halfwords: the 16-byte
// operand reads of one phase land in distinct banks). Tiles, B addressing,
// the tile table and the outputs are those of k_gemm_mfma_fold.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;

template <bool B_K_CONTIG, int FOLD>
__global__ __launch_bounds__(256) void k_gemm_bf16x3_fold(const GemmDesc* __restrict__ descs,
                                                          const int2* __restrict__ tiles) {
  // [hi / lo][buffer][set][n][k]: 24 KB, also the inverse's exchange space
  __shared__ __attribute__((aligned(16))) u16 Bs[2][2][2][64][24];
  auto& Bh = Bs[0];
  auto& Bl = Bs[1];
  const int2 tm = tiles[blockIdx.x];
  const GemmDesc& g = descs[tm.x];
  const int M = g.M, N = g.N, K = g.K, N1 = g.N1, L = g.L, Kp = g.Kp;
  const int t = tm.y;
  const int tiles_m = (M + 63) >> 6;
  const int m0 = (t % tiles_m) * 64;
  const int n0 = (t / tiles_m) * 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const float sg0 = FOLD == 1 && m0 >= g.Mh ? -1.f : 1.f;
  const float sg1 = FOLD == 1 && m0 + 32 >= g.Mh ? -1.f : 1.f;
  const int bsel = wm >> 5;
  // A: this lane's operand row of the wave's 32 rows, 8 k per 16-column step
  const int arow = FOLD == 2 ? bsel * g.Ap + (m0 >> 1) + (lane & 31) : m0 + wm + (lane & 31);
  const unsigned aoff = unsigned(arow * Kp + 8 * (lane >> 5)) * 2u;
  const brsrc rAh = make_rsrc(g.A, unsigned((FOLD == 2 ? 2 : 1) * g.Ap * Kp) * 2u);
  const brsrc rAl = make_rsrc(g.Alo, unsigned((FOLD == 2 ? 2 : 1) * g.Ap * Kp) * 2u);
  const brsrc rB = make_rsrc(g.B, (g.b_elems + 4u) * 4u);
  const int bK = int(g.bK);
  int bcol;
  if (B_K_CONTIG) {
    const int n = min(n0 + (tid >> 2), N - 1);
    bcol = int((n % N1) * g.b1 + (n / N1) * g.b2);
  } else {
    const int n = min(n0 + lane, N - 1);
    bcol = int((n % N1) * g.b1 + (n / N1) * g.b2);
  }
  const int kt = tid >> 6;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;

  v4f rah, ral;     // this lane's A operand (hi, lo) of the next step
  float rb[2][4];   // B elements, as in k_gemm_mfma_fold
  auto gload = [&](const int k0) {
    rah = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rAh, int(aoff), int(k0 * 2), 0));
    ral = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rAl, int(aoff), int(k0 * 2), 0));
    if (B_K_CONTIG) {
      const int k = k0 + 4 * (tid & 3);
      if (FOLD == 1) {
        const v4f d = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol + k) * 4u), 0, 0));
        const v4f m = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol + L - 4 - k) * 4u), 0, 0));
        rb[0][0] = d.x; rb[0][1] = d.y; rb[0][2] = d.z; rb[0][3] = d.w;
        rb[1][0] = m.w; rb[1][1] = m.z; rb[1][2] = m.y; rb[1][3] = m.x;
      } else {
        const v4f f0 = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol + 2 * k) * 4u), 0, 0));
        const v4f f1 = __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(rB, int(unsigned(bcol + 2 * k + 4) * 4u), 0, 0));
        rb[0][0] = f0.x; rb[1][0] = f0.y; rb[0][1] = f0.z; rb[1][1] = f0.w;
        rb[0][2] = f1.x; rb[1][2] = f1.y; rb[0][3] = f1.z; rb[1][3] = f1.w;
      }
    } else {
#pragma unroll
      for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = k0 + kt + 4 * r;
          const int row = FOLD == 1 ? (w == 0 ? k : L - 1 - k) : 2 * k + w;
          rb[w][r] = bload(rB, unsigned(row * bK + bcol) * 4u, 0u);
        }
    }
  };
  // B staged as [set][n][k], split into hi / lo once
  auto sstore = [&](const int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = B_K_CONTIG ? 4 * (tid & 3) + r : kt + 4 * r;
      const int n = B_K_CONTIG ? (tid >> 2) : lane;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float v = FOLD == 1 ? rb[0][r] + (h ? sg1 : sg0) * rb[1][r] : rb[h][r];
        const short hi = bf16_rn(v);
        Bh[buf][h][n][k] = u16(hi);
        Bl[buf][h][n][k] = u16(bf16_rn(v - bf16_f(hi)));
      }
    }
  };
  gload(0);
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += 16, buf ^= 1) {
    const v4f ah = rah, al = ral;
    sstore(buf);
    __syncthreads();
    if (k0 + 16 < K) gload(k0 + 16);
    const v4f bh = *reinterpret_cast<const v4f*>(&Bh[buf][bsel][wn + (lane & 31)][8 * (lane >> 5)]);
    const v4f bl = *reinterpret_cast<const v4f*>(&Bl[buf][bsel][wn + (lane & 31)][8 * (lane >> 5)]);
    const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah), Al = __builtin_bit_cast(bf16x8, al);
    const bf16x8 Bhh = __builtin_bit_cast(bf16x8, bh), Bll = __builtin_bit_cast(bf16x8, bl);
    // K-contiguous B: the output is contiguous along m, so the instruction's
    // columns take m: C^T = B^T A^T
    if (B_K_CONTIG) {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bll, Ah, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bhh, Al, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bhh, Ah, acc, 0, 0, 0);
    } else {
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Bhh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bll, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bhh, acc, 0, 0, 0);
    }
  }
  // outputs exactly as k_gemm_mfma_fold
  f32x16 oth;
  if (FOLD == 2) {
    __syncthreads();
    float* __restrict__ xs = reinterpret_cast<float*>(&Bs[0][0][0][0][0]);  // 4096 of its 6144 floats
#pragma unroll
    for (int r = 0; r < 16; ++r) xs[((bsel * 2 + (wn >> 5)) * 16 + r) * 64 + lane] = acc[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) oth[r] = xs[(((bsel ^ 1) * 2 + (wn >> 5)) * 16 + r) * 64 + lane];
  }
  gfloat* __restrict__ C = ARD_G(g.C);
  const int mb = FOLD == 2 ? (m0 >> 1) : m0 + wm;
  const int mlim = FOLD == 2 ? (L + 1) / 2 : M;
  const int jl = (B_K_CONTIG ? mb : n0 + wn) + (lane & 31);
  if (jl >= (B_K_CONTIG ? mlim : N)) return;
  long long cn = 0;
  if (!B_K_CONTIG) cn = (long long)(jl % N1) * g.c1 + (long long)(jl / N1) * g.c2;
  const int nb = n0 + wn + 4 * (lane >> 5);
  int nq = 0, nr = 0;
  if (B_K_CONTIG) {
    nq = int(float(nb) * (1.f / float(N1)));
    nr = nb - nq * N1;
    while (nr < 0) { nr += N1; --nq; }
    while (nr >= N1) { nr -= N1; ++nq; }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int il = (B_K_CONTIG ? n0 + wn : mb) + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (il >= (B_K_CONTIG ? N : mlim)) continue;
    const int i = B_K_CONTIG ? jl : il;
    if (B_K_CONTIG) {
      int q = nq, rr = nr + 8 * (r >> 2) + (r & 3);
      while (rr >= N1) { rr -= N1; ++q; }
      cn = (long long)rr * g.c1 + (long long)q * g.c2;
    }
    if (FOLD == 1) {
      const int k = i < g.Mh ? 2 * i : 2 * (i - g.Mh) + 1;
      if (k < L) C[(long long)k * g.cM + cn] = g.alpha * acc[r];
    } else {
      const float e = bsel == 0 ? acc[r] : oth[r], o = bsel == 0 ? oth[r] : acc[r];
      if (bsel == 0) C[(long long)i * g.cM + cn] = g.alpha * (e + o);
      else if (L - 1 - i != i) C[(long long)(L - 1 - i) * g.cM + cn] = g.alpha * (e - o);
    }
  }
}

int launch_gemm_bf16_fold(const GemmDesc* d_descs, const int2* d_tiles, int total_tiles, bool b_k_contig, int fold,
                          hipStream_t s) {
  if (total_tiles <= 0) return 0;
  dim3 grid(total_tiles), block(256);
  if (fold == 1 && b_k_contig) hipLaunchKernelGGL((k_gemm_bf16x3_fold<true, 1>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 1) hipLaunchKernelGGL((k_gemm_bf16x3_fold<false, 1>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 2 && b_k_contig) hipLaunchKernelGGL((k_gemm_bf16x3_fold<true, 2>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 2) hipLaunchKernelGGL((k_gemm_bf16x3_fold<false, 2>), grid, block, 0, s, d_descs, d_tiles);
  else return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_gemm_fold(const GemmDesc* d_descs, const int2* d_tiles, int total_tiles, bool b_k_contig, int fold,
                     hipStream_t s, int ep) {
  if (total_tiles <= 0) return 0;
  dim3 grid(total_tiles), block(256);
  if (ep == 1 && fold == 1 && !b_k_contig)
    hipLaunchKernelGGL((k_gemm_mfma_fold<false, 1, 1>), grid, block, 0, s, d_descs, d_tiles);
  else if (ep != 0) return -1;
  else if (fold == 1 && b_k_contig) hipLaunchKernelGGL((k_gemm_mfma_fold<true, 1, 0>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 1) hipLaunchKernelGGL((k_gemm_mfma_fold<false, 1, 0>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 2 && b_k_contig) hipLaunchKernelGGL((k_gemm_mfma_fold<true, 2, 0>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 2) hipLaunchKernelGGL((k_gemm_mfma_fold<false, 2, 0>), grid, block, 0, s, d_descs, d_tiles);
  else return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// variant: 0 vector FMA, 1 matrix cores (FP32), 2 matrix cores (BF16 split, FP32 accumulate),
// 3 matrix cores FP32 with even/odd folding (batched face passes; others use 1)
int g_gemm_variant = 1;

int launch_gemm(const GemmDesc* d_descs, int ndesc, int total_tiles, bool b_k_contig, bool c_m_contig, int ep,
                hipStream_t s, int variant) {
  if (total_tiles <= 0) return 0;
  dim3 grid(total_tiles), block(256);
  if (variant < 0) variant = g_gemm_variant;
  if (variant == 3 || variant == 5) variant = 1;  // folding applies to the batched form's face passes; dense passes use 1
  if (variant == 2) {
    if (b_k_contig && ep == 0) hipLaunchKernelGGL((k_gemm_bf16x3<true, 0>), grid, block, 0, s, d_descs, ndesc);
    else if (!b_k_contig && ep == 0) hipLaunchKernelGGL((k_gemm_bf16x3<false, 0>), grid, block, 0, s, d_descs, ndesc);
    else if (!b_k_contig && ep == 1) hipLaunchKernelGGL((k_gemm_bf16x3<false, 1>), grid, block, 0, s, d_descs, ndesc);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  if (variant == 1) {
    if (b_k_contig && ep == 0) hipLaunchKernelGGL((k_gemm_mfma<true, 0>), grid, block, 0, s, d_descs, ndesc);
    else if (!b_k_contig && ep == 0) hipLaunchKernelGGL((k_gemm_mfma<false, 0>), grid, block, 0, s, d_descs, ndesc);
    else if (!b_k_contig && ep == 1) hipLaunchKernelGGL((k_gemm_mfma<false, 1>), grid, block, 0, s, d_descs, ndesc);
    else return -1;
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  if (b_k_contig && c_m_contig && ep == 0)
    hipLaunchKernelGGL((k_gemm<true, true, 0>), grid, block, 0, s, d_descs, ndesc);
  else if (!b_k_contig && !c_m_contig && ep == 0)
    hipLaunchKernelGGL((k_gemm<false, false, 0>), grid, block, 0, s, d_descs, ndesc);
  else if (!b_k_contig && !c_m_contig && ep == 1)
    hipLaunchKernelGGL((k_gemm<false, false, 1>), grid, block, 0, s, d_descs, ndesc);
  else if (b_k_contig && !c_m_contig && ep == 0)
    hipLaunchKernelGGL((k_gemm<true, false, 0>), grid, block, 0, s, d_descs, ndesc);
  else if (!b_k_contig && c_m_contig && ep == 0)
    hipLaunchKernelGGL((k_gemm<false, true, 0>), grid, block, 0, s, d_descs, ndesc);
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---------------------------------------------------------------------------
// PML slab update, one thread per cell (ARD/pml_partition.cpp:113-200).
// Out-of-slab neighbours read 0 (the reference's sentinel element, :6-15).
// ---------------------------------------------------------------------------
// r180 = 1 / (180 h^2), r12 = 1 / (12 h), rdt = 1 / dt, formed on the host in FP32
// One block of 256 cells (block t of slab P).
__device__ __forceinline__ void pml_cells(const PmlDesc& P, int t, int par, float r180, float r12, float dt,
                                          float rdt, float c0) {
  const int W = P.W, H = P.H, D = P.D;
  const unsigned cell = unsigned(t) * 256u + threadIdx.x;  // slabs hold < 2^29 cells (checked at setup)
  if (cell >= unsigned(W * H * D)) return;
  const unsigned row = cell / unsigned(W);
  const int i = int(cell - row * unsigned(W));
  const int j = int(row % unsigned(H));
  const int k = int(row / unsigned(H));
  const gfloat* __restrict__ p = ARD_CG(P.p[par]);
  gfloat* __restrict__ pn = ARD_G(P.p[par ^ 1]);  // holds p_old on entry
  const gfloat* __restrict__ fx = ARD_CG(P.phi[par][0]);
  const gfloat* __restrict__ fy = ARD_CG(P.phi[par][1]);
  const gfloat* __restrict__ fz = ARD_CG(P.phi[par][2]);
  // Neighbour reads through buffer resources over the slab: a neighbour
  // outside the slab gets an offset past the end, which reads 0 (the zero
  // sentinel of ARD/pml_partition.cpp:6-15). One offset per axis and distance
  // serves every array read there.
  const unsigned nbytes = unsigned(W * H * D) * 4u;  // slabs hold < 2^29 cells
  const brsrc rp = make_rsrc(P.p[par], nbytes);
  const brsrc rfx = make_rsrc(P.phi[par][0], nbytes);
  const brsrc rfy = make_rsrc(P.phi[par][1], nbytes);
  const brsrc rfz = make_rsrc(P.phi[par][2], nbytes);
  const unsigned c4b = cell * 4u;
  constexpr unsigned kOut = 0x80000000u;
  unsigned ox[7], oy[7], oz[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    const int d = m - 3;
    ox[m] = unsigned(i + d) < unsigned(W) ? c4b + unsigned(d * 4) : kOut;
    oy[m] = unsigned(j + d) < unsigned(H) ? c4b + unsigned(d * 4 * W) : kOut;
    oz[m] = unsigned(k + d) < unsigned(D) ? c4b + unsigned(d * 4 * W * H) : kOut;
  }
  const float c6[7] = {2.f, -27.f, 270.f, -490.f, 270.f, -27.f, 2.f};
  const float c4[5] = {1.f, -8.f, 0.f, 8.f, -1.f};
  float px[7], py[7], pz[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    px[m] = bload(rp, ox[m], 0u);
    py[m] = bload(rp, oy[m], 0u);
    pz[m] = bload(rp, oz[m], 0u);
  }
  float dxx = 0.f, dyy = 0.f, dzz = 0.f;
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    dxx += c6[m] * px[m];
    dyy += c6[m] * py[m];
    dzz += c6[m] * pz[m];
  }
  // divisions by the step's constants as products with their reciprocals
  // (within 1.5 ulp of the quotients)
  dxx *= r180;
  dyy *= r180;
  dzz *= r180;
  const float z = ARD_CG(P.zprof)[P.axis == 0 ? i : (P.axis == 1 ? j : k)];
  const float zx = P.axis == 0 ? z : 0.f, zy = P.axis == 1 ? z : 0.f, zz = P.axis == 2 ? z : 0.f;
  const float pc = p[cell];
  const float po = pn[cell];
  const float term1 = 2.f * pc;
  const float term2 = -po;
  const float term3 = c0 * c0 * (dxx + dyy + dzz);
  const float term4 = -(zx + zy + zz) * (pc - po) * rdt;
  const float term5 = -(zx * zy + zy * zz + zx * zz) * pc;
  float dpx = 0.f, dpy = 0.f, dpz = 0.f, dux = 0.f, duy = 0.f, duz = 0.f;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    if (m == 2) continue;  // zero coefficient
    dpx += c4[m] * bload(rfx, ox[m + 1], 0u);
    dpy += c4[m] * bload(rfy, oy[m + 1], 0u);
    dpz += c4[m] * bload(rfz, oz[m + 1], 0u);
    dux += c4[m] * px[m + 1];
    duy += c4[m] * py[m + 1];
    duz += c4[m] * pz[m + 1];
  }
  dpx *= r12; dpy *= r12; dpz *= r12;
  dux *= r12; duy *= r12; duz *= r12;
  const float term6 = dpx + dpy + dpz;
  gfloat* __restrict__ f = ARD_G(P.f);
  const float fc = f[cell];
  pn[cell] = term1 + term2 + dt * dt * (term3 + term4 + term5 + term6 + fc);
  const float phx = fx[cell], phy = fy[cell], phz = fz[cell];
  ARD_G(P.phi[par ^ 1][0])[cell] = phx - dt * zx * phx + dt * (zy + zz - zx) * dux;
  ARD_G(P.phi[par ^ 1][1])[cell] = phy - dt * zy * phy + dt * (zx + zz - zy) * duy;
  ARD_G(P.phi[par ^ 1][2])[cell] = phz - dt * zz * phz + dt * (zx + zy - zz) * duz;
  if (fc != 0.f) f[cell] = 0.f;  // force cleared after the update (:199)
}

__global__ __launch_bounds__(256) void k_pml(const PmlDesc* __restrict__ descs, int ndesc, int par, float r180,
                                             float r12, float dt, float rdt, float c0) {
  // contiguous block ranges per XCD: the stencil's y and z neighbours, read
  // by nearby blocks, stay in one L2
  const long long b = xcd_block(int(blockIdx.x), int(gridDim.x));
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  pml_cells(descs[lo], int(b - descs[lo].block0), par, r180, r12, dt, rdt, c0);
}

// ---------------------------------------------------------------------------
// LDS-tiled PML slab update, same arithmetic as k_pml. A block of 256 threads
// owns a tile that spans the slab's whole thickness L along its normal axis
// and 256 cells across it, 256 L cells, L per thread in memory order. Tile
// shape SH across the normal (first x second of the other axes in x, y, z
// order): SH 0 32 x 8 (normal z or y) or 16 x 16 (normal x); SH 1 the other
// way round (the autotuner's candidates: less halo against longer rows). p is staged with a 3-cell halo across
// the normal, each phi_a with a 2-cell halo along a when a lies across the
// normal, all from buffer loads that read 0 outside the slab; along the
// normal the tile covers the slab, so neighbours there outside [0, L) are the
// zero sentinel and are skipped. Each p and phi value is loaded from global
// memory about twice per tile instead of 21 and 5 times per cell.
// ---------------------------------------------------------------------------
template <int SH, int AX>
struct PmlShape {  // tile extents across the normal: first and second other axis
  static constexpr bool square = (SH == 0) == (AX == 0);
  static constexpr int T0 = square ? 16 : 32, T1 = square ? 16 : 8;
};

template <int L, int AX, int SH>
struct PmlTile {
  using S = PmlShape<SH, AX>;
  static constexpr int TX = AX == 0 ? L : S::T0;
  static constexpr int TY = AX == 1 ? L : (AX == 0 ? S::T0 : S::T1);
  static constexpr int TZ = AX == 2 ? L : S::T1;
  static_assert(S::T0 * S::T1 == 256, "one cell per thread in each normal layer");
  // p halo per axis (none along the normal) and its staged box
  static constexpr int HX = AX == 0 ? 0 : 3, HY = AX == 1 ? 0 : 3, HZ = AX == 2 ? 0 : 3;
  static constexpr int PX = TX + 2 * HX, PY = TY + 2 * HY, PZ = TZ + 2 * HZ;
  // phi_a halo along a (none along the normal)
  static constexpr int GX = AX == 0 ? 0 : 2, GY = AX == 1 ? 0 : 2, GZ = AX == 2 ? 0 : 2;
  static constexpr int NP = PX * PY * PZ;
  static constexpr int N0 = (TX + 2 * GX) * TY * TZ, N1 = TX * (TY + 2 * GY) * TZ, N2 = TX * TY * (TZ + 2 * GZ);
  static constexpr int total = NP + N0 + N1 + N2;
};

template <int L, int SH>
constexpr int pml_tile_lds() {
  constexpr int a = PmlTile<L, 0, SH>::total, b = PmlTile<L, 1, SH>::total, c = PmlTile<L, 2, SH>::total;
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

// Staging of the box [x0, x0 + BX) x [y0, y0 + BY) x [z0, z0 + BZ) of one
// slab array (x fastest), cells outside the slab reading 0: every load of
// every box is issued first (pml_fetch), then written to LDS (pml_put).
template <int BX, int BY, int BZ>
struct PmlBox {
  static constexpr int N = BX * BY * BZ, Q = (N + 255) / 256;  // elements, loads per thread
};

template <int BX, int BY, int BZ, int Q>
__device__ __forceinline__ void pml_fetch(float (&v)[Q], brsrc r, int x0, int y0, int z0, int W, int H, int D) {
  constexpr int N = PmlBox<BX, BY, BZ>::N;
  static_assert(Q == PmlBox<BX, BY, BZ>::Q, "one register per 256 box elements");
  constexpr unsigned kOut = 0x80000000u;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = int(threadIdx.x) + 256 * q;
    const int x = e % BX, y = (e / BX) % BY, z = e / (BX * BY);
    const unsigned gx = unsigned(x0 + x), gy = unsigned(y0 + y), gz = unsigned(z0 + z);
    const bool in = (e < N) & (gx < unsigned(W)) & (gy < unsigned(H)) & (gz < unsigned(D));
    const unsigned off = ((gz * unsigned(H) + gy) * unsigned(W) + gx) * 4u;
    v[q] = bload(r, in ? off : kOut, 0u);
  }
}

template <int BX, int BY, int BZ, int Q>
__device__ __forceinline__ void pml_put(float* __restrict__ s, const float (&v)[Q]) {
  static_assert(Q == PmlBox<BX, BY, BZ>::Q, "one register per 256 box elements");
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int e = int(threadIdx.x) + 256 * q;
    if (e < PmlBox<BX, BY, BZ>::N) s[e] = v[q];
  }
}

// s[c + d * stride] for a neighbour at d cells along the normal from a cell
// at normal position ln, 0 (the zero sentinel) outside [0, L); the address is
// clamped into the tile so every lane reads inside the staged box.
template <int L>
__device__ __forceinline__ float pml_normal(const float* __restrict__ s, int c, int ln, int d, int stride) {
  const int nd = min(max(ln + d, 0), L - 1);
  const float v = s[c + (nd - ln) * stride];
  return unsigned(ln + d) < unsigned(L) ? v : 0.f;
}

template <int L, int AX, int SH>
__device__ __forceinline__ void pml_tile(const PmlDesc& P, int t, float* __restrict__ sm, int par, float r180,
                                         float r12, float dt, float rdt, float c0) {
  using T = PmlTile<L, AX, SH>;
  using BP = PmlBox<T::PX, T::PY, T::PZ>;
  using B0 = PmlBox<T::TX + 2 * T::GX, T::TY, T::TZ>;
  using B1 = PmlBox<T::TX, T::TY + 2 * T::GY, T::TZ>;
  using B2 = PmlBox<T::TX, T::TY, T::TZ + 2 * T::GZ>;
  const int W = P.W, H = P.H, D = P.D;
  const int ta = t % P.nb, tb = t / P.nb;
  const int x0 = AX == 0 ? 0 : ta * T::TX;
  const int y0 = AX == 0 ? ta * T::TY : (AX == 1 ? 0 : tb * T::TY);
  const int z0 = AX == 2 ? 0 : tb * T::TZ;
  const unsigned nbytes = unsigned(W * H * D) * 4u;  // slabs hold < 2^29 cells (checked at setup)
  constexpr unsigned kOut = 0x80000000u;
  // the thread's L cells in memory order: tile coordinates, byte offsets
  // (kOut beyond the slab, where loads read 0 and stores are dropped), and
  // their p_old, force and zeta
  int lx[L], ly[L], lz[L];
  unsigned coff[L];
  float po_r[L], f_r[L], z_r[L];
  const brsrc rpn = make_rsrc(P.p[par ^ 1], nbytes), rf = make_rsrc(P.f, nbytes);
  {
    const brsrc rz = make_rsrc(P.zprof, unsigned(L) * 4u);
#pragma unroll
    for (int r = 0; r < L; ++r) {
      const int e = int(threadIdx.x) + 256 * r;
      lx[r] = e % T::TX;
      ly[r] = (e / T::TX) % T::TY;
      lz[r] = AX == 2 ? r : e / (T::TX * T::TY);  // a z-normal tile has 256 cells per plane
      const int i = x0 + lx[r], j = y0 + ly[r], k = z0 + lz[r];
      coff[r] = ((i < W) & (j < H) & (k < D)) ? unsigned((k * H + j) * W + i) * 4u : kOut;
      po_r[r] = bload(rpn, coff[r], 0u);
      f_r[r] = bload(rf, coff[r], 0u);
      z_r[r] = bload(rz, unsigned(AX == 0 ? lx[r] : (AX == 1 ? ly[r] : lz[r])) * 4u, 0u);
    }
  }
  float* __restrict__ sp = sm;
  float* __restrict__ s0 = sp + T::NP;
  float* __restrict__ s1 = s0 + T::N0;
  float* __restrict__ s2 = s1 + T::N1;
  {
    float vp[BP::Q], v0[B0::Q], v1[B1::Q], v2[B2::Q];
    pml_fetch<T::PX, T::PY, T::PZ>(vp, make_rsrc(P.p[par], nbytes), x0 - T::HX, y0 - T::HY, z0 - T::HZ, W, H, D);
    pml_fetch<T::TX + 2 * T::GX, T::TY, T::TZ>(v0, make_rsrc(P.phi[par][0], nbytes), x0 - T::GX, y0, z0, W, H, D);
    pml_fetch<T::TX, T::TY + 2 * T::GY, T::TZ>(v1, make_rsrc(P.phi[par][1], nbytes), x0, y0 - T::GY, z0, W, H, D);
    pml_fetch<T::TX, T::TY, T::TZ + 2 * T::GZ>(v2, make_rsrc(P.phi[par][2], nbytes), x0, y0, z0 - T::GZ, W, H, D);
    pml_put<T::PX, T::PY, T::PZ>(sp, vp);
    pml_put<T::TX + 2 * T::GX, T::TY, T::TZ>(s0, v0);
    pml_put<T::TX, T::TY + 2 * T::GY, T::TZ>(s1, v1);
    pml_put<T::TX, T::TY, T::TZ + 2 * T::GZ>(s2, v2);
  }
  __syncthreads();
  const brsrc rq0 = make_rsrc(P.phi[par ^ 1][0], nbytes), rq1 = make_rsrc(P.phi[par ^ 1][1], nbytes),
              rq2 = make_rsrc(P.phi[par ^ 1][2], nbytes);
  const float c6[7] = {2.f, -27.f, 270.f, -490.f, 270.f, -27.f, 2.f};
  const float c4[5] = {1.f, -8.f, 0.f, 8.f, -1.f};
  constexpr int SX1 = T::TX + 2 * T::GX;  // row length of the phi_x box
  constexpr int SY1 = T::TY + 2 * T::GY;  // column height of the phi_y box
  constexpr int SPZ = T::PX * T::PY;      // plane of the p box
#pragma unroll
  for (int r = 0; r < L; ++r) {
    // cells beyond the slab in a partial tile compute on zeros; their stores are dropped
    const int ln = AX == 0 ? lx[r] : (AX == 1 ? ly[r] : lz[r]);  // position along the normal
    const int cp = ((lz[r] + T::HZ) * T::PY + ly[r] + T::HY) * T::PX + lx[r] + T::HX;
    float px[7], py[7], pz[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      const int d = m - 3;
      px[m] = AX == 0 ? pml_normal<L>(sp, cp, ln, d, 1) : sp[cp + d];
      py[m] = AX == 1 ? pml_normal<L>(sp, cp, ln, d, T::PX) : sp[cp + d * T::PX];
      pz[m] = AX == 2 ? pml_normal<L>(sp, cp, ln, d, SPZ) : sp[cp + d * SPZ];
    }
    float dxx = 0.f, dyy = 0.f, dzz = 0.f;
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      dxx += c6[m] * px[m];
      dyy += c6[m] * py[m];
      dzz += c6[m] * pz[m];
    }
    dxx *= r180;
    dyy *= r180;
    dzz *= r180;
    const float z = z_r[r];
    const float zx = AX == 0 ? z : 0.f, zy = AX == 1 ? z : 0.f, zz = AX == 2 ? z : 0.f;
    const float pc = px[3];
    const float po = po_r[r];
    const float term1 = 2.f * pc;
    const float term2 = -po;
    const float term3 = c0 * c0 * (dxx + dyy + dzz);
    const float term4 = -(zx + zy + zz) * (pc - po) * rdt;
    const float term5 = -(zx * zy + zy * zz + zx * zz) * pc;
    const int c0i = (lz[r] * T::TY + ly[r]) * SX1 + lx[r] + T::GX;
    const int c1i = (lz[r] * SY1 + ly[r] + T::GY) * T::TX + lx[r];
    const int c2i = ((lz[r] + T::GZ) * T::TY + ly[r]) * T::TX + lx[r];
    float dpx = 0.f, dpy = 0.f, dpz = 0.f, dux = 0.f, duy = 0.f, duz = 0.f;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      if (m == 2) continue;  // zero coefficient
      const int d = m - 2;
      const float gx = AX == 0 ? pml_normal<L>(s0, c0i, ln, d, 1) : s0[c0i + d];
      const float gy = AX == 1 ? pml_normal<L>(s1, c1i, ln, d, T::TX) : s1[c1i + d * T::TX];
      const float gz = AX == 2 ? pml_normal<L>(s2, c2i, ln, d, T::TX * T::TY) : s2[c2i + d * T::TX * T::TY];
      dpx += c4[m] * gx;
      dpy += c4[m] * gy;
      dpz += c4[m] * gz;
      dux += c4[m] * px[m + 1];
      duy += c4[m] * py[m + 1];
      duz += c4[m] * pz[m + 1];
    }
    dpx *= r12; dpy *= r12; dpz *= r12;
    dux *= r12; duy *= r12; duz *= r12;
    const float term6 = dpx + dpy + dpz;
    const float fc = f_r[r];
    bstore(term1 + term2 + dt * dt * (term3 + term4 + term5 + term6 + fc), rpn, coff[r]);
    const float phx = s0[c0i], phy = s1[c1i], phz = s2[c2i];
    bstore(phx - dt * zx * phx + dt * (zy + zz - zx) * dux, rq0, coff[r]);
    bstore(phy - dt * zy * phy + dt * (zx + zz - zy) * duy, rq1, coff[r]);
    bstore(phz - dt * zz * phz + dt * (zx + zy - zz) * duz, rq2, coff[r]);
    bstore(0.f, rf, fc != 0.f ? coff[r] : kOut);  // force cleared after the update (ARD/pml_partition.cpp:199)
  }
}

template <int L, int SH>
__global__ __launch_bounds__(256) void k_pml_tiled(const PmlDesc* __restrict__ descs, int ndesc, int par,
                                                   float r180, float r12, float dt, float rdt, float c0) {
  __shared__ float sm[pml_tile_lds<L, SH>()];
  const long long b = xcd_block(int(blockIdx.x), int(gridDim.x));
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const PmlDesc& P = descs[lo];
  const int t = int(b - P.block0);
  if (P.nb == 0) pml_cells(P, t, par, r180, r12, dt, rdt, c0);  // a slab not L cells thick
  else if (P.axis == 0) pml_tile<L, 0, SH>(P, t, sm, par, r180, r12, dt, rdt, c0);
  else if (P.axis == 1) pml_tile<L, 1, SH>(P, t, sm, par, r180, r12, dt, rdt, c0);
  else pml_tile<L, 2, SH>(P, t, sm, par, r180, r12, dt, rdt, c0);
}

long long pml_slab_blocks(PmlDesc& d, int L, int shape) {
  const long long n = (long long)d.W * d.H * d.D;
  const int ext[3] = {d.W, d.H, d.D};
  if (L < 3 || L > 8 || ext[d.axis] != L) {  // one thread per cell
    d.nb = 0;
    return (n + 255) / 256;
  }
  // tile extents across the normal: first and second axis other than it
  const int a0 = d.axis == 0 ? 1 : 0, a1 = d.axis == 2 ? 1 : 2;
  const bool square = (shape == 0) == (d.axis == 0);  // PmlShape
  const int t0 = square ? 16 : 32, t1 = square ? 16 : 8;
  d.nb = (ext[a0] + t0 - 1) / t0;
  return (long long)d.nb * ((ext[a1] + t1 - 1) / t1);
}

int launch_pml(const PmlDesc* d_descs, int ndesc, long long total_blocks, int par, int L, float dh, float dt,
               float c0, hipStream_t s, int shape) {
  if (total_blocks <= 0) return 0;
  const float r180 = 1.f / (180.f * dh * dh), r12 = 1.f / (12.f * dh), rdt = 1.f / dt;
  const dim3 grid(static_cast<unsigned>(total_blocks)), block(256);
#define ARD_PML_CASE(LL)                                                                                             \
  case LL:                                                                                                           \
    if (shape == 1) hipLaunchKernelGGL((k_pml_tiled<LL, 1>), grid, block, 0, s, d_descs, ndesc, par, r180, r12, dt, \
                                       rdt, c0);                                                                     \
    else hipLaunchKernelGGL((k_pml_tiled<LL, 0>), grid, block, 0, s, d_descs, ndesc, par, r180, r12, dt, rdt, c0);   \
    break;
  switch (L) {
    ARD_PML_CASE(3)
    ARD_PML_CASE(4)
    ARD_PML_CASE(5)
    ARD_PML_CASE(6)
    ARD_PML_CASE(7)
    ARD_PML_CASE(8)
    default: hipLaunchKernelGGL(k_pml, grid, block, 0, s, d_descs, ndesc, par, r180, r12, dt, rdt, c0); break;
  }
#undef ARD_PML_CASE
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---------------------------------------------------------------------------
// Interface forcing (ARD/boundary.cpp:21-173). One thread per face element
// computes the 6 outputs m = -3..2 from the 3 cells on each side.
// ---------------------------------------------------------------------------
// The three cells of one side of an interface element: normal coordinates
// c0, c0 + 1, c0 + 2 along axis ax (0 x, 1 y, 2 z) at tangential coordinates
// (a, b), the other two axes in x, y, z order. Face-shell storage of the
// batched form: a cell belongs to the x-shell [6][D][H] if it is within 3 of
// an x face, else to the y-shell [6][D][W], else to the z-shell [6][H][W];
// plane q < 3 is cell q from the low face, q >= 3 cell q - 6 from the high
// one. The three cells lie within 3 of the faces normal to ax, so the
// tangential coordinates alone pick their shell: one shell w, the first
// cell's index and a stride, formed with selects (no branch on the lane's
// cell) so that a thread's reads issue together. A dense field: in = false
// when (a, b) lies outside it (the index is then a clamped cell's and the
// caller reads 0, the zero sentinel).
struct SideRun {
  long long off, stride;
  int w;
  bool in;
};

__device__ __forceinline__ SideRun side_run(const PartAccess& A, int ax, int c0, int a, int b) {
  const int W = A.W, H = A.H, D = A.D;
  const int x = ax == 0 ? c0 : a, y = ax == 1 ? c0 : (ax == 0 ? a : b), z = ax == 2 ? c0 : b;
  SideRun s;
  if (A.kind == 2) {  // uniform across the block
    const bool sx = x < 3 || x >= W - 3;
    const bool sy = y < 3 || y >= H - 3;
    const int qx = x < 3 ? x : x - W + 6, qy = y < 3 ? y : y - H + 6, qz = z < 3 ? z : z - D + 6;
    // shells hold 6 D H, 6 D W, 6 H W < 2^30 cells: batched partitions hold fewer than
    // 2^30 cells and DCT partitions are at least 6 cells along every axis (both checked)
    const unsigned ox = unsigned(qx * D + z) * unsigned(H) + unsigned(y);
    const unsigned oy = unsigned(qy * D + z) * unsigned(W) + unsigned(x);
    const unsigned oz = unsigned(qz * H + y) * unsigned(W) + unsigned(x);
    // a step along the normal inside each shell (the z-shell only for ax = 2,
    // the y-shell only for ax = 1 or 2)
    const int tx = ax == 0 ? D * H : (ax == 1 ? 1 : H);
    const int ty = ax == 1 ? D * W : W;
    s.w = sx ? 0 : (sy ? 1 : 2);
    s.off = sx ? ox : (sy ? oy : oz);
    s.stride = sx ? tx : (sy ? ty : H * W);
    s.in = true;
    return s;
  }
  s.in = !(x < 0 || x >= W || y < 0 || y >= H || z < 0 || z >= D);
  const int xc = min(max(x, 0), W - 1), yc = min(max(y, 0), H - 1), zc = min(max(z, 0), D - 1);
  s.off = ((long long)zc * H + yc) * W + xc;
  s.stride = ax == 0 ? 1 : (ax == 1 ? W : (long long)W * H);
  s.w = 0;
  return s;
}

// Base of a side's cells: shell w of sh (batched form) or the dense field.
// The three shell bases are block-uniform: held in scalar registers, so the
// compiler selects between them instead of loading a lane's base.
__device__ __forceinline__ float* run_base(const PartAccess& A, float* const* sh, float* dense, int w) {
  if (A.kind != 2) return dense;
  float* b0 = sh[0];
  float* b1 = sh[1];
  float* b2 = sh[2];
  asm volatile("" : "+s"(b0), "+s"(b1), "+s"(b2));
  return w == 0 ? b0 : (w == 1 ? b1 : b2);
}

__global__ __launch_bounds__(256) void k_interface(const BndDesc* __restrict__ descs, int ndesc,
                                                   const PartAccess* __restrict__ parts, int par,
                                                   long long total) {
  // each boundary's elements start on a block boundary: one lookup per block
  const long long b0 = (long long)blockIdx.x * 256;
  int lo = 0, hi = ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].elem0 <= b0) lo = mid;
    else hi = mid - 1;
  }
  const BndDesc& B = descs[lo];
  const long long e = b0 + threadIdx.x;
  // elements of one boundary write distinct cells, so their order is free:
  // the first tangential coordinate runs fastest (contiguous in the face slabs)
  const unsigned le = unsigned(e - B.elem0);
  if (e >= total || le >= unsigned(B.t0n) * unsigned(B.t1n)) return;
  const unsigned lv = le / unsigned(B.t0n);
  const int u = B.t0s + int(le - lv * unsigned(B.t0n));
  const int v = B.t1s + int(lv);
  const PartAccess& S0 = parts[B.side[0]];
  const PartAccess& S1 = parts[B.side[1]];
  const int u0 = u - B.org[0][0], v0 = v - B.org[0][1];
  const int u1 = u - B.org[1][0], v1 = v - B.org[1][1];
  const int n0 = B.type == 0 ? S0.W : (B.type == 1 ? S0.H : S0.D);
  // the element's cells: side 0 at normal n0-3..n0-1, side 1 at 0..2
  const SideRun r0 = side_run(S0, B.type, n0 - 3, u0, v0);
  const SideRun r1 = side_run(S1, B.type, 0, u1, v1);
  const float* p0 = run_base(S0, S0.ps, S0.p[par], r0.w);
  const float* p1 = run_base(S1, S1.ps, S1.p[par], r1.w);
  float pv[6];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    pv[n] = ARD_CG(p0)[r0.off + n * r0.stride];
    pv[3 + n] = ARD_CG(p1)[r1.off + n * r1.stride];
  }
  float pl[3], pr[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    pl[n] = r0.in ? pv[n] : 0.f;
    pr[n] = r1.in ? pv[3 + n] : 0.f;
  }
  const float C[6][6] = {{0, 0, -2, 2, 0, 0},       {0, -2, 27, -27, 2, 0},  {-2, 27, -270, 270, -27, 2},
                         {2, -27, 270, -270, 27, -2}, {0, 2, -27, 27, -2, 0}, {0, 0, 2, -2, 0, 0}};
  const unsigned mask = B.mask ? ((const __attribute__((address_space(1))) uint8_t*)B.mask)[le] : 0x3fu;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    if (!(mask & (1u << r))) continue;
    const int m = r - 3;
    float s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int n = 0; n < 3; ++n) s1 += C[r][n] * pl[n];
#pragma unroll
    for (int n = 0; n < 3; ++n) s2 += C[r][n + 3] * pr[n];
    if (m < 0) {  // side 0, cell n0 + m = the run's cell r
      const float sip = B.self[0] ? (s1 + s2) : s2;
      ARD_G(run_base(S0, S0.fs, S0.f, r0.w))[r0.off + r * r0.stride] = sip * B.k[0];
    } else {  // side 1, cell m
      const float sip = B.self[1] ? (s1 + s2) : s1;
      ARD_G(run_base(S1, S1.fs, S1.f, r1.w))[r1.off + m * r1.stride] = sip * B.k[1];
    }
  }
}

int launch_interface(const BndDesc* d_descs, int ndesc, const PartAccess* d_parts, int par, long long total,
                     hipStream_t s) {
  if (total <= 0) return 0;
  const long long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_interface, dim3(unsigned(blocks)), dim3(256), 0, s, d_descs, ndesc, d_parts, par, total);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---------------------------------------------------------------------------
// Sources and recorders. The step index lives in device memory (d_step) so
// that a captured graph replays with the right source time.
// ---------------------------------------------------------------------------
__global__ void k_sources(const SourceDesc* __restrict__ s, int n, const int* __restrict__ d_step, float courant) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = *d_step;
  float v;
  if (s[i].kind == 0) {
    // GaussianSource::SampleValue (ARD/gaussian_source.cpp:16-20), FP32
    const float x = 3.14159265358979f * ((2.f * courant * float(t)) / 6.f - 2.f);
    v = 1e9f * expf(-(x * x));
  } else {
    v = t < s[i].n ? s[i].wave[t] : 0.f;
  }
  *s[i].fptr = v;
}

int launch_sources(const SourceDesc* d_src, int n, const int* d_step, float courant, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_sources, dim3((n + 63) / 64), dim3(64), 0, s, d_src, n, d_step, courant);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

__global__ void k_record(const RecDesc* __restrict__ r, int n, float* __restrict__ trace, int cap,
                         int* __restrict__ d_step) {
  const int t = *d_step;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    if (t < cap) trace[(long long)t * n + i] = r[i].p[r[i].cell];
  __syncthreads();
  if (threadIdx.x == 0) *d_step = t + 1;
}

int launch_record(const RecDesc* d_rec, int n, float* trace, int cap, int* d_step, hipStream_t s) {
  hipLaunchKernelGGL(k_record, dim3(1), dim3(256), 0, s, d_rec, n, trace, cap, d_step);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace ard
