// Even/odd folded per-axis transforms on the matrix cores (FP32) with the
// operands loaded straight into registers: no LDS staging and no barriers.
//
// The folded transforms are those of k_gemm_mfma_fold (ard_kernels.hip): with
// c_L(k, L-1-j) = (-1)^k c_L(k, j),
//   forward (REDFT10)  X[2m]   = sum_j Fe[m][j] (x[j] + x[L-1-j])
//                      X[2m+1] = sum_j Fo[m][j] (x[j] - x[L-1-j])
//   inverse (REDFT01)  E[j] = sum_k Ie[j][k] X[2k],  O[j] = sum_k Io[j][k] X[2k+1]
//                      x[j] = E + O,  x[L-1-j] = E - O
// for j, k, m below P = ceil(L/2). One wave owns 32 "pair rows" p (m forward,
// j inverse) by 32 columns and keeps both accumulators (even and odd rows,
// or E and O), so each output pair is formed in the lanes that hold it.
//
// Operand order. v_mfma_f32_16x16x4_f32 takes, per lane l, one element of
// the first operand (row l % 16, k-slot l / 16) and one of the second (k-slot
// l / 16, column l % 16) and adds the four-term product sums over the slots
// (on gfx950 it reaches 151 TFLOP/s, against 130 for the 32x32x2 form).
// Which k goes into which (instruction, slot) is free as long as both
// operands agree, so lane l takes k = kc + 4 (l / 16) + s at instruction s of
// a 16-column chunk kc: each lane reads 4 consecutive table columns (one
// 16-byte load per table row) and, when B is K-contiguous, 4 consecutive
// elements of its column, so the four lanes of a row or column read 64
// contiguous bytes. A wave covers 32 pair rows x 32 columns as 2 x 2 blocks
// of 16 x 16. B K-contiguous computes C^T = B^T A^T (the output is contiguous
// along p, which the lanes then index), otherwise C = A B (contiguous along n).
//
// Tables: the folded tables of the batched form ([Ap][Kp] forward with the
// odd rows from Mh; [2][Ap][Kp] inverse), zero beyond ceil(L/2) columns, so B
// elements of columns k >= ceil(L/2) meet zeros (they are read from the
// neighbouring column or past the array, and are finite). Loads use buffer
// resources: past the operand arrays they read 0.
#include <hip/hip_runtime.h>

#include "ard_kernels.h"

namespace ard {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v4f ld4(brsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(r, int(voff), int(soff), 0));
}

// One chunk's operands of a lane, per 16-row (or 16-column) block: the two
// table rows' 4 columns and the matching B elements (forward: x and its
// mirror; inverse: X even and odd).
struct Chunk {
  v4f a0[2], a1[2];
  v4f b0[2], b1[2];
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <bool B_K_CONTIG, int FOLD>
__global__ __launch_bounds__(256) void k_fold_direct(const GemmDesc* __restrict__ descs,
                                                     const int2* __restrict__ tiles) {
  const int2 tm = tiles[blockIdx.x];
  const GemmDesc& g = descs[tm.x];
  const int P = g.M, N = g.N, N1 = g.N1, L = g.L, Kp = g.Kp;
  // block tile: 32 pair rows x 128 columns, one 32-column slice per wave (the
  // four waves share the table rows)
  const int tiles_m = (P + 31) >> 5;
  // the wave index as a uniform value: tile bounds, loop trip counts and the
  // loads' scalar offsets stay in scalar registers
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p0 = (tm.y % tiles_m) * 32;
  const int n0 = (tm.y / tiles_m) * 128 + wave * 32;
  if (p0 >= P || n0 >= N) return;  // no barriers: a wave past the tile's edge just ends
  const int q = lane >> 4, li = lane & 15;

  // table rows of this lane in row block b: forward even row p and odd row
  // Mh + p; inverse E row p and O row Ap + p (p = p0 + 16 b + l % 16)
  const brsrc rA = make_rsrc(g.A, unsigned((FOLD == 2 ? 2 : 1) * g.Ap * Kp) * 4u);
  const unsigned ao0 = unsigned((p0 + li) * Kp + 4 * q) * 4u;
  const unsigned ao1 = unsigned(((FOLD == 2 ? g.Ap : g.Mh) + p0 + li) * Kp + 4 * q) * 4u;
  const unsigned ablk = unsigned(16 * Kp) * 4u;  // next row block
  // this lane's B columns n0 + 16 b + l % 16 (clamped; outputs past N are not stored)
  int bcol[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int n = min(n0 + 16 * b + li, N - 1);
    bcol[b] = int((n % N1) * g.b1 + (n / N1) * g.b2);
  }
  const brsrc rB = make_rsrc(g.B, (g.b_elems + 4u) * 4u);
  const unsigned bK4 = unsigned(g.bK) * 4u;

  auto load = [&](const int kc_, Chunk& c) {
    const int kc = __builtin_amdgcn_readfirstlane(kc_);  // uniform (scalar offsets below)
    const unsigned ka = unsigned(kc) * 4u;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      c.a0[b] = ld4(rA, ao0 + b * ablk, ka);
      c.a1[b] = ld4(rA, ao1 + b * ablk, ka);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (B_K_CONTIG) {
        if (FOLD == 1) {
          // x[k], k = kc + 4q + s, and x[L-1-k] = element 3 - s of the 4 from L-4-kc-4q
          c.b0[b] = ld4(rB, unsigned(bcol[b] + 4 * q) * 4u, ka);
          const v4f m = ld4(rB, unsigned(bcol[b] + L - 4 - 4 * q - kc) * 4u, 0u);  // negative: past the array, 0
          c.b1[b] = v4f{m.w, m.z, m.y, m.x};
        } else {
          // X[2k], X[2k+1] for k = kc + 4q + s: 8 consecutive elements
          const unsigned vo = unsigned(bcol[b] + 8 * q) * 4u;
          const v4f f0 = ld4(rB, vo, 2u * ka), f1 = ld4(rB, vo + 16u, 2u * ka);
          c.b0[b] = v4f{f0.x, f0.z, f1.x, f1.z};
          c.b1[b] = v4f{f0.y, f0.w, f1.y, f1.w};
        }
      } else {
        if (FOLD == 1) {
          // rows k = kc + 4q + s (uniform part (kc + s) bK in the scalar
          // offset) and L-1-k (negative rows wrap past the array: 0)
          const unsigned vo = unsigned(4 * q * int(g.bK) + bcol[b]) * 4u;
          const unsigned vm = unsigned((L - 1 - 4 * q - kc) * int(g.bK) + bcol[b]) * 4u;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            c.b0[b][t] = bload(rB, vo, unsigned(kc + t) * bK4);
            c.b1[b][t] = bload(rB, vm - unsigned(t) * bK4, 0u);
          }
        } else {
          // rows 2k and 2k+1
          const unsigned vo = unsigned(8 * q * int(g.bK) + bcol[b]) * 4u;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            c.b0[b][t] = bload(rB, vo, unsigned(2 * (kc + t)) * bK4);
            c.b1[b][t] = bload(rB, vo, unsigned(2 * (kc + t) + 1) * bK4);
          }
        }
      }
    }
  };

  // accumulators [even / E, odd / O][row block][column block]
  f32x4 acc[2][2][2];
#pragma unroll
  for (int e = 0; e < 2; ++e)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[e][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  auto mma = [&](const Chunk& c) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      float b0[2], b1[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        b0[j] = c.b0[j][s];
        b1[j] = c.b1[j][s];
        if (FOLD == 1) {  // x[j] +- x[L-1-j], rounded once
          const float sum = b0[j] + b1[j], dif = b0[j] - b1[j];
          b0[j] = sum;
          b1[j] = dif;
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (B_K_CONTIG) {
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[j], c.a0[i][s], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[j], c.a1[i][s], acc[1][i][j], 0, 0, 0);
          } else {
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(c.a0[i][s], b0[j], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(c.a1[i][s], b1[j], acc[1][i][j], 0, 0, 0);
          }
        }
    }
  };
  Chunk cur, nxt;
  load(0, cur);
  const int nchunks = Kp >> 4;  // Kp: a multiple of 16
  if (nchunks > 1) {
    // written as do-while: the compiler keeps this form of the loop uniform
    int ch = 1;
    do {
      load(16 * ch, nxt);
      // the next chunk's loads stay ahead of this chunk's products (the
      // scheduler would otherwise sink them to the loop's end)
      __builtin_amdgcn_sched_barrier(0);
      mma(cur);
      cur = nxt;
    } while (++ch < nchunks);
  }
  mma(cur);

  // Outputs. Accumulator r of lane l holds (row 4 (l/16) + r, column l % 16)
  // of its 16 x 16 block: rows are n and columns p when B is K-contiguous,
  // else rows p and columns n. Stores go through a buffer resource and an
  // output that does not exist gets an offset past its range (dropped), so
  // the epilogue has no branches (a divergent one here makes the compiler
  // treat the chunk loop above as divergent).
  const brsrc rC = make_rsrc(g.C, 0xFFFFFFF0u);
  constexpr unsigned kDrop = 0xFFFFFFF0u;
  const float alpha = g.alpha;
  auto put = [&](const bool ok, const long long off, const float v) { bstore(v, rC, ok ? unsigned(off) * 4u : kDrop); };
  auto out = [&](const int pp, const int nn, const float e, const float o) {
    const bool ok = pp < P && nn < N;
    const long long cn = (long long)(nn % N1) * g.c1 + (long long)(nn / N1) * g.c2;
    if (FOLD == 1) {
      put(ok, (long long)(2 * pp) * g.cM + cn, alpha * e);
      put(ok && 2 * pp + 1 < L, (long long)(2 * pp + 1) * g.cM + cn, alpha * o);
    } else {
      put(ok, (long long)pp * g.cM + cn, alpha * (e + o));
      put(ok && L - 1 - pp != pp, (long long)(L - 1 - pp) * g.cM + cn, alpha * (e - o));
    }
  };
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int pp = p0 + 16 * i + (B_K_CONTIG ? li : 4 * q + r);
        const int nn = n0 + 16 * j + (B_K_CONTIG ? 4 * q + r : li);
        out(pp, nn, acc[0][i][j][r], acc[1][i][j][r]);
      }
}

}  // namespace

int launch_fold_direct(const GemmDesc* d_descs, const int2* d_tiles, int total_tiles, bool b_k_contig, int fold,
                       hipStream_t s) {
  if (total_tiles <= 0) return 0;
  dim3 grid(total_tiles), block(256);
  if (fold == 1 && b_k_contig) hipLaunchKernelGGL((k_fold_direct<true, 1>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 1) hipLaunchKernelGGL((k_fold_direct<false, 1>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 2 && b_k_contig) hipLaunchKernelGGL((k_fold_direct<true, 2>), grid, block, 0, s, d_descs, d_tiles);
  else if (fold == 2) hipLaunchKernelGGL((k_fold_direct<false, 2>), grid, block, 0, s, d_descs, d_tiles);
  else return -1;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace ard
