// Device-side descriptors shared by the ARD kernels and the host context.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ard_bnd.h"

namespace ard {

// Workgroups are dealt to the 8 XCDs in turn (block b to XCD b % 8). This
// renumbering gives each XCD one contiguous range of logical blocks, so tiles
// that share face spectra and cache lines share that XCD's L2.
__device__ __forceinline__ int xcd_block(int bid, int n) {
  const int x = bid & 7, i = bid >> 3, q = n >> 3, r = n & 7;
  return x * q + min(x, r) + i;
}

// Pointers read from descriptor tables are generic pointers to the compiler,
// which then emits flat memory instructions; these casts state that they
// point to global memory, so it can use global and scalar loads.
typedef __attribute__((address_space(1))) float gfloat;
#define ARD_G(p) ((gfloat*)(p))
#define ARD_CG(p) ((const gfloat*)(p))
// read-only tables: constant address space (uniform reads become scalar loads)
typedef __attribute__((address_space(4))) const float cfloat;
#define ARD_CC(p) ((cfloat*)(p))

// A uniform pointer read once from a descriptor and held in scalar registers
// from then on: the empty asm makes the value opaque, so the compiler cannot
// re-read it from the descriptor at each use (a dependent load pair whose
// wait would also wait for every load issued before it).
__device__ __forceinline__ const gfloat* pinned(const void* p) {
  unsigned long long v = reinterpret_cast<unsigned long long>(p);
  asm volatile("" : "+s"(v));
  return reinterpret_cast<const gfloat*>(v);
}

// Raw buffer resource over n bytes (gfx9 word 3: 32-bit float data, no
// swizzle): loads at byte offsets at or past n return 0.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t brsrc;
__device__ __forceinline__ brsrc make_rsrc(const void* p, unsigned nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(nbytes), 0x00020000);
}
__device__ __forceinline__ float bload(brsrc r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 0));
}
// stores at byte offsets at or past n are dropped
__device__ __forceinline__ void bstore(float v, brsrc r, unsigned voff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(voff), 0, 0);
}

// One problem of a grouped, strided GEMM  C = alpha * A * B  (A: M x K row-major).
// B(k, n) = B[k*bK + (n % N1)*b1 + (n / N1)*b2], C(m, n) likewise with cM, c1, c2.
// Every per-axis DCT pass is one such problem: A is the cosine matrix of the
// axis and B the volume (or face planes) viewed along that axis.
struct GemmDesc {
  const float* A;
  const float* B;
  float* C;
  int M, N, K, N1;
  long long bK, b1, b2;
  long long cM, c1, c2;
  float alpha;
  int tiles_n;  // ceil(N / 64)
  int tile0;    // first tile of this problem in the launch
  // Modal epilogue (used by the last forward pass of the plain form):
  // C holds the previous modes on entry and receives the next ones,
  //   C = a*cur - b*C + g*(alpha*AB)   (ARD/dct_partition.cpp:66-70).
  const float* ep_cur;
  const float* ep_a;
  const float* ep_g;
  float ep_b;
  // Even/odd folded transforms along an axis of length L (k_gemm_mfma_fold):
  // forward, rows [0, Mh) give the even outputs 2m from x[j] + x[L-1-j], rows
  // Mh.. the odd outputs from x[j] - x[L-1-j]; inverse, row j gives
  // E = A_E X_even and O = A_O X_odd and stores x[j] = E + O, x[L-1-j] = E - O.
  int L;
  int Mh;
  // Extents for buffer-resource loads (k_gemm_mfma_fold): A holds one
  // [Ap][Kp] matrix (forward) or two (inverse: E then O), rows padded to a
  // multiple of 64 (forward) or 32 (inverse) and columns to 16 with zeros; B
  // is b_elems floats from B. Inverse: M = 64 ceil(ceil(L/2) / 32), 64 tile
  // rows (32 of E, 32 of O) per 32 outputs.
  int Ap, Kp;
  unsigned b_elems;
  // BF16 split folded transforms (k_gemm_bf16x3_fold): A holds the tables'
  // BF16 high parts, Alo their low parts (same layout, 2-byte elements)
  const void* Alo;
  // FFT-style per-axis transform (k_dct_fft, transform variant 4): the plan of
  // the length K (fft_plan_table), complex lines (pairs of lines) per tile,
  // direction (0 REDFT10, 1 REDFT01) and the radices of the stages
  const void* plan;
  int tlc;
  int inv;
  int nrad;
  unsigned char rad[12];
};

// A PML slab (ARD/pml_partition.h). Buffers are double-buffered by step parity:
// p[par] is the current pressure, p[par^1] the previous one and receives the
// next; phi[par] current, phi[par^1] next.
struct PmlDesc {
  int W, H, D;
  int axis;  // axis along which zeta is nonzero (x = 0, y = 1, z = 2)
  float* p[2];
  float* phi[2][3];
  float* f;
  const float* zprof;  // zeta along that axis (length = W, H or D)
  long long block0;    // first block of this slab in a grouped launch
  int nb;              // tiled update: tiles along the first axis across the normal (0: one thread per cell)
};

// Pressure / force access to one partition as seen by the interface kernel.
// kind 0: dense DCT fields (plain form); 1: PML (p double-buffered);
// 2: DCT face shells (batched form).
struct PartAccess {
  int kind;
  int W, H, D;
  float* p[2];
  float* f;
  // batched shells: x-shell [6][D][H], y-shell [6][D][W], z-shell [6][H][W]
  float* ps[3];
  float* fs[3];
};

struct SourceDesc {
  int part;
  long long cell;   // flat cell index in the DCT partition
  int kind;         // 0 Gaussian, 1 waveform
  const float* wave;
  int n;
  float* fptr;      // where the forcing value is written (plain: F + cell)
};

struct RecDesc {
  const float* p;   // plain: pressure field of the owner
  long long cell;
};

}  // namespace ard
