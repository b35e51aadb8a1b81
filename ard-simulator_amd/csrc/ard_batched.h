// Descriptors and launchers of the batched face-shell form (ard_batched_kernels.hip).
#pragma once

#include "ard_device.h"

#define ARD_MAX_REC_PER_PART 4
#define ARD_MAX_SRC_PER_PART 4  // interior point sources per partition

namespace ard {

// sum / difference of the face pair planes (x, n-1-x), x = 0..2
// The unfold walks its planes in items of `vec` consecutive cells (4 when
// the plane length and every pointer allow 16-byte accesses, else 1); each
// descriptor's items start on a 256-item boundary of the grid (block-uniform
// descriptor lookup).
struct UnfoldDesc {
  const float* part;  // [nchunk][6][plane] partial reductions: q < 3 even modes, q >= 3 odd
  float* dst;         // [6][plane] in face-pair order
  long long plane;
  int nchunk;
  int vec;
};

// 1D factors of a point in mode space: fx[W], fy[H], fz[D]; for a source the
// product fx*fy*fz*val is its forcing spectrum, for a recorder the weights
// of its pressure sample.
struct SrcTab {
  const float* fx;
  const float* fy;
  const float* fz;
  const float* val;
};
struct RecTab {
  const float* fx;
  const float* fy;
  const float* fz;
};

struct BSourceDesc {
  int kind;  // 0 Gaussian, 1 waveform, 2 constant
  const float* wave;
  int n;
  float value;
  float* dst;  // face-slab cell or value slot
};

struct ModalDesc {
  int W, H, D;
  int nbx, nby, nbz, kz_chunk;
  float dt;
  float gscale;       // .999 * 4 * s_f * s_i * dt^2 / 4 (force spectrum to modes)
  const float* wx2;   // [W] u_x = (c0 pi (kx+1) / Lx)^2 dt^2 / 4, and likewise for y, z
  const float* wy2;
  const float* wz2;
  float* M[2];       // [D][H][Wp]
  int Wp;            // mode row pitch (W rounded up to 32)
  const float* Ex;  // [W][3]  2 c_W(kx, x)
  const float* Rx;  // [W][3]  e(kx) c_W(kx, x)
  const float* Ey;
  const float* Ry;
  const float* Ez;
  const float* Rz;
  const float* yt;  // [H][8] {Ey[ky][0..2], Ry[ky][0..2], u_y, 0}
  const float* zt;  // [D][8] likewise for z
  // face spectra of the six raw slabs of each face pair (slab q at x = q for
  // q < 3, at x = W - 6 + q for q >= 3; likewise y, z), folded by the pass:
  // q and 5 - q summed for even wavenumbers along the normal, subtracted for odd
  const float* Gx;  // [D][H][6]
  const float* Gy;  // [6][D][W]
  const float* Gz;  // [6][H][W]
  float* Tx;        // [nbx][6][D][H] partial reductions
  float* Ty;        // [nby][6][D][W]
  float* Tz;        // [nbz][6][H][W]
  const float* pty;  // [H][8] y factors of the point forces (4) and recorders (4); zero past the counts
  const float* ptz;  // [D][8] likewise along z
  int nsrc;
  const SrcTab* src;
  int nrec;
  const RecTab* rec;
  float* rec_part;  // [nbx*nby*nbz][ARD_MAX_REC_PER_PART]
};

struct BRecDesc {
  const float* part;  // this recorder's partial of the owner's first block
  int nblocks;
  int stride;         // floats between consecutive blocks' partials
};

// Interior point sources of a partition past the modal pass's
// ARD_MAX_SRC_PER_PART (k_xsrc), and recorders past its
// ARD_MAX_REC_PER_PART (k_xrec): rows of 16 (kz, ky) mode rows per block.
constexpr int kXRows = 16;
struct XsrcDesc {
  float* M[2];  // modes by step parity; the pass adjusts M[par ^ 1]
  int W, H, D, Wp;
  const float* wx2;
  const float* wy2;
  const float* wz2;
  float gscale;
  int n;
  const SrcTab* src;
};
struct XrecDesc {
  const float* M[2];  // modes by step parity; the pass reads the new ones, M[par ^ 1]
  int W, H, D, Wp;
  int n;
  const RecTab* rec;
  float* part;  // [blocks][n]
};

int launch_unfold(const UnfoldDesc* d, const long long* starts, int n, long long total, hipStream_t s);
int launch_bsources(const BSourceDesc* d, int n, const int* d_step, float courant, hipStream_t s);
// ring: planes of mode rows in flight per wave through LDS (0: one, in registers; 2 or 3 with 4 rows per wave)
int launch_modal(int rpw, bool pts, int ring, const ModalDesc* d, const long long* starts, int n, long long blocks,
                 int par, hipStream_t s);
// records the step's samples, advances *d_step and writes the next step's source samples
// the extra point passes (before and after the modal pass)
int launch_xsrc(const XsrcDesc* d, const long long* starts, int n, long long blocks, int par, hipStream_t s);
int launch_xrec(const XrecDesc* d, const long long* starts, int n, long long blocks, int par, hipStream_t s);
int launch_brecord(const BRecDesc* d, int n, float* trace, int cap, int* d_step, const BSourceDesc* src, int nsrc,
                   float courant, hipStream_t s);

}  // namespace ard
