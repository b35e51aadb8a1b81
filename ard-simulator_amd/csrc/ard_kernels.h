// Launch helpers of the ARD kernels (ard_kernels.hip). Each returns 0 on
// success, negative on a launch failure.
#pragma once

#include <vector>

#include "ard_device.h"

namespace ard {

// GEMM implementation used by launch_gemm: 0 vector FMA, 1 matrix cores (FP32).
extern int g_gemm_variant;
int launch_gemm(const GemmDesc* d_descs, int ndesc, int total_tiles, bool b_k_contig, bool c_m_contig, int ep,
                hipStream_t s, int variant = -1);  // variant -1: g_gemm_variant
// Even/odd folded transform on the matrix cores (FP32): fold 1 forward, 2 inverse.
// d_tiles[b] = (descriptor, tile within it) of block b (fold_tile_table).
// ep 1: the modal epilogue of GemmDesc (forward, B contiguous along n only).
int launch_gemm_fold(const GemmDesc* d_descs, const int2* d_tiles, int total_tiles, bool b_k_contig, int fold,
                     hipStream_t s, int ep = 0);
// The same folded transforms in BF16 split precision (three BF16 products,
// FP32 accumulate) on v_mfma_f32_32x32x16_bf16; the descriptors' A / Alo are
// the folded tables' BF16 high and low parts.
int launch_gemm_bf16_fold(const GemmDesc* d_descs, const int2* d_tiles, int total_tiles, bool b_k_contig, int fold,
                          hipStream_t s);
// The same FP32 folded transforms with each wave's operands loaded straight
// into registers (no LDS, no barriers; ard_fold_direct.hip): the descriptors
// of launch_gemm_fold with M = ceil(L/2) pair rows and tiles_n = ceil(N/128);
// block tiles of 32 pair rows x 128 columns (fold_tile_table with tile_m 32).
int launch_fold_direct(const GemmDesc* d_descs, const int2* d_tiles, int total_tiles, bool b_k_contig, int fold,
                       hipStream_t s);
// Blocks of one slab's update for PML thickness L (pml_layers) and tile
// shape (0 or 1, see k_pml_tiled); sets d.nb. Slabs L cells thick along their
// normal, for L from 3 to 8, take the LDS-tiled update; others (thicker
// layers, or slabs of another thickness added through ard_add_pml_partition)
// one thread per cell, in the same launch. A launch takes the shape its
// descriptors' blocks were counted for.
long long pml_slab_blocks(PmlDesc& d, int L, int shape = 0);
int launch_pml(const PmlDesc* d_descs, int ndesc, long long total_blocks, int par, int L, float dh, float dt,
               float c0, hipStream_t s, int shape = 0);
int launch_interface(const BndDesc* d_descs, int ndesc, const PartAccess* d_parts, int par, long long total,
                     hipStream_t s);
int launch_sources(const SourceDesc* d_src, int n, const int* d_step, float courant, hipStream_t s);
int launch_record(const RecDesc* d_rec, int n, float* trace, int cap, int* d_step, hipStream_t s);

// FFT-style per-axis passes (ard_fft.hip, transform variant 4) on the same
// descriptors as launch_gemm: fft_factor splits L into radices (4, 2, 3, 5, 7,
// 11, 13; false for other prime factors or L > 4096), fft_lines gives the
// complex lines per tile, fft_plan_table the device table of a length.
bool fft_factor(int L, std::vector<int>& rad);
int fft_lines(int L);
std::vector<float> fft_plan_table(int L, const std::vector<int>& rad);
size_t fft_lds_bytes(const GemmDesc& g);
int launch_fft(const GemmDesc* d_descs, int ndesc, int total_tiles, bool b_k_contig, bool c_m_contig, int ep,
               size_t lds_bytes, hipStream_t s);

}  // namespace ard
