// Host-side boundary set-up (ard_masks.cpp): interface descriptors, face-slab
// addressing and last-writer masks, on the topology alone.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "ard_bnd.h"
#include "ard_topology.h"

namespace ard {

// Interface descriptors of t's boundaries in list order and their face
// element counts; "" or an error message (its ARD_ERR_* code in *code).
std::string boundary_descs(const Topology& t, float h, float c0, std::vector<BndDesc>& out,
                           std::vector<long long>& elems, int* code);
// Which face slab stores cell (x,y,z) of a W x H x D partition (0 x, 1 y,
// 2 z; -1 for an interior cell), and its offset there in *off.
int shell_of(int W, int H, int D, int x, int y, int z, long long* off);
// Last-writer masks: for each face element of each boundary, bit r set when
// output r is not overwritten by a later boundary; mstart[b] = first element of b.
void final_writer_masks(const Topology& t, const std::vector<BndDesc>& descs, const std::vector<long long>& elems,
                        std::vector<uint8_t>& mask, std::vector<long long>& mstart);

}  // namespace ard
