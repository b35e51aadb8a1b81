// Host-side boundary set-up shared by both forms: the interface descriptors
// of the topology's boundaries (ARD/boundary.cpp:21-173 resolved per side),
// the face-slab address of a cell (host twin of shell_addr, ard_kernels.hip)
// and the last-writer masks (ARD/simulation.cpp:313-317 replayed serially).
// Plain C++ on the topology, no device calls: built by the host compiler, so
// that it can also be checked under AddressSanitizer and UBSan.
#include "ard_masks.h"

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ard.h"

namespace ard {

std::string boundary_descs(const Topology& t, float h, float c0, std::vector<BndDesc>& out,
                           std::vector<long long>& elems, int* code) {
  out.clear();
  elems.clear();
  for (const BoundaryGeom& b : t.bnds) {
    const PartGeom& A = t.parts[b.a];
    BndDesc d{};
    d.type = b.type;
    bool a_first;
    if (b.type == X_BOUNDARY) a_first = (b.xs <= A.x1() && b.xe >= A.x1());
    else if (b.type == Y_BOUNDARY) a_first = (b.ys <= A.y1() && b.ye >= A.y1());
    else a_first = (b.zs <= A.z1() && b.ze >= A.z1());
    d.side[0] = a_first ? b.a : b.b;
    d.side[1] = a_first ? b.b : b.a;
    if (b.type == X_BOUNDARY) {
      d.t0s = b.ys; d.t0n = b.ye - b.ys; d.t1s = b.zs; d.t1n = b.ze - b.zs;
    } else if (b.type == Y_BOUNDARY) {
      d.t0s = b.xs; d.t0n = b.xe - b.xs; d.t1s = b.zs; d.t1n = b.ze - b.zs;
    } else {
      d.t0s = b.xs; d.t0n = b.xe - b.xs; d.t1s = b.ys; d.t1n = b.ye - b.ys;
    }
    if (d.t0n <= 0 || d.t1n <= 0) return *code = ARD_ERR_INVALID, "empty boundary";
    for (int s = 0; s < 2; ++s) {
      const PartGeom& P = t.parts[d.side[s]];
      if (b.type == X_BOUNDARY) { d.org[s][0] = P.y0; d.org[s][1] = P.z0; }
      else if (b.type == Y_BOUNDARY) { d.org[s][0] = P.x0; d.org[s][1] = P.z0; }
      else { d.org[s][0] = P.x0; d.org[s][1] = P.y0; }
      // fi = sip * c0^2 / (180 h^2), scaled by the boundary's absorption (ARD/boundary.cpp:64-65)
      d.k[s] = b.absorption * (c0 * c0 / (180.0f * h * h));
      d.self[s] = P.is_pml ? 0 : 1;  // include_self_terms_ (ARD/partition.h:38, ARD/pml_partition.cpp:20)
      // tangential ranges must lie inside each side
      const int n0 = b.type == X_BOUNDARY ? P.h : P.w;
      const int n1 = b.type == Z_BOUNDARY ? P.h : P.d;
      if (d.t0s - d.org[s][0] < 0 || d.t0s + d.t0n - d.org[s][0] > n0 || d.t1s - d.org[s][1] < 0 ||
          d.t1s + d.t1n - d.org[s][1] > n1)
        return *code = ARD_ERR_RANGE, "boundary face extends past a partition";
      const int nn = b.type == X_BOUNDARY ? P.w : (b.type == Y_BOUNDARY ? P.h : P.d);
      if (nn < 3) return *code = ARD_ERR_RANGE, "partition thinner than the 3-cell interface stencil";
    }
    d.mask = nullptr;
    d.elem0 = 0;
    out.push_back(d);
    elems.push_back((long long)d.t0n * d.t1n);
  }
  return "";
}

// Which face slab stores cell (x,y,z) of a W x H x D partition, and where
// (host twin of shell_addr, ard_kernels.hip); -1 for an interior cell.
int shell_of(int W, int H, int D, int x, int y, int z, long long* off) {
  if (x < 3 || x >= W - 3) {
    const int q = x < 3 ? x : x - W + 6;
    *off = ((long long)q * D + z) * H + y;
    return 0;
  }
  if (y < 3 || y >= H - 3) {
    const int q = y < 3 ? y : y - H + 6;
    *off = ((long long)q * D + z) * W + x;
    return 1;
  }
  if (z < 3 || z >= D - 3) {
    const int q = z < 3 ? z : z - D + 6;
    *off = ((long long)q * H + y) * W + x;
    return 2;
  }
  return -1;
}

// Last-writer masks: replays the boundaries in list order over the written
// cells (ARD/simulation.cpp:313-317 run serially) and keeps, for each face
// element, the outputs that no later boundary overwrites (bit r = output r).
void final_writer_masks(const Topology& t, const std::vector<BndDesc>& descs, const std::vector<long long>& elems,
                        std::vector<uint8_t>& mask, std::vector<long long>& mstart) {
  const int np = int(t.parts.size());
  std::vector<std::vector<int>> owner[3];
  for (int w = 0; w < 3; ++w) owner[w].resize(np);
  std::vector<std::vector<int>> powner(np);
  for (int i = 0; i < np; ++i) {
    const PartGeom& p = t.parts[i];
    if (p.is_pml) powner[i].assign(size_t(p.cells()), -1);
    else {
      owner[0][i].assign(size_t(6) * p.d * p.h, -1);
      owner[1][i].assign(size_t(6) * p.d * p.w, -1);
      owner[2][i].assign(size_t(6) * p.h * p.w, -1);
    }
  }
  auto target = [&](const BndDesc& d, int side, int u, int v, int m, int** slot) {
    const PartGeom& p = t.parts[d.side[side]];
    const int uu = u - d.org[side][0], vv = v - d.org[side][1];
    const int n0 = d.type == 0 ? p.w : (d.type == 1 ? p.h : p.d);
    const int nrm = side == 0 ? n0 + m : m;
    int x, y, z;
    if (d.type == 0) { x = nrm; y = uu; z = vv; }
    else if (d.type == 1) { x = uu; y = nrm; z = vv; }
    else { x = uu; y = vv; z = nrm; }
    if (p.is_pml) {
      *slot = &powner[d.side[side]][(size_t(z) * p.h + y) * p.w + x];
    } else {
      long long off;
      const int w = shell_of(p.w, p.h, p.d, x, y, z, &off);
      *slot = w >= 0 ? &owner[w][d.side[side]][size_t(off)] : nullptr;
    }
  };
  for (int b = 0; b < int(descs.size()); ++b) {
    const BndDesc& d = descs[b];
    for (long long e = 0; e < elems[b]; ++e) {
      const int u = d.t0s + int(e % d.t0n), v = d.t1s + int(e / d.t0n);  // k_interface's element order
      for (int r = 0; r < 6; ++r) {
        int* s;
        target(d, r < 3 ? 0 : 1, u, v, r - 3, &s);
        if (s) *s = b;
      }
    }
  }
  mask.clear();
  mstart.assign(descs.size(), 0);
  for (int b = 0; b < int(descs.size()); ++b) {
    const BndDesc& d = descs[b];
    mstart[b] = (long long)mask.size();
    for (long long e = 0; e < elems[b]; ++e) {
      const int u = d.t0s + int(e % d.t0n), v = d.t1s + int(e / d.t0n);
      uint8_t m = 0;
      for (int r = 0; r < 6; ++r) {
        int* s;
        target(d, r < 3 ? 0 : 1, u, v, r - 3, &s);
        if (s && *s == b) m |= uint8_t(1u << r);
      }
      mask.push_back(m);
    }
  }
}

}  // namespace ard
