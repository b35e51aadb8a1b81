// Topology rules of the reference's Simulation constructor, restated for the
// MI355X library. See ard_topology.h for the citations.
#include "ard_topology.h"

#include <algorithm>
#include <functional>

namespace ard {

int Topology::add_dct(int x0, int y0, int z0, int w, int h, int d) {
  PartGeom p;
  p.x0 = x0; p.y0 = y0; p.z0 = z0; p.w = w; p.h = h; p.d = d;
  // Partition ctor: all borders free (ARD/partition.cpp:22-51)
  p.left_free.assign(h, 1);
  p.right_free.assign(h, 1);
  p.top_free.assign(w, 1);
  p.bottom_free.assign(w, 1);
  parts.push_back(p);
  return int(parts.size()) - 1;
}

int Topology::add_pml(int parent, int type, int x0, int y0, int z0, int w, int h, int d) {
  int id = add_dct(x0, y0, z0, w, h, d);
  parts[id].is_pml = true;
  parts[id].pml_type = type;
  parts[id].parent = parent;
  return id;
}

int Topology::add_boundary(const BoundaryGeom& b) {
  bnds.push_back(b);
  return int(bnds.size()) - 1;
}

// ARD/boundary.cpp:289-334. Only X and Y faces are found (2.5D scenes): the
// z-range is taken from `a`.
bool Topology::find_boundary(int ia, int ib, float absorption, BoundaryGeom* out) const {
  const PartGeom& a = parts[ia];
  const PartGeom& b = parts[ib];
  const int x_ov = std::min(a.x1(), b.x1()) - std::max(a.x0, b.x0);
  const int y_ov = std::min(a.y1(), b.y1()) - std::max(a.y0, b.y0);
  BoundaryGeom g;
  g.a = ia; g.b = ib; g.absorption = absorption;
  g.zs = a.z0; g.ze = a.z1();
  if (x_ov == 0 && y_ov > 0) {
    g.type = X_BOUNDARY;
    g.xs = (a.x1() == b.x0) ? a.x1() - 3 : b.x1() - 3;
    g.xe = g.xs + 6;
    g.ys = a.y0 + std::max(0, b.y0 - a.y0);
    g.ye = g.ys + y_ov;
  } else if (y_ov == 0 && x_ov > 0) {
    g.type = Y_BOUNDARY;
    g.xs = a.x0 + std::max(0, b.x0 - a.x0);
    g.xe = g.xs + x_ov;
    g.ys = (a.y1() == b.y0) ? a.y1() - 3 : b.y1() - 3;
    g.ye = g.ys + 6;
  } else {
    return false;
  }
  *out = g;
  return true;
}

// ARD/partition.cpp:96-143.
void Topology::mark_boundary(int ip, const BoundaryGeom& b) {
  PartGeom& p = parts[ip];
  if (b.type == X_BOUNDARY) {
    std::vector<char>& m = (b.xs < p.x0 && b.xe > p.x0) ? p.left_free : p.right_free;
    for (int j = b.ys; j < b.ye; ++j) m[j - p.y0] = 0;
  } else if (b.type == Y_BOUNDARY) {
    std::vector<char>& m = (b.ys < p.y0 && b.ye > p.y0) ? p.top_free : p.bottom_free;
    for (int j = b.xs; j < b.xe; ++j) m[j - p.x0] = 0;
  }
}

namespace {
// The run search over one border mask (ARD/simulation.cpp:63-98, repeated for
// each side). Emits (start, length) for each run. A run that ends on the last
// element keeps that element even when it is not free, and a lone free
// element at the very end starts a run that is never emitted.
void find_runs(const std::vector<char>& freev, const std::function<void(int, int)>& emit) {
  const int n = int(freev.size());
  int start = 0, end = 0;
  bool started = false;
  for (int i = 0; i < n; ++i) {
    if (started) {
      end++;
      if (freev[i] && i != n - 1) continue;
      if (start != end) {
        if (i != n - 1) end--;
        emit(start, end - start + 1);
        started = false;
      }
    } else if (freev[i]) {
      start = end = i;
      started = true;
    }
  }
}
}  // namespace

void Topology::build() {
  const int ndct = int(parts.size());
  for (int i = 0; i < ndct; ++i)
    for (int j = i + 1; j < ndct; ++j) {
      BoundaryGeom b;
      if (find_boundary(i, j, 1.0f, &b)) {
        bnds.push_back(b);
        mark_boundary(i, b);
        mark_boundary(j, b);
      }
    }
  const int L = pml_layers;
  for (int c = 0; c < ndct; ++c) {
    const PartGeom P = parts[c];  // copy: add_pml grows the vector
    auto pml_with_boundary = [&](int type, int x0, int y0, int z0, int w, int h, int d) {
      const int id = add_pml(c, type, x0, y0, z0, w, h, d);
      BoundaryGeom b;
      find_boundary(id, c, absorption, &b);
      bnds.push_back(b);
    };
    find_runs(P.left_free, [&](int s, int n) { pml_with_boundary(P_LEFT, P.x0 - L, P.y0 + s, P.z0, L, n, P.d); });
    find_runs(P.right_free, [&](int s, int n) { pml_with_boundary(P_RIGHT, P.x1(), P.y0 + s, P.z0, L, n, P.d); });
    find_runs(P.top_free, [&](int s, int n) { pml_with_boundary(P_TOP, P.x0 + s, P.y0 - L, P.z0, n, L, P.d); });
    find_runs(P.bottom_free, [&](int s, int n) { pml_with_boundary(P_BOTTOM, P.x0 + s, P.y1(), P.z0, n, L, P.d); });
    // front and back slabs always exist, with explicitly placed Z boundaries (ARD/simulation.cpp:214-265)
    const int idf = add_pml(c, P_FRONT, P.x0, P.y0, P.z0 - L, P.w, P.h, L);
    BoundaryGeom bf;
    bf.type = Z_BOUNDARY; bf.absorption = absorption; bf.a = idf; bf.b = c;
    bf.xs = P.x0; bf.xe = P.x1(); bf.ys = P.y0; bf.ye = P.y1(); bf.zs = P.z0 - 3; bf.ze = P.z0 + 3;
    bnds.push_back(bf);
    const int idb = add_pml(c, P_BACK, P.x0, P.y0, P.z1(), P.w, P.h, L);
    BoundaryGeom bb = bf;
    bb.a = idb; bb.zs = P.z1() - 3; bb.ze = P.z1() + 3;
    bnds.push_back(bb);
  }
}

std::string Topology::resolve(bool recorder_local) {
  for (auto& s : sources) {
    s.part = -1;
    for (int i = 0; i < int(parts.size()); ++i) {
      const PartGeom& p = parts[i];
      if (p.is_pml) continue;
      if (s.x >= p.x0 && s.x < p.x1() && s.y >= p.y0 && s.y < p.y1() && s.z >= p.z0 && s.z < p.z1()) {
        s.part = i;  // half-open boxes: at most one owner
        break;
      }
    }
  }
  for (size_t r = 0; r < recorders.size(); ++r) {
    RecorderDef& rc = recorders[r];
    if (rc.preset) continue;
    rc.part = -1;
    for (int i = 0; i < int(parts.size()); ++i) {
      const PartGeom& p = parts[i];
      if (p.is_pml) continue;
      if (p.x0 < rc.x - 5 && p.x1() > rc.x + 4 && p.y0 < rc.y - 5 && p.y1() > rc.y + 4 && p.z0 < rc.z - 5 &&
          p.z1() > rc.z + 4) {
        rc.part = i;
        break;
      }
    }
    if (rc.part < 0)
      return "recorder " + std::to_string(r) + " is not 5 cells inside any DCT partition";
    const PartGeom& p = parts[rc.part];
    long long flat;
    if (recorder_local)
      flat = (long long)(rc.z - p.z0) * p.h * p.w + (long long)(rc.y - p.y0) * p.w + (rc.x - p.x0);
    else
      flat = (long long)rc.z * p.h * p.w + (long long)rc.y * p.w + rc.x;
    if (flat < 0 || flat >= p.cells())
      return "recorder " + std::to_string(r) + " addresses a cell outside its partition's array";
    rc.lx = int(flat % p.w);
    rc.ly = int((flat / p.w) % p.h);
    rc.lz = int(flat / ((long long)p.w * p.h));
  }
  return "";
}

}  // namespace ard
