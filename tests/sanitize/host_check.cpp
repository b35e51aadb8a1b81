// Host-side scene set-up of libard (topology, interface descriptors,
// last-writer masks: ard_topology.cpp, ard_masks.cpp) exercised under
// AddressSanitizer and UBSan by tests/test_sanitize_cpu.py, with the
// invariants that hold for every scene: one mask per face element, and on
// every DCT face-slab cell written by some boundary, exactly one final writer.
#include <cstdio>
#include <fstream>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../ard-simulator_amd/csrc/ard_masks.h"
#include "../../ard-simulator_amd/csrc/ard_topology.h"

using namespace ard;
typedef std::vector<std::vector<int>> Boxes;

static int check(const char* name, const Boxes& boxes, int layers, bool local) {
  Topology t;
  t.pml_layers = layers;
  for (const auto& b : boxes) t.add_dct(b[0], b[1], b[2], b[3], b[4], b[5]);
  t.build();
  // a recorder at the centre of the first box large enough
  for (int i = 0; i < int(boxes.size()); ++i) {
    const auto& b = boxes[i];
    if (b[3] > 11 && b[4] > 11 && b[5] > 11) {
      RecorderDef r;
      r.x = b[0] + b[3] / 2; r.y = b[1] + b[4] / 2; r.z = b[2] + b[5] / 2;
      t.recorders.push_back(r);
      SourceDef s;
      s.x = r.x - 1; s.y = r.y; s.z = r.z;
      t.sources.push_back(s);
      break;
    }
  }
  const std::string err = t.resolve(local);
  if (!err.empty()) { std::printf("FAIL %s resolve: %s\n", name, err.c_str()); return 1; }
  std::vector<BndDesc> d;
  std::vector<long long> e;
  int code = 0;
  const std::string e2 = boundary_descs(t, 0.05f, 343.0f, d, e, &code);
  if (!e2.empty()) { std::printf("FAIL %s descs: %s\n", name, e2.c_str()); return 1; }
  std::vector<uint8_t> mask;
  std::vector<long long> start;
  final_writer_masks(t, d, e, mask, start);
  long long total = 0;
  for (long long n : e) total += n;
  if ((long long)mask.size() != total) { std::printf("FAIL %s mask size\n", name); return 1; }
  // final writers per written DCT face-slab cell: count the set bits that land there
  std::map<std::pair<int, long long>, int> finals, writes;
  for (size_t b = 0; b < d.size(); ++b)
    for (long long k = 0; k < e[b]; ++k) {
      const int u = d[b].t0s + int(k % d[b].t0n), v = d[b].t1s + int(k / d[b].t0n);
      for (int r = 0; r < 6; ++r) {
        const int side = r < 3 ? 0 : 1, m = r - 3;
        const PartGeom& p = t.parts[d[b].side[side]];
        if (p.is_pml) continue;
        const int uu = u - d[b].org[side][0], vv = v - d[b].org[side][1];
        const int n0 = d[b].type == 0 ? p.w : (d[b].type == 1 ? p.h : p.d);
        const int nrm = side == 0 ? n0 + m : m;
        int x, y, z;
        if (d[b].type == 0) { x = nrm; y = uu; z = vv; }
        else if (d[b].type == 1) { x = uu; y = nrm; z = vv; }
        else { x = uu; y = vv; z = nrm; }
        long long off;
        const int w = shell_of(p.w, p.h, p.d, x, y, z, &off);
        if (w < 0) { std::printf("FAIL %s forced cell outside the face slabs\n", name); return 1; }
        const auto key = std::make_pair(d[b].side[side] * 3 + w, off);
        writes[key]++;
        if (mask[size_t(start[b] + k)] & (1u << r)) finals[key]++;
      }
    }
  for (const auto& kv : writes)
    if (finals[kv.first] != 1) { std::printf("FAIL %s: a cell with %d final writers\n", name, finals[kv.first]); return 1; }
  std::printf("ok %s: %zu partitions, %zu boundaries, %lld face elements, %zu written cells\n", name, t.parts.size(),
              d.size(), total, writes.size());
  return 0;
}

int main(int argc, char** argv) {
  int fails = 0;
  fails += check("two_rooms", {{0, 0, 0, 30, 30, 30}, {30, 0, 0, 30, 30, 30}}, 5, false);
  fails += check("scene3", {{0, 0, 0, 60, 60, 40}, {0, 60, 0, 40, 20, 40}, {60, 0, 0, 20, 40, 40},
                            {80, 0, 0, 20, 20, 40}, {0, 80, 0, 20, 20, 40}}, 5, true);
  fails += check("odd", {{0, 0, 0, 17, 13, 15}, {17, 0, 0, 11, 19, 15}, {0, 13, 0, 17, 9, 15}}, 3, true);
  fails += check("three_boxes_thick_pml", {{0, 0, 0, 16, 12, 16}, {16, 0, 0, 12, 20, 16}, {0, 12, 0, 16, 8, 16}}, 10, true);
  if (argc > 1) {  // the seeded hall, cell units
    std::ifstream f(argv[1]);
    Boxes b;
    std::vector<int> v(6);
    while (f >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5]) b.push_back(v);
    fails += check("hall", b, 5, true);
  }
  // seeded guillotine scenes with uneven sides
  std::mt19937 rng(12345);
  for (int s = 0; s < 20; ++s) {
    std::vector<std::vector<int>> rects = {{0, 0, 6 + int(rng() % 40), 6 + int(rng() % 40)}};
    for (int k = 0; k < 4; ++k) {
      auto r = rects[rng() % rects.size()];
      const bool ax = rng() & 1;
      const int L = ax ? r[2] : r[3];
      if (L < 12) continue;
      const int c = 6 + int(rng() % (L - 11));
      std::vector<int> a = r, b = r;
      if (ax) { a[2] = c; b[0] += c; b[2] -= c; } else { a[3] = c; b[1] += c; b[3] -= c; }
      for (auto& q : rects)
        if (q == r) { q = a; break; }
      rects.push_back(b);
    }
    Boxes b;
    const int d = 6 + int(rng() % 20);
    for (auto& r : rects) b.push_back({r[0], r[1], 0, r[2], r[3], d});
    fails += check(("guillotine_" + std::to_string(s)).c_str(), b, 3 + s % 6, true);
  }
  std::printf("%s\n", fails ? "FAILED" : "ALL OK");
  return fails ? 1 : 0;
}
