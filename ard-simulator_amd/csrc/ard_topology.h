// Scene topology of the ARD solver: partitions, boundaries, sources, recorders.
//
// Host-side C++ that reproduces the topology rules of the reference's
// Simulation constructor (ARD/simulation.cpp:19-266), Boundary::FindBoundary
// (ARD/boundary.cpp:289-334), Partition::AddBoundary (ARD/partition.cpp:96-143)
// and Recorder::FindPartition (ARD/recorder.cpp:26-38). Shared by the C-ABI
// (ard_build_topology) and the C++ class mirror in host/.
#pragma once

#include <string>
#include <vector>

namespace ard {

enum PmlType { P_LEFT = 0, P_RIGHT, P_TOP, P_BOTTOM, P_FRONT, P_BACK };
enum BoundaryType { X_BOUNDARY = 0, Y_BOUNDARY, Z_BOUNDARY };

struct PartGeom {
  bool is_pml = false;
  int pml_type = -1;
  int parent = -1;
  int x0 = 0, y0 = 0, z0 = 0, w = 0, h = 0, d = 0;
  int x1() const { return x0 + w; }
  int y1() const { return y0 + h; }
  int z1() const { return z0 + d; }
  long long cells() const { return (long long)w * h * d; }
  // Free-border masks of z-layer 0 (the only layer the PML search reads,
  // ARD/simulation.cpp:69,107,145,183); 1 = free.
  std::vector<char> left_free, right_free, top_free, bottom_free;
};

struct BoundaryGeom {
  int type = 0;
  int a = 0, b = 0;
  int xs = 0, xe = 0, ys = 0, ye = 0, zs = 0, ze = 0;
  float absorption = 1.0f;
};

struct SourceDef {
  int x = 0, y = 0, z = 0;  // global cells
  int part = -1;            // owning DCT partition
  std::vector<float> wave;  // empty: Gaussian pulse
};

struct RecorderDef {
  int x = 0, y = 0, z = 0;  // global cells
  int part = -1;            // owning DCT partition
  int lx = 0, ly = 0, lz = 0;  // partition-local cell actually sampled
  bool preset = false;      // part and local cell given (a batch's copy of a resolved recorder)
};

struct Topology {
  std::vector<PartGeom> parts;
  std::vector<BoundaryGeom> bnds;
  std::vector<SourceDef> sources;
  std::vector<RecorderDef> recorders;
  int pml_layers = 5;
  float absorption = 1.0f;

  int add_dct(int x0, int y0, int z0, int w, int h, int d);
  int add_pml(int parent, int type, int x0, int y0, int z0, int w, int h, int d);
  int add_boundary(const BoundaryGeom& b);
  // Boundary::FindBoundary between parts a and b; false when they do not share a face.
  bool find_boundary(int a, int b, float absorption, BoundaryGeom* out) const;
  // Partition::AddBoundary: clears the free-border mask cells the boundary covers.
  void mark_boundary(int part, const BoundaryGeom& b);
  // The Simulation constructor's boundary search and PML generation.
  void build();
  // Assigns sources to DCT partitions (ARD/simulation.cpp:37-48) and recorders
  // to their owner (ARD/recorder.cpp:26-38). recorder_local selects the
  // sampled cell; returns an error message or "".
  std::string resolve(bool recorder_local);
};

}  // namespace ard
