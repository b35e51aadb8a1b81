// The interface descriptor shared by the interface kernels and the host-side
// boundary set-up (plain data, no device code, so host-only files can use it).
#pragma once

#include <cstdint>

namespace ard {

// One interface (ARD/boundary.cpp:21-173) with its sides resolved: side 0 is
// the left/top/front partition, side 1 the right/bottom/back one.
struct BndDesc {
  int type;        // 0 X, 1 Y, 2 Z
  int side[2];     // partition index of each side
  int t0s, t0n;    // first tangential coordinate range (global): X: y, Y: x, Z: x
  int t1s, t1n;    // second tangential range (global): X: z, Y: z, Z: y
  int org[2][2];   // per side: origin of the two tangential axes (global)
  float k[2];      // absorption * c0^2 / (180 h^2), per side (same value)
  int self[2];     // include_self_terms_ per side
  const uint8_t* mask;  // per element, bit m+3 set: write output m (last-writer rule); null = all
  long long elem0; // first element of this boundary in a grouped launch
};

}  // namespace ard
