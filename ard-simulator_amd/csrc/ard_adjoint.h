// Discrete adjoint of the plain time step: device descriptors and launchers
// (ard_adjoint_kernels.hip), driven by ard_adjoint (ard_adjoint.cpp).
#pragma once

#include "ard_device.h"

namespace ard {

// One PML slab in the reverse sweep. Inputs: the adjoints lP, lPhi, lQ of the
// state after the step (lP also collects the interface transpose) and the
// forward state p, q, phi before the step; outputs: the adjoints op, oq,
// ophi, of of the state before the step.
struct AdjPmlDesc {
  int W, H, D;
  int axis;             // damping axis (x = 0, y = 1, z = 2)
  int face;             // PML type: index of the damping scale it belongs to
  const float* zeta;    // damping profile along the axis, scaled
  const float* prof;    // the same profile for scale 1
  const float* lP;
  const float* lPhi[3];
  const float* lQ;
  const float* p;
  const float* q;
  const float* phi[3];
  float* op;
  float* oq;
  float* ophi[3];
  float* of;
  long long block0;
};

struct AdjSource {
  float* lF;       // adjoint of the owner's force volume
  long long cell;
  int index;       // row of the source in the gradient output
};

int launch_adj_record(const RecDesc* d_rec, int n, const float* d_g, hipStream_t s);
int launch_adj_interface(const BndDesc* d_descs, int ndesc, const PartAccess* d_parts, long long total,
                         hipStream_t s);
int launch_adj_pml(const AdjPmlDesc* d_descs, int ndesc, long long total_blocks, float dh, float dt, float c0,
                   double* d_dtheta, hipStream_t s);
int launch_adj_modal(float* t, float* lM, float* lMp, const float* a, const float* g, float b, long long n,
                     hipStream_t s);
int launch_adj_add(float* lF, const float* t, long long n, hipStream_t s);
int launch_adj_sources(const AdjSource* d_src, int n, float* d_du, int step, int nsteps, hipStream_t s);

}  // namespace ard
