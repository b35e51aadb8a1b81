// A batch of independent rooms stepped together (ard_batch_*, include/ard.h):
// the rooms are laid side by side along x in one context, far enough apart
// that no partition or PML slab of one room touches another, so the merged
// topology is the union of the rooms' own topologies and one step graph (or
// one plain step) advances all of them. No counterpart in the reference,
// whose Simulation holds a single scene (ARD/simulation.cpp:14-294).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "ard_ctx_impl.h"

using namespace ard;

struct ard_batch {
  ard_ctx* merged = nullptr;
  std::vector<int> rec0, nrec;  // each room's recorders in the merged context
  std::vector<int> part0;       // each room's first DCT partition in the merged context
};

extern "C" {

int ard_batch_create(ard_ctx* const* members, int n, ard_batch** out) {
  if (!members || !out || n <= 0) return set_error(ARD_ERR_INVALID, "null argument or empty batch");
  std::vector<Ctx*> m(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    m[size_t(i)] = reinterpret_cast<Ctx*>(members[i]);
    Ctx* c = m[size_t(i)];
    if (!c) return set_error(ARD_ERR_INVALID, "null member");
    if (c->finalized || !c->topology_built)
      return set_error(ARD_ERR_STATE, "members must have their topology built and not be finalized");
    const Ctx* a = m[0];
    if (c->device != a->device || c->prm.h != a->prm.h || c->prm.dt != a->prm.dt || c->prm.c0 != a->prm.c0 ||
        c->prm.pml_layers != a->prm.pml_layers || c->prm.absorption != a->prm.absorption ||
        c->prm.form != a->prm.form || !std::equal(c->pml_scale, c->pml_scale + 6, a->pml_scale))
      return set_error(ARD_ERR_INVALID, "members differ in device, h, dt, c0, PML, absorption or form");
  }
  // each member's recorders as it would sample them on its own
  for (Ctx* c : m) {
    std::string err = c->topo.resolve(c->prm.recorder_mode == ARD_REC_LOCAL);
    if (!err.empty()) return set_error(ARD_ERR_RANGE, err);
  }
  for (Ctx* c : m)
    if (c->cubes != m[0]->cubes) return set_error(ARD_ERR_INVALID, "members differ in recorder cubes");
  ard_params p = m[0]->prm;
  p.recorder_mode = ARD_REC_LOCAL;  // recorders are copied resolved
  ard_ctx* h = nullptr;
  int r = ard_ctx_create(m[0]->device, &p, &h);
  if (r != ARD_OK) return r;
  Ctx* M = reinterpret_cast<Ctx*>(h);
  M->cubes = m[0]->cubes;
  std::copy(m[0]->pml_scale, m[0]->pml_scale + 6, M->pml_scale);
  auto* b = new ard_batch();
  b->merged = h;
  auto bail = [&](int code) {
    ard_ctx_destroy(h);
    delete b;
    return code;
  };
  const int gap = 2 * p.pml_layers + 8;  // cells between rooms, beyond both rooms' PML slabs
  std::vector<int> shift(static_cast<size_t>(n));
  std::vector<std::vector<int>> idmap(static_cast<size_t>(n));
  int next_x = 0;
  for (int i = 0; i < n; ++i) {
    const Topology& T = m[size_t(i)]->topo;
    int x0 = INT_MAX, x1 = INT_MIN;
    for (const PartGeom& g : T.parts)
      if (!g.is_pml) {
        x0 = std::min(x0, g.x0);
        x1 = std::max(x1, g.x1());
      }
    if (x0 == INT_MAX) return bail(set_error(ARD_ERR_INVALID, "a member has no DCT partition"));
    shift[size_t(i)] = next_x - x0;
    next_x += (x1 - x0) + gap;
    b->part0.push_back(int(M->topo.parts.size()));
    idmap[size_t(i)].assign(T.parts.size(), -1);
    for (size_t k = 0; k < T.parts.size(); ++k) {
      const PartGeom& g = T.parts[k];
      if (g.is_pml) continue;
      idmap[size_t(i)][k] = int(M->topo.parts.size());
      if ((r = ard_add_dct_partition(h, g.x0 + shift[size_t(i)], g.y0, g.z0, g.w, g.h, g.d)) < 0) return bail(r);
    }
  }
  if ((r = ard_build_topology(h)) != ARD_OK) return bail(r);
  for (int i = 0; i < n; ++i) {
    const Topology& T = m[size_t(i)]->topo;
    for (const SourceDef& s : T.sources)
      if ((r = ard_add_source(h, s.x + shift[size_t(i)], s.y, s.z, s.wave.empty() ? nullptr : s.wave.data(),
                              int(s.wave.size()))) < 0)
        return bail(r);
    b->rec0.push_back(int(M->topo.recorders.size()));
    b->nrec.push_back(int(T.recorders.size()));
    for (const RecorderDef& rc : T.recorders) {
      RecorderDef q = rc;
      q.x += shift[size_t(i)];
      q.part = idmap[size_t(i)][size_t(rc.part)];
      q.preset = true;
      M->topo.recorders.push_back(q);
    }
  }
  // host-only members give a host-only batch (topology inspection); others are finalized
  if (M->device >= 0 && (r = ard_finalize(h)) != ARD_OK) return bail(r);
  *out = b;
  return ARD_OK;
}

int ard_batch_context(ard_batch* b, ard_ctx** out) {
  if (!b || !out) return set_error(ARD_ERR_INVALID, "null argument");
  *out = b->merged;
  return ARD_OK;
}

int ard_batch_step(ard_batch* b, int nsteps) {
  if (!b) return set_error(ARD_ERR_INVALID, "null batch");
  return ard_step(b->merged, nsteps);
}

int ard_batch_fetch_recorders(ard_batch* b, int room, float* out, int t0, int nt) {
  if (!b || (!out && nt > 0)) return set_error(ARD_ERR_INVALID, "null argument");
  if (room < 0 || room >= int(b->rec0.size())) return set_error(ARD_ERR_RANGE, "no such room");
  const int all = ard_num_recorders(b->merged);
  std::vector<float> buf(size_t(std::max(nt, 0)) * size_t(std::max(all, 1)));
  int r = ard_fetch_recorders(b->merged, buf.data(), t0, nt);
  if (r != ARD_OK) return r;
  const int k0 = b->rec0[size_t(room)], k = b->nrec[size_t(room)];
  for (int t = 0; t < nt; ++t)
    for (int j = 0; j < k; ++j) out[size_t(t) * k + j] = buf[size_t(t) * all + k0 + j];
  return ARD_OK;
}

int ard_batch_room_partition(ard_batch* b, int room, int part) {
  if (!b) return set_error(ARD_ERR_INVALID, "null batch");
  if (room < 0 || room >= int(b->part0.size())) return set_error(ARD_ERR_RANGE, "no such room");
  return b->part0[size_t(room)] + part;
}

int ard_batch_destroy(ard_batch* b) {
  if (!b) return set_error(ARD_ERR_INVALID, "null batch");
  ard_ctx_destroy(b->merged);
  delete b;
  return ARD_OK;
}

}  // extern "C"
