// Internal state of an ARD context (the object behind the C-ABI handle).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ard.h"
#include "ard_device.h"
#include "ard_masks.h"
#include "ard_topology.h"

namespace ard {

struct BatchedState;
struct AdjState;      // ard_adjoint.cpp
struct OutputState;   // ard_output.hip
struct ViewState;     // ard_view.hip

// Dense fields of one DCT partition in the plain form (ARD/dct_partition.h,
// ARD/dct_volume.h): force F and pressure p in the value domain, modes M[2]
// (current / previous by step parity, stored multiplied by the inverse
// transform scale), modal coefficients a and g.
struct DctPlain {
  float* F = nullptr;
  float* p = nullptr;
  float* M[2] = {nullptr, nullptr};
  float* a = nullptr;
  float* g = nullptr;
  float* t[2] = {nullptr, nullptr};  // pass intermediates of its own (folded passes, all partitions in one launch)
  int gemm0 = 0;  // first of its 12 GEMM problems (2 parities x 6 passes)
  int tiles[6] = {0, 0, 0, 0, 0, 0};
  int fft_tiles[6] = {0, 0, 0, 0, 0, 0};  // the same passes as FFT-style transforms (variant 4)
};

// PML slab buffers (ARD/pml_partition.h:8-25); shared by both forms.
struct PmlBuffers {
  float* p[2] = {nullptr, nullptr};
  float* phi[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  float* f = nullptr;
  float* zprof = nullptr;
  int desc = -1;  // index into the PML descriptor table
};

// HIP-event timing of the step phases (ard_set_kernel_timing).
struct PhaseTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<int> phase_of;  // phase of the event pair (2i, 2i+1)
  double total_ms[ARD_NUM_PHASES] = {};
  long long count[ARD_NUM_PHASES] = {};
  int begin(hipStream_t s, int phase);  // records the start event; returns the pair index or -1
  void end(hipStream_t s, int pair);
  int collect();                        // synchronises and accumulates
  void release();
};

struct Ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  ard_params prm{};
  float pml_scale[6] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
  Topology topo;
  bool finalized = false;
  bool topology_built = false;
  int time_step = 0;
  float courant = 0.f;  // c0 * dt / h in FP32, as ARD/gaussian_source.cpp:18 forms it

  std::vector<void*> allocs;
  std::map<int, std::pair<float*, float*>> tables;  // axis length -> (REDFT10, REDFT01) matrices
  std::map<int, std::pair<float*, float*>> fold_tables;  // axis length -> even/odd folded (forward, inverse)
  std::map<int, float*> fft_plans;                       // axis length -> FFT plan table (variant 4)
  std::map<int, std::array<void*, 4>> fold_bf16;         // axis length -> folded tables' BF16 hi / lo parts

  // step counter and recorder trace, device-resident so captured graphs replay correctly
  int* d_step = nullptr;
  float* d_trace = nullptr;
  int trace_cap = 0;
  // recorder cubes (ard_set_recorder_cubes): [trace_cap][recorders][1000]
  bool cubes = false;
  float* d_cube = nullptr;
  OutputState* out = nullptr;
  ViewState* view = nullptr;  // visualisation plane (ard_render_view)

  // plain form
  std::vector<DctPlain> dct;       // by partition id
  std::vector<PmlBuffers> pml;     // by partition id
  float* tmp[2] = {nullptr, nullptr};
  std::vector<GemmDesc> h_gemm;
  GemmDesc* d_gemm = nullptr;
  int plain_variant = 1;       // transform kernel of the plain form, fixed at finalize (1, 0, 2, 3 or 4)
  GemmDesc* d_fft = nullptr;   // FFT twins of h_gemm (variant 4)
  // variant 3: pass j of every partition in one launch (even/odd folded),
  // per step parity: descriptors, tile table, blocks
  struct PassGroup {
    GemmDesc* d = nullptr;
    int2* tiles = nullptr;
    int blocks = 0;
  };
  PassGroup pass_group[2][6];
  // every PML slab in one launch, every boundary in one launch (last-writer
  // masks keep the boundary list order's result)
  PmlDesc* d_pml_grp = nullptr;
  long long pml_grp_blocks = 0;
  BndDesc* d_bnd_grp = nullptr;
  long long bnd_grp_total = 0;
  const float* plain_graph_trace = nullptr;  // trace buffer the plain step graphs were captured with
  size_t fft_lds = 0;          // dynamic LDS of the plain form's FFT passes
  std::vector<PmlDesc> h_pml;      // block0 = 0 each (one launch per slab)
  PmlDesc* d_pml = nullptr;
  std::vector<long long> pml_blocks;
  std::vector<BndDesc> h_bnd;      // elem0 = 0 each (one launch per boundary)
  BndDesc* d_bnd = nullptr;
  std::vector<long long> bnd_elems;
  std::vector<PartAccess> h_parts;
  PartAccess* d_parts = nullptr;
  SourceDesc* d_src = nullptr;
  int nsrc = 0;
  RecDesc* d_rec = nullptr;

  PhaseTimer timer;

  // batched form
  BatchedState* bs = nullptr;
  // adjoint of the plain form (ard_adjoint.cpp), set up on first use
  AdjState* adj = nullptr;
  // a batched context's plain-form twin, where its adjoint runs (ard_adjoint.cpp)
  Ctx* twin = nullptr;
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};

  ~Ctx();
  int dalloc(void** out, size_t bytes);
  int dzero(void* p, size_t bytes);
  int upload(void* dst, const void* src, size_t bytes);
  int cos_tables(int n, const float** fwd, const float** inv);
};

int hip_fail(hipError_t e, const char* what);
int set_error(int code, const std::string& msg);
void modal_tables(int W, int H, int D, double h, double dt, double c0, std::vector<float>& a, std::vector<float>& g);

// shared set-up (ard_plain.cpp)
int setup_common(Ctx& c);               // step counter, PML slabs, sources/recorders tables
int build_boundary_descs(Ctx& c, std::vector<BndDesc>& out, std::vector<long long>& elems);
std::vector<float> pml_profile(const PartGeom& p, int layers, double h, double c0, double scale);
void destroy_adjoint(Ctx& c);
int ensure_trace(Ctx& c, int steps_needed);

// plain form (ard_plain.cpp)
int finalize_plain(Ctx& c);
int step_plain(Ctx& c, int n);

// batched form (ard_batched.hip)
int finalize_batched(Ctx& c);
int step_batched(Ctx& c, int n);
int batched_pressure_field(Ctx& c, int part, const float** out);
int batched_set_force(Ctx& c, int part, int x, int y, int z, float f);
void destroy_batched(Ctx& c);
int batched_config(Ctx& c, int* out4);
int batched_set_tiling(Ctx& c, int rpw, int kzc);
int batched_set_option(Ctx& c, int option, int value);
int batched_option(Ctx& c, int option, int* value);
int batched_modes(Ctx& c, int part, const float* M[2], int* wp);
// The state arrays of a DCT partition in the batched form (modes, face-slab
// force and pressure) as (pointer, floats), for the non-finite check.
void batched_state_arrays(Ctx& c, int part, std::vector<std::pair<const float*, long long>>& out);
// ard_check_finite: the lowest partition id whose state holds a NaN or an
// infinity, or -1; synchronous.
int check_finite(Ctx& c, int* bad_part);
int fold_selftest(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y, int kind);
// The even/odd folded twin of a dense per-axis pass (k_gemm_mfma_fold) and
// the tile table of a launch of such passes (ard_batched.hip).
int fold_twin(Ctx& c, GemmDesc& g, bool inverse);
int fold_pass_tiles(Ctx& c, const std::vector<GemmDesc>& v, int2** d);
// FFT-style transforms (ard_fft.hip): the FFT fields of a dense pass
// descriptor (its plan cached in the context; ARD_ERR_UNSUPPORTED for a
// length fft_factor refuses), its tile count, and the entry self-test
int fft_setup(Ctx& c, GemmDesc& g, bool inverse);
int fft_tiles(const GemmDesc& g);
int fft_selftest(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y);

// recorder cubes (ard_output.hip)
int check_cubes(const Ctx& c);  // host-side range check
int setup_cubes(Ctx& c);        // after the form's finalize
int enqueue_cubes(Ctx& c, int par, hipStream_t st);
void destroy_output(Ctx& c);

// visualisation plane (ard_view.hip)
int render_view(Ctx& c, int view, int coord, unsigned int* pixels, int* dims);
void destroy_view(Ctx& c);

// field access (either form)
int pressure_field_ptr(Ctx& c, int part, const float** out);
int set_force_cell(Ctx& c, int part, int x, int y, int z, float f);

}  // namespace ard
