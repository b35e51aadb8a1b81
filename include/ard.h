/*
 * ard.h -- C-ABI of the MI355X ARD time-stepping library (libard.so).
 *
 * The reference (pjx3/ARD-simulator) exposes this path as C++ classes, not as
 * an FFI. Each entry point below names the reference interface it replaces
 * (paths relative to the reference root, ARD = ARD-simulator-190113/). A
 * context holds one scene: DCT partitions, PML partitions, boundaries,
 * sources and recorders, all device-resident. Plain pointers and sizes only.
 *
 * Errors: every function returns ARD_OK (0) or a negative ARD_ERR_* code;
 * ard_last_error() gives a thread-local message. The reference asserts or
 * crashes instead (ARD/dct_volume.cpp:23,25; ARD/recorder.cpp:50).
 *
 * Threading: a context is externally synchronised (one host thread at a
 * time). Host pointers are borrowed for the duration of the call.
 *
 * Coordinates are integer cells. Volumes are z-major with x fastest,
 * idx = z*H*W + y*W + x (ARD/dct_volume.cpp:70-88).
 */
#ifndef ARD_H_
#define ARD_H_

#ifdef __cplusplus
extern "C" {
#endif

#define ARD_OK 0
#define ARD_ERR_INVALID (-1)     /* bad argument */
#define ARD_ERR_STATE (-2)       /* call not allowed in the context's state */
#define ARD_ERR_HIP (-3)         /* HIP runtime failure (message has the HIP error) */
#define ARD_ERR_RANGE (-4)       /* coordinate or index out of range */
#define ARD_ERR_UNSUPPORTED (-5) /* valid input this form cannot run */

typedef struct ard_ctx ard_ctx;

/* Stepping forms. */
#define ARD_FORM_PLAIN 0   /* per-partition launches, dense fields: the reference's structure */
#define ARD_FORM_BATCHED 1 /* all partitions per phase in one launch, face-shell transforms, hipGraph */

/* Recorder addressing. REFERENCE: the recorder's global cell coordinates
 * index the owning partition's array (ARD/recorder.cpp:44-55 passes global
 * coordinates to get_pressure, ARD/dct_volume.cpp:70-73 does not subtract the
 * origin). LOCAL: the cell at those global coordinates. */
#define ARD_REC_REFERENCE 0
#define ARD_REC_LOCAL 1

/* PML face types (ARD/pml_partition.h:39-46) and boundary types (ARD/boundary.h:30-34). */
#define ARD_P_LEFT 0
#define ARD_P_RIGHT 1
#define ARD_P_TOP 2
#define ARD_P_BOTTOM 3
#define ARD_P_FRONT 4
#define ARD_P_BACK 5
#define ARD_X_BOUNDARY 0
#define ARD_Y_BOUNDARY 1
#define ARD_Z_BOUNDARY 2

typedef struct ard_params {
  float h;          /* Simulation::m_dh (ARD/simulation.h:43) */
  float dt;         /* Simulation::m_dt */
  float c0;         /* Simulation::m_c0 */
  int pml_layers;   /* Simulation::m_pml_layers */
  float absorption; /* Partition::m_absorption (ARD/partition.h:45) */
  int form;         /* ARD_FORM_* */
  int recorder_mode;/* ARD_REC_* */
  int autotune;     /* batched form: time candidate tilings and transform kernels at ard_finalize */
  int reserved[8];  /* zero */
} ard_params;

/* Replaces the statics set in ARD/main.cpp:29-45 plus the GPU set-up of
 * initVkGPU (ARD/utils_VkFFT.cpp:564-610). */
int ard_ctx_create(int device, const ard_params* params, ard_ctx** out);
int ard_ctx_destroy(ard_ctx* ctx);
const char* ard_last_error(void);

/* Per-face scale of the PML damping profile zeta (ARD/pml_partition.cpp:61-92
 * has scale 1). Before ard_finalize. */
int ard_set_pml_scale(ard_ctx* ctx, const float scale[6]);

/* DctPartition(xs,ys,zs,w,h,d) (ARD/dct_partition.cpp:7-37). Returns the id (>= 0). */
int ard_add_dct_partition(ard_ctx* ctx, int x0, int y0, int z0, int w, int h, int d);
/* PmlPartition(neighbor, type, xs,ys,zs,w,h,d) (ARD/pml_partition.cpp:17-93). Returns the id. */
int ard_add_pml_partition(ard_ctx* ctx, int neighbor, int type, int x0, int y0, int z0, int w, int h, int d);
/* Boundary(type, absorption, a, b, xs,xe,ys,ye,zs,ze) (ARD/boundary.cpp:7-15). Registration
 * order is write order: a cell forced by two boundaries keeps the later one's value. */
int ard_add_boundary(ard_ctx* ctx, int type, int a, int b, int xs, int xe, int ys, int ye, int zs, int ze,
                     float absorption);
/* Boundary::FindBoundary (ARD/boundary.cpp:289-334) on two boxes in cell units, box = {x0, y0,
 * z0, w, h, d}: 1 and bnd_info9 = {type, -1, -1, xs,xe,ys,ye,zs,ze} when they share an x or y
 * face (the 2.5D search: z from box a), 0 when they do not, negative on error. Needs no context. */
int ard_find_boundary(const int* box_a6, const int* box_b6, int* bnd_info9);
/* The topology part of Simulation::Simulation (ARD/simulation.cpp:19-266): boundaries between
 * the DCT partitions added so far (Boundary::FindBoundary, ARD/boundary.cpp:289-334), border
 * masks (Partition::AddBoundary, ARD/partition.cpp:96-143) and the PML slabs with their boundaries. */
int ard_build_topology(ard_ctx* ctx);

/* A source at global cell (x,y,z). wave == NULL: GaussianSource (ARD/gaussian_source.cpp:16-20);
 * otherwise sample t of the step t is wave[t] (0 past n). Assigned like ARD/simulation.cpp:37-48. */
int ard_add_source(ard_ctx* ctx, int x, int y, int z, const float* wave, int n);
/* Recorder at global cell (x,y,z); owner found as Recorder::FindPartition (ARD/recorder.cpp:26-38). */
int ard_add_recorder(ard_ctx* ctx, int x, int y, int z);

/* Allocates device state and tables. After it, topology calls fail with ARD_ERR_STATE.
 * Limit of the batched form (the reference has none; the plain form has none either):
 * fewer than 2^30 cells per DCT partition; beyond it ARD_ERR_UNSUPPORTED (split the
 * partition or use the plain form). Point sources and recorders per partition are not
 * limited: the modal pass takes up to 4 of each in a DCT partition's interior, and a
 * partition with more adds one pass over its modes for the extra sources (before the
 * modal pass) and one for the extra recorders (after it). */
int ard_finalize(ard_ctx* ctx);

/* Topology queries. part_info[9] = is_pml, pml_type, parent, x0,y0,z0, w,h,d.
 * bnd_info[9] = type, a, b, xs,xe,ys,ye,zs,ze. */
int ard_num_partitions(ard_ctx* ctx);
int ard_num_boundaries(ard_ctx* ctx);
int ard_num_recorders(ard_ctx* ctx);
int ard_partition_info(ard_ctx* ctx, int part, int* info9);
int ard_boundary_info(ard_ctx* ctx, int bnd, int* info9);
int ard_recorder_partition(ard_ctx* ctx, int rec);

/* Non-finite state check: reads every state array of the context (DCT force, pressure and
 * modes, or in the batched form the modes and the face-slab force and pressure; PML pressure,
 * phi and force) and sets *bad_partition to the lowest partition id holding a NaN or an
 * infinity, -1 if none. Synchronous; one read of the state (about 0.2 ms for 80M cells), so a
 * caller checks every so many steps. No counterpart in the reference, which has no failure
 * detection (a diverging run fills its traces with NaN, ARD/recorder.cpp:40-57). */
int ard_check_finite(ard_ctx* ctx, int* bad_partition);

/* n calls of Simulation::Update (ARD/simulation.cpp:300-317) followed by the recorder point
 * sample of Recorder::RecordField (ARD/recorder.cpp:40-57), all on the device. Asynchronous
 * with respect to the host. */
int ard_step(ard_ctx* ctx, int n);
/* Simulation::time_step_ (ARD/simulation.h:48): number of steps taken. */
int ard_time_step(ard_ctx* ctx);
/* Recorder samples of steps [t0, t0+nt) into out[nt][num_recorders] (the response_<id>.txt
 * column of ARD/recorder.cpp:55). Synchronises. */
int ard_fetch_recorders(ard_ctx* ctx, float* out, int t0, int nt);

/* Recorder cubes: every step each recorder also samples the 10 x 10 x 10 cells (x+k, y+j, z+i),
 * i, j, k in [-5, 5), that Recorder::RecordField writes to out_<id>.txt (ARD/recorder.cpp:44-54),
 * addressed like the point sample (ard_params.recorder_mode). Before ard_finalize; ard_finalize
 * fails with ARD_ERR_RANGE if a cube reaches outside its partition's array (undefined in the
 * reference). The batched form evaluates the cubes from the modes on the device. */
int ard_set_recorder_cubes(ard_ctx* ctx, int on);
/* Cube samples of steps [t0, t0+nt) into out[nt][num_recorders][1000], element
 * 100 (i+5) + 10 (j+5) + (k+5): the reference's write order. Synchronises. */
int ard_fetch_recorder_cubes(ard_ctx* ctx, float* out, int t0, int nt);

/* The visualisation of Simulation::Update (ARD/simulation.cpp:320-420) on the device: the pressure
 * on one plane, colour-mapped (white at 0, blue for p >= 0.1, red for p <= -0.1) into the
 * context's pixel array, which persists between calls as the reference's pixels_ does (cells no
 * drawn partition covers keep their last colour; zero at first).
 *   view ARD_VIEW_XY: plane z = coord; pixel (y - y_start) * size_x + (x - x_start)
 *   view ARD_VIEW_YZ: plane x = coord; pixel (z - z_start) * size_y + (y - y_start)
 * over the bounding box of all partitions, PML slabs included. coord < 0: the first source's z
 * (xy) or x (yz), the reference's choice. Partitions are drawn in list order and read the plane
 * with its global coordinate as their local index, as Partition::get_xy_plane / get_yz_plane
 * do (ARD/partition.cpp:58-78); cells past a DCT partition's array (undefined there) read 0.
 * Pixels are SDL_MapRGBA(RGBA8888, 255, r, g, b) values: 0xFF000000 | r << 16 | g << 8 | b.
 * dims[2] receives (columns, rows): (size_x, size_y) or (size_y, size_z); pixels receives
 * rows * columns values (pixels NULL: only dims, nothing drawn). Synchronises. */
#define ARD_VIEW_XY 0
#define ARD_VIEW_YZ 1
int ard_render_view(ard_ctx* ctx, int view, int coord, unsigned int* pixels, int* dims);

/* Partition::get_pressure_field (ARD/partition.h:52): the whole pressure volume, w*h*d floats.
 * Synchronises. The batched form evaluates DCT partitions from their modes on demand. */
int ard_copy_pressure_field(ard_ctx* ctx, int part, float* out);
/* Partition::get_pressure(x,y,z) (ARD/partition.h:56), partition-local cell. */
int ard_get_pressure(ard_ctx* ctx, int part, int x, int y, int z, float* out);
/* Partition::set_force(x,y,z,f) (ARD/partition.h:57), partition-local cell. DCT forces
 * persist across steps; PML forces are cleared by the next update (ARD/pml_partition.cpp:199). */
int ard_set_force(ard_ctx* ctx, int part, int x, int y, int z, float f);

/* Scaled 3D DCT of a device volume: VkFFT_DCT(vkGPU, type, w, h, d, in, out).execute()
 * (ARD/utils_VkFFT.h:38-60, ARD/utils_VkFFT.cpp:623-691), which computes FFTW's unnormalised
 * REDFT10 (type 2) or REDFT01 (type 3) along x, y and z (ARD/dct_volume.cpp:17,19), followed by
 * DctVolume's scaling: out = scale * transform(in). scale = 1/(2 sqrt(2 w h d)) (type 2) and
 * 1/sqrt(2 w h d) (type 3) give DctVolume::ExecuteDct / ExecuteIdct (ARD/dct_volume.cpp:38-68);
 * scale = 1 gives VkFFT_DCT's output. in and out are device pointers to w*h*d floats, z-major
 * with x fastest, and must not overlap (DctVolume transforms out of place). stream is a
 * hipStream_t (NULL: the null stream). Allocates a workspace for the call (the intermediate
 * volume, the cosine tables, the pass descriptors: dct3d workspace bytes, about 4 w h d + 4 (w^2
 * + h^2 + d^2)) and returns when the transform is done. */
int ard_dct3d(int device, int type, int w, int h, int d, const float* in, float* out, float scale, void* stream);

/* ard_dct3d for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference. Same arguments and
 * result as ard_dct3d; the workspace is a static device array of ARD_DCT3D_SCRATCH_BYTES per GPU
 * (held until the process ends), so volumes whose workspace exceeds it (about 4 w h d +
 * 4 (w^2 + h^2 + d^2) bytes: 250^3 cells fit, 256^3 do not) return ARD_ERR_RANGE. Calls no hipMalloc or
 * hipFree; asynchronous on `stream` (it returns once the work is enqueued). Every call on one
 * GPU shares its workspace: calls that may overlap on one GPU (other streams, other threads)
 * race, so serialise them. */
#define ARD_DCT3D_SCRATCH_BYTES (64u << 20)
int ard_dct3d_noalloc(int device, int type, int w, int h, int d, const float* in, float* out, float scale,
                      void* stream);

/* Blocks until all work queued on the context's stream is done. */
int ard_synchronize(ard_ctx* ctx);

/* Kernel timing (no counterpart in the reference, which only prints wall-clock
 * totals, ARD/main.cpp:49,124-126). While on, ard_step launches each phase
 * directly on the context's stream between two HIP events (no graph replay)
 * and accumulates the per-launch durations; ard_kernel_times returns, per
 * phase, the average duration of one launch in ms and the number of launches
 * since timing was switched on. Phases: */
#define ARD_PH_SOURCES 0
#define ARD_PH_PML 1
#define ARD_PH_FOLD 2 /* batched form, no longer launched: the forward transforms read the raw force slabs */
#define ARD_PH_FACE_FWD1 3
#define ARD_PH_FACE_FWD2 4
#define ARD_PH_MODAL 5
#define ARD_PH_UNFOLD 6
#define ARD_PH_FACE_INV1 7
#define ARD_PH_FACE_INV2 8
#define ARD_PH_INTERFACE 9
#define ARD_PH_RECORD 10
#define ARD_PH_DCT_FWD 11 /* plain form: the three DCT-II passes with the modal update */
#define ARD_PH_DCT_INV 12 /* plain form: the three DCT-III passes */
#define ARD_NUM_PHASES 13
int ard_set_kernel_timing(ard_ctx* ctx, int on);
int ard_kernel_times(ard_ctx* ctx, float* avg_ms, int* launches, int nphase);

/* Self-test of the dense transform kernels (no counterpart in the reference):
 * C[M][N] = A[M][K] * B on device `device`, where B is K x N row-major
 * (b_layout 0) or N x K row-major (b_layout 1). variant: 0 vector FMA,
 * 1 matrix cores FP32, 2 matrix cores BF16 split precision (each operand
 * x = hi + lo in BF16, three products, FP32 accumulate). Synchronous;
 * allocates and frees its own buffers. */
int ard_selftest_transform(int device, int variant, int b_layout, int M, int N, int K, const float* A, const float* B,
                           float* C);
/* Self-test of the even/odd folded transform kernel (variant 3; no
 * counterpart in the reference): one unnormalised transform of length L along
 * the first axis of X, FFTW's REDFT10 (inverse = 0) or REDFT01 (inverse = 1)
 * (ARD/dct_volume.cpp:17,19), for N columns. b_layout 0: X and Y are [L][N]
 * row-major; 1: [N][L]. Built exactly as the batched form builds its face
 * passes. L >= 6 (the smallest partition side, which the kernel assumes).
 * Synchronous; allocates and frees its own buffers. */
int ard_selftest_fold(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y);
/* The same self-test for the folded transform in BF16 split precision on the
 * matrix cores (variant 2's batched face passes: every operand x = hi + lo in
 * BF16, three products on v_mfma_f32_32x32x16_bf16, FP32 accumulate). No
 * counterpart in the reference. */
int ard_selftest_fold_bf16(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y);
/* The same self-test for the folded FP32 transform with each wave's operands
 * loaded straight into registers (variant 5's batched face passes: no LDS
 * staging, no barriers). No counterpart in the reference. */
int ard_selftest_fold_direct(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y);
/* Self-test of the FFT-style transform kernel (variant 4; no counterpart in the
 * reference, whose FFTW plans it replaces, ARD/dct_volume.cpp:17,19): the same
 * transform as ard_selftest_fold, for L >= 2 whose prime factors are at most 13
 * (ARD_ERR_UNSUPPORTED otherwise). Synchronous; allocates and frees its own
 * buffers. */
int ard_selftest_fft(int device, int inverse, int b_layout, int L, int N, const float* X, float* Y);
/* Selects the dense transform kernels used by contexts finalized afterwards
 * (0 vector FMA, 1 matrix cores FP32, 2 matrix cores BF16 split, 3 matrix
 * cores FP32 with the even/odd folding c(k, n-1-j) = (-1)^k c(k, j), half the
 * multiply-adds; 5 the same folded FP32 transforms with each wave's operands
 * loaded straight into registers, no LDS staging; 3 and 5 apply to the
 * batched form's face passes, other passes use 1). Default 1. The batched
 * form's autotuner may choose among 0, 1, 3, 4 and 5 (same precision); it
 * keeps 2. 4: FFT-style passes (Makhoul's reordering and mixed-radix FFTs in
 * LDS, O(L log L) per line) for every pass, in either form, when every axis
 * length's prime factors are at most 13 (else 1). */
int ard_set_transform_variant(int variant);
/* Tiling of the batched form after ard_finalize (chosen by the autotuner when
 * ard_params.autotune is set): out[4] = rows per wave, kz planes per block,
 * transform variant, number of candidates timed. For a finalized plain-form
 * context: out[2] = the transform variant of its passes, the rest 0. No
 * counterpart in the reference. */
int ard_batched_config(ard_ctx* ctx, int* out4);
/* Sets the batched form's modal-pass tiling after ard_finalize: rows per wave
 * (2, 4, 6 or 8) and kz planes per block (16 to 4096, rounded to an even
 * chunk per partition), as the autotuner would. ARD_ERR_STATE for a context
 * that is not a finalized batched one, ARD_ERR_INVALID for other values. No
 * counterpart in the reference: it pins the tiling, for tests and callers that
 * skip the autotuner. */
int ard_set_batched_tiling(ard_ctx* ctx, int rows_per_wave, int kz_per_block);
/* Schedule options of the batched form after ard_finalize (chosen by the
 * autotuner when ard_params.autotune is set). Results do not depend on them.
 *   ARD_OPT_RING      0 or 2: planes of mode rows the modal pass keeps in
 *                     flight per wave; 2 streams them into LDS by
 *                     direct-to-LDS loads, 0 (default) keeps one plane in
 *                     registers. With 4 rows per wave only (other tilings
 *                     use 0).
 *   ARD_OPT_PML_SHAPE 0 or 1: tile shape of the tiled PML update across a
 *                     slab's normal, 32 x 8 cells for y- and z-normal slabs
 *                     and 16 x 16 for x-normal ones (0, default), or the
 *                     other way round (1).
 *   ARD_OPT_PML_SPLIT 0..100: percent of the PML update's tile blocks
 *                     (leading slabs, whole slabs) launched beside the
 *                     forward face transforms; the rest run beside the
 *                     unfold and the inverse face transforms (default 0:
 *                     the whole PML beside the unfold and inverse ones).
 * ARD_ERR_STATE for a context that is not a finalized batched one,
 * ARD_ERR_INVALID for other options or values. No counterpart in the
 * reference: they schedule the step of ARD/simulation.cpp:300-317 on one
 * device. */
#define ARD_OPT_RING 2
#define ARD_OPT_PML_SHAPE 3
#define ARD_OPT_PML_SPLIT 4
int ard_set_batched_option(ard_ctx* ctx, int option, int value);
int ard_batched_option(ard_ctx* ctx, int option, int* value);

/* Discrete adjoint of the step (no counterpart in the reference, which
 * has no gradients; it transposes the step of ARD/simulation.cpp:300-317).
 * Runs `nsteps` steps from the zero state with a checkpoint every
 * checkpoint_stride steps (<= 0: ceil(sqrt(nsteps))), then the reverse sweep,
 * and returns the gradients of a loss L(y) of the recorder traces given
 * dl_drec[t][r] = dL/dy(t, r) (the layout of ard_fetch_recorders):
 *   dl_dwave[s][t]  dL/d(force sample of source s injected at step t), for every
 *                   source in registration order (zero rows for sources outside
 *                   every DCT partition); may be NULL
 *   dl_dscale[6]    dL/d(damping scale of each PML face, ard_set_pml_scale); may be NULL
 * A plain-form context is left in its forward state after nsteps (time step
 * nsteps, the traces of that run). A batched context runs the adjoint on a
 * plain-form twin built on its first call from the context's topology,
 * parameters and PML damping scales (the same discrete step; gradients agree
 * with its own forward run to FP32 rounding); its own state is not touched.
 * Forces set with ard_set_force are not part of the run (it starts from the
 * zero state). Allocates its checkpoints for the call. */
int ard_adjoint(ard_ctx* ctx, int nsteps, const float* dl_drec, float* dl_dwave, float* dl_dscale,
                int checkpoint_stride);

/* A batch of independent rooms advanced together (no counterpart in the
 * reference, whose Simulation holds one scene, ARD/simulation.cpp:14-294).
 * Each member is a context whose partitions, sources and recorders are
 * registered and whose topology is built (ard_build_topology), not finalized;
 * all members share device, h, dt, c0, PML layers and scales, absorption and
 * form. The batch lays the rooms side by side in one context of its own (far
 * enough apart that no two touch, so every room keeps its own topology and
 * recorder addressing) and finalizes it: one step graph per step for all
 * rooms in the batched form. Members are left unchanged and stay owned by
 * the caller. A batch of host-only members is built but not finalized. */
typedef struct ard_batch ard_batch;
int ard_batch_create(ard_ctx* const* members, int n, ard_batch** out);
int ard_batch_step(ard_batch* batch, int nsteps);
/* Room `room`'s recorders, [nt][its recorder count], steps t0 .. t0+nt-1. */
int ard_batch_fetch_recorders(ard_batch* batch, int room, float* out, int t0, int nt);
/* The batch's own context (partition info, kernel timing, fields); owned by the batch. */
int ard_batch_context(ard_batch* batch, ard_ctx** out);
/* Id in the batch's context of the room's DCT partition `part` (the room's own id). */
int ard_batch_room_partition(ard_batch* batch, int room, int part);
int ard_batch_destroy(ard_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* ARD_H_ */
