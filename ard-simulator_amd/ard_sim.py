"""Python mirror of the reference's host API over libard.so (ctypes).

The reference drives a scene with ``Simulation(partitions, sources)``,
``Simulation::Update()`` and ``Recorder::RecordField`` (ARD/main.cpp:62-150).
:class:`Simulation` below keeps those names and meanings on top of the C-ABI
declared in include/ard.h; every call goes to the HIP library. There is no CPU
fallback: constructing a GPU simulation without libard.so built raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libard.so")

ARD_OK = 0
FORM_PLAIN, FORM_BATCHED = 0, 1
REC_REFERENCE, REC_LOCAL = 0, 1
OPT_RING = 2  # ard_set_batched_option
OPT_PML_SHAPE = 3
OPT_PML_SPLIT = 4

EXPORTED = [
    "ard_ctx_create", "ard_ctx_destroy", "ard_last_error", "ard_set_pml_scale", "ard_add_dct_partition",
    "ard_add_pml_partition", "ard_add_boundary", "ard_build_topology", "ard_add_source", "ard_add_recorder",
    "ard_finalize", "ard_num_partitions", "ard_num_boundaries", "ard_num_recorders", "ard_partition_info",
    "ard_boundary_info", "ard_recorder_partition", "ard_step", "ard_time_step", "ard_fetch_recorders",
    "ard_copy_pressure_field", "ard_get_pressure", "ard_set_force", "ard_synchronize", "ard_set_kernel_timing",
    "ard_kernel_times", "ard_selftest_transform", "ard_set_transform_variant", "ard_batched_config",
    "ard_set_batched_tiling", "ard_set_batched_option", "ard_batched_option",
    "ard_adjoint", "ard_batch_create", "ard_batch_step", "ard_batch_fetch_recorders", "ard_batch_context",
    "ard_batch_room_partition", "ard_batch_destroy", "ard_set_recorder_cubes", "ard_fetch_recorder_cubes",
    "ard_render_view", "ard_selftest_fold", "ard_dct3d", "ard_dct3d_noalloc", "ard_selftest_fft",
    "ard_selftest_fold_bf16", "ard_check_finite", "ard_find_boundary", "ard_selftest_fold_direct",
]
PHASES = ["sources", "pml", "fold", "face_fwd1", "face_fwd2", "modal", "unfold", "face_inv1", "face_inv2",
          "interface", "record", "dct_fwd", "dct_inv"]


class ArdError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_float), ("dt", ctypes.c_float), ("c0", ctypes.c_float), ("pml_layers", ctypes.c_int),
        ("absorption", ctypes.c_float), ("form", ctypes.c_int), ("recorder_mode", ctypes.c_int),
        ("autotune", ctypes.c_int), ("reserved", ctypes.c_int * 8),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ArdError("libard.so is not built (run __graft_entry__.build())")
        # One HIP runtime per process: torch ships its own libamdhip64 (the
        # soname libard.so needs). Loaded first, it serves both, so torch's
        # device pointers and streams are valid here; loaded after ROCm's,
        # it would be a second runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ard_ctx_create.argtypes = [i, ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.ard_ctx_destroy.argtypes = [vp]
        L.ard_last_error.restype = ctypes.c_char_p
        L.ard_set_pml_scale.argtypes = [vp, fp]
        L.ard_add_dct_partition.argtypes = [vp, i, i, i, i, i, i]
        L.ard_add_pml_partition.argtypes = [vp, i, i, i, i, i, i, i, i]
        L.ard_add_boundary.argtypes = [vp, i, i, i, i, i, i, i, i, i, f]
        L.ard_build_topology.argtypes = [vp]
        L.ard_add_source.argtypes = [vp, i, i, i, fp, i]
        L.ard_add_recorder.argtypes = [vp, i, i, i]
        L.ard_finalize.argtypes = [vp]
        for n in ("ard_num_partitions", "ard_num_boundaries", "ard_num_recorders", "ard_time_step",
                  "ard_synchronize"):
            getattr(L, n).argtypes = [vp]
        L.ard_partition_info.argtypes = [vp, i, ip]
        L.ard_boundary_info.argtypes = [vp, i, ip]
        L.ard_recorder_partition.argtypes = [vp, i]
        L.ard_step.argtypes = [vp, i]
        L.ard_fetch_recorders.argtypes = [vp, fp, i, i]
        L.ard_copy_pressure_field.argtypes = [vp, i, fp]
        L.ard_get_pressure.argtypes = [vp, i, i, i, i, fp]
        L.ard_set_force.argtypes = [vp, i, i, i, i, f]
        L.ard_set_kernel_timing.argtypes = [vp, i]
        L.ard_kernel_times.argtypes = [vp, fp, ip, i]
        L.ard_selftest_transform.argtypes = [i, i, i, i, i, i, fp, fp, fp]
        L.ard_set_transform_variant.argtypes = [i]
        L.ard_selftest_fold.argtypes = [i, i, i, i, i, fp, fp]
        L.ard_selftest_fft.argtypes = [i, i, i, i, i, fp, fp]
        L.ard_selftest_fold_bf16.argtypes = [i, i, i, i, i, fp, fp]
        L.ard_selftest_fold_direct.argtypes = [i, i, i, i, i, fp, fp]
        L.ard_dct3d.argtypes = [i, i, i, i, i, vp, vp, f, vp]
        L.ard_batched_config.argtypes = [vp, ip]
        L.ard_set_batched_tiling.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.ard_set_batched_option.argtypes = [vp, i, i]
        L.ard_check_finite.argtypes = [vp, ip]
        L.ard_find_boundary.argtypes = [ip, ip, ip]
        L.ard_batched_option.argtypes = [vp, i, ip]
        L.ard_adjoint.argtypes = [vp, i, fp, fp, fp, i]
        L.ard_batch_create.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(vp)]
        L.ard_batch_step.argtypes = [vp, i]
        L.ard_batch_fetch_recorders.argtypes = [vp, i, fp, i, i]
        L.ard_batch_context.argtypes = [vp, ctypes.POINTER(vp)]
        L.ard_batch_room_partition.argtypes = [vp, i, i]
        L.ard_batch_destroy.argtypes = [vp]
        L.ard_set_recorder_cubes.argtypes = [vp, i]
        L.ard_fetch_recorder_cubes.argtypes = [vp, fp, i, i]
        L.ard_render_view.argtypes = [vp, i, i, ctypes.POINTER(ctypes.c_uint32), ip]
        _lib = L
    return _lib


def _check(r):
    if r < 0:
        raise ArdError("libard error %d: %s" % (r, lib().ard_last_error().decode()))
    return r


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def dct3d(inp, out, dct_type=2, scale=1.0, stream=None, device=0):
    """out = scale * DCT-II (dct_type 2) or DCT-III (3) of the device volume inp
    ([d][h][w] float32, contiguous; torch tensors or raw device addresses with
    shape given by `out`'s tensor) through ard_dct3d. stream: a HIP stream handle
    (int) or None for the null stream."""
    d, h, w = out.shape
    _check(lib().ard_dct3d(device, dct_type, w, h, d, ctypes.c_void_p(inp.data_ptr()),
                           ctypes.c_void_p(out.data_ptr()), float(scale), ctypes.c_void_p(stream or 0)))
    return out


def selftest_fold(X, inverse=False, b_layout=0, device=0, entry="ard_selftest_fold"):
    """Unnormalised REDFT10 (or REDFT01 if inverse) along the length-L axis of X
    through the folded matrix-core kernel (or, with entry="ard_selftest_fft",
    the FFT-style one): X is [L][N] (b_layout 0) or [N][L]."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    L, N = X.shape if b_layout == 0 else X.shape[::-1]
    Y = np.zeros_like(X)
    _check(getattr(lib(), entry)(device, int(inverse), b_layout, L, N, _fp(X), _fp(Y)))
    return Y


def find_boundary(box_a, box_b):
    """Boundary::FindBoundary on two boxes (x0, y0, z0, w, h, d) in cells:
    (type, xs, xe, ys, ye, zs, ze) or None (ard_find_boundary)."""
    a = (ctypes.c_int * 6)(*[int(v) for v in box_a])
    b = (ctypes.c_int * 6)(*[int(v) for v in box_b])
    out = (ctypes.c_int * 9)()
    r = lib().ard_find_boundary(a, b, out)
    _check(r)
    return None if r == 0 else (out[0],) + tuple(out[3:9])


def selftest_transform(A, B, variant=1, b_layout=0, device=0):
    """C = A @ B through the library's dense transform kernel (B given as K x N,
    or as N x K when b_layout is 1)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, K = A.shape
    N = B.shape[1] if b_layout == 0 else B.shape[0]
    C = np.zeros((M, N), dtype=np.float32)
    _check(lib().ard_selftest_transform(device, variant, b_layout, M, N, K, _fp(A), _fp(B), _fp(C)))
    return C


def set_transform_variant(v):
    _check(lib().ard_set_transform_variant(v))


class Simulation:
    """A scene on one MI355X: the reference's Simulation with Update() on the GPU.

    ``scene`` is an ard_scenes.Scene. ``device=-1`` builds a host-only context
    whose topology can be inspected without a GPU.
    """

    def __init__(self, scene, form=FORM_PLAIN, device=0, pml_scale=None, finalize=True, autotune=False,
                 cubes=False):
        self.L = lib()
        self.scene = scene
        self.form = form
        p = Params(scene.h, scene.dt, scene.c0, scene.pml_layers, scene.absorption, form, scene.recorder_mode,
                   int(autotune))
        h = ctypes.c_void_p()
        _check(self.L.ard_ctx_create(device, ctypes.byref(p), ctypes.byref(h)))
        self.ctx = h
        self._keep = []
        if pml_scale is not None:
            a = np.ascontiguousarray(pml_scale, dtype=np.float32)
            _check(self.L.ard_set_pml_scale(self.ctx, _fp(a)))
        for b in scene.boxes:
            _check(self.L.ard_add_dct_partition(self.ctx, *b))
        _check(self.L.ard_build_topology(self.ctx))
        for s in scene.sources:
            if s.wave is None:
                _check(self.L.ard_add_source(self.ctx, s.x, s.y, s.z, None, 0))
            else:
                w = np.ascontiguousarray(s.wave, dtype=np.float32)
                self._keep.append(w)
                _check(self.L.ard_add_source(self.ctx, s.x, s.y, s.z, _fp(w), len(w)))
        for r in scene.recorders:
            _check(self.L.ard_add_recorder(self.ctx, *r))
        self.nrec = len(scene.recorders)
        if cubes:
            _check(self.L.ard_set_recorder_cubes(self.ctx, 1))
        if finalize:
            _check(self.L.ard_finalize(self.ctx))

    def close(self):
        if getattr(self, "ctx", None):
            self.L.ard_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    # --- topology -------------------------------------------------------
    def partitions(self):
        out, buf = [], (ctypes.c_int * 9)()
        for i in range(_check(self.L.ard_num_partitions(self.ctx))):
            _check(self.L.ard_partition_info(self.ctx, i, buf))
            out.append(tuple(buf))
        return out

    def boundaries(self):
        out, buf = [], (ctypes.c_int * 9)()
        for i in range(_check(self.L.ard_num_boundaries(self.ctx))):
            _check(self.L.ard_boundary_info(self.ctx, i, buf))
            out.append(tuple(buf))
        return out

    def recorder_partition(self, r):
        return _check(self.L.ard_recorder_partition(self.ctx, r))

    # --- stepping ---------------------------------------------------------
    def update(self, n=1):
        """n calls of Simulation::Update (asynchronous)."""
        _check(self.L.ard_step(self.ctx, n))

    @property
    def time_step(self):
        return _check(self.L.ard_time_step(self.ctx))

    def synchronize(self):
        _check(self.L.ard_synchronize(self.ctx))

    def recorders(self, t0=0, nt=None):
        """Recorder samples [nt][num_recorders] of steps t0..t0+nt-1."""
        if nt is None:
            nt = self.time_step - t0
        out = np.zeros((nt, max(self.nrec, 1)), dtype=np.float32)
        _check(self.L.ard_fetch_recorders(self.ctx, _fp(out), t0, nt))
        return out[:, : self.nrec]

    def recorder_cubes(self, t0=0, nt=None):
        """Recorder cubes [nt][num_recorders][10][10][10] (z, y, x offsets -5..4)
        of steps t0..t0+nt-1: the out_<id>.txt rows of Recorder::RecordField."""
        if nt is None:
            nt = self.time_step - t0
        out = np.zeros((nt, self.nrec, 10, 10, 10), dtype=np.float32)
        _check(self.L.ard_fetch_recorder_cubes(self.ctx, _fp(out), t0, nt))
        return out

    def render_view(self, view=0, coord=-1):
        """The visualisation plane of Simulation::Update (ARD/simulation.cpp:320-420):
        uint32 pixels [rows][columns] of view 0 (xy at z = coord) or 1 (yz at
        x = coord); coord < 0 places it at the first source."""
        dims = (ctypes.c_int * 2)()
        _check(self.L.ard_render_view(self.ctx, view, coord, None, dims))
        out = np.zeros((dims[1], dims[0]), dtype=np.uint32)
        _check(self.L.ard_render_view(self.ctx, view, coord,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), dims))
        return out

    def pressure_field(self, part):
        info = self.partitions()[part]
        w, h, d = info[6:9]
        out = np.zeros((d, h, w), dtype=np.float32)
        _check(self.L.ard_copy_pressure_field(self.ctx, part, _fp(out)))
        return out

    def get_pressure(self, part, x, y, z):
        v = ctypes.c_float()
        _check(self.L.ard_get_pressure(self.ctx, part, x, y, z, ctypes.byref(v)))
        return v.value

    def set_force(self, part, x, y, z, f):
        _check(self.L.ard_set_force(self.ctx, part, x, y, z, f))

    def set_kernel_timing(self, on):
        _check(self.L.ard_set_kernel_timing(self.ctx, int(on)))

    def kernel_times(self):
        """{phase: (average ms per launch, launches)} since timing was switched on."""
        n = len(PHASES)
        ms = np.zeros(n, dtype=np.float32)
        cnt = np.zeros(n, dtype=np.int32)
        _check(self.L.ard_kernel_times(self.ctx, _fp(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n))
        return {PHASES[i]: (float(ms[i]), int(cnt[i])) for i in range(n) if cnt[i]}

    def adjoint(self, dl_drec, checkpoint_stride=0):
        """Gradients of a recorder loss by the discrete adjoint.

        dl_drec: [nsteps][nrec] array of dL/dy. Re-runs nsteps steps from the
        zero state with checkpoints, then the reverse sweep (a batched context
        runs both on a plain-form twin of itself). Returns
        (dL/dwave [nsources][nsteps] float32, dL/dscale [6] float32)."""
        g = np.ascontiguousarray(dl_drec, dtype=np.float32)
        nsteps = g.shape[0]
        if g.ndim != 2 or g.shape[1] != self.nrec:
            raise ValueError("dl_drec must be [nsteps][%d]" % self.nrec)
        dw = np.zeros((len(self.scene.sources), nsteps), dtype=np.float32)
        ds = np.zeros(6, dtype=np.float32)
        _check(self.L.ard_adjoint(self.ctx, nsteps, _fp(g), _fp(dw), _fp(ds), int(checkpoint_stride)))
        return dw, ds

    def set_batched_tiling(self, rows_per_wave, kz_per_block):
        """Pins the modal-pass tiling of a finalized batched context."""
        _check(self.L.ard_set_batched_tiling(self.ctx, int(rows_per_wave), int(kz_per_block)))

    def batched_config(self):
        """(rows per wave, kz planes per block, transform variant, candidates timed)."""
        out = (ctypes.c_int * 4)()
        _check(self.L.ard_batched_config(self.ctx, out))
        return tuple(out)

    def check_finite(self):
        """Lowest partition id whose state holds a NaN or an infinity, or -1 (ard_check_finite)."""
        out = ctypes.c_int(0)
        _check(self.L.ard_check_finite(self.ctx, ctypes.byref(out)))
        return out.value

    def set_batched_option(self, option, value):
        """Sets a schedule option of a finalized batched context (OPT_RING, OPT_PML_SHAPE, OPT_PML_SPLIT)."""
        _check(self.L.ard_set_batched_option(self.ctx, int(option), int(value)))

    def batched_option(self, option):
        """A schedule option of a finalized batched context (OPT_RING, OPT_PML_SHAPE, OPT_PML_SPLIT)."""
        out = ctypes.c_int(0)
        _check(self.L.ard_batched_option(self.ctx, int(option), ctypes.byref(out)))
        return out.value


class Batch:
    """A batch of independent rooms advanced together (ard_batch_create).

    ``scenes``: ard_scenes.Scene objects sharing h, dt, c0, PML and form.
    Each is registered in a member context (topology built, not finalized);
    the batch owns one merged context and frees the members at once."""

    def __init__(self, scenes, form=FORM_BATCHED, device=0, autotune=False):
        self.L = lib()
        self.members = [Simulation(sc, form=form, device=device, finalize=False, autotune=autotune)
                        for sc in scenes]
        arr = (ctypes.c_void_p * len(scenes))(*[m.ctx for m in self.members])
        h = ctypes.c_void_p()
        _check(self.L.ard_batch_create(arr, len(scenes), ctypes.byref(h)))
        self.batch = h
        self.nrec = [len(sc.recorders) for sc in scenes]
        for m in self.members:
            m.close()
        self.members = []
        c = ctypes.c_void_p()
        _check(self.L.ard_batch_context(self.batch, ctypes.byref(c)))
        self.ctx = c

    def close(self):
        if getattr(self, "batch", None):
            self.L.ard_batch_destroy(self.batch)
            self.batch = None

    def __del__(self):
        self.close()

    def update(self, n=1):
        _check(self.L.ard_batch_step(self.batch, n))

    def set_kernel_timing(self, on):
        """Per-phase kernel timing of the batch's context (ard_set_kernel_timing)."""
        _check(self.L.ard_set_kernel_timing(self.ctx, int(on)))

    def kernel_times(self):
        """{phase: (average ms per launch, launches)} since timing was switched on."""
        n = len(PHASES)
        ms = np.zeros(n, dtype=np.float32)
        cnt = np.zeros(n, dtype=np.int32)
        _check(self.L.ard_kernel_times(self.ctx, _fp(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n))
        return {PHASES[i]: (float(ms[i]), int(cnt[i])) for i in range(n) if cnt[i]}

    def synchronize(self):
        _check(self.L.ard_synchronize(self.ctx))

    def batched_config(self):
        """(rows per wave, kz planes per block, transform variant, candidates timed)."""
        out = (ctypes.c_int * 4)()
        _check(self.L.ard_batched_config(self.ctx, out))
        return tuple(out)

    def recorders(self, room, t0, nt):
        out = np.zeros((nt, max(self.nrec[room], 1)), dtype=np.float32)
        _check(self.L.ard_batch_fetch_recorders(self.batch, room, _fp(out), t0, nt))
        return out[:, : self.nrec[room]]

    def partitions(self):
        out, buf = [], (ctypes.c_int * 9)()
        for i in range(_check(self.L.ard_num_partitions(self.ctx))):
            _check(self.L.ard_partition_info(self.ctx, i, buf))
            out.append(tuple(buf))
        return out

    def boundaries(self):
        out, buf = [], (ctypes.c_int * 9)()
        for i in range(_check(self.L.ard_num_boundaries(self.ctx))):
            _check(self.L.ard_boundary_info(self.ctx, i, buf))
            out.append(tuple(buf))
        return out

    def room_partition(self, room, part):
        return _check(self.L.ard_batch_room_partition(self.batch, room, part))
