"""ctypes wrapper of the CPU oracle (oracle/ard_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product path never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libard_oracle.so")
_lib = None

P_LEFT, P_RIGHT, P_TOP, P_BOTTOM, P_FRONT, P_BACK = range(6)
X_BOUNDARY, Y_BOUNDARY, Z_BOUNDARY = range(3)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, d, d, d, i, d]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_pml_scale.argtypes = [vp, dp]
        L.orc_set_recorder_mode.argtypes = [vp, i]
        L.orc_add_dct.argtypes = [vp, i, i, i, i, i, i]
        L.orc_add_source.argtypes = [vp, i, i, i, dp, i]
        L.orc_add_recorder.argtypes = [vp, i, i, i]
        L.orc_finalize.argtypes = [vp]
        L.orc_num_partitions.argtypes = [vp]
        L.orc_num_boundaries.argtypes = [vp]
        L.orc_partition_info.argtypes = [vp, i, ip]
        L.orc_boundary_info.argtypes = [vp, i, ip]
        L.orc_recorder_partition.argtypes = [vp, i]
        L.orc_step.argtypes = [vp, i, dp]
        L.orc_get_field.argtypes = [vp, i, i, dp]
        L.orc_set_field.argtypes = [vp, i, i, dp]
        L.orc_dct3d.argtypes = [i, i, i, i, dp, dp, i, i]
        L.orc_set_fft_dct.argtypes = [vp, i]
        L.orc_mode_tables.argtypes = [i, d, d, d, i, i, i, dp, dp]
        L.orc_gaussian.argtypes = [i, d, d, d, i, dp]
        L.orc_pml_profile.argtypes = [i, d, d, i, i, i, i, i, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleSim:
    """The reference's Simulation (ARD/simulation.cpp) restated on the CPU."""

    def __init__(self, h, dt, c0, pml_layers=5, absorption=1.0, double=False):
        self.L = lib()
        self.h = self.L.orc_create(int(double), h, dt, c0, pml_layers, absorption)
        self.nrec = 0
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_pml_scale(self, scale6):
        a = np.ascontiguousarray(scale6, dtype=np.float64)
        assert self.L.orc_set_pml_scale(self.h, _dp(a)) == 0

    def set_fft_dct(self, on):
        """True: O(n log n) transforms (default); False: the dense per-axis sums."""
        assert self.L.orc_set_fft_dct(self.h, int(bool(on))) == 0

    def set_recorder_mode(self, local):
        assert self.L.orc_set_recorder_mode(self.h, int(local)) == 0

    def add_dct(self, xs, ys, zs, w, h, d):
        r = self.L.orc_add_dct(self.h, xs, ys, zs, w, h, d)
        assert r >= 0
        return r

    def add_source(self, x, y, z, wave=None):
        if wave is None:
            return self.L.orc_add_source(self.h, x, y, z, None, 0)
        a = np.ascontiguousarray(wave, dtype=np.float64)
        self._keep.append(a)
        return self.L.orc_add_source(self.h, x, y, z, _dp(a), len(a))

    def add_recorder(self, x, y, z):
        self.nrec += 1
        return self.L.orc_add_recorder(self.h, x, y, z)

    def finalize(self):
        r = self.L.orc_finalize(self.h)
        if r != 0:
            raise ValueError("oracle finalize failed: %d" % r)

    def partitions(self):
        out = []
        buf = (ctypes.c_int * 9)()
        for i in range(self.L.orc_num_partitions(self.h)):
            self.L.orc_partition_info(self.h, i, buf)
            out.append(tuple(buf))
        return out

    def boundaries(self):
        out = []
        buf = (ctypes.c_int * 9)()
        for i in range(self.L.orc_num_boundaries(self.h)):
            self.L.orc_boundary_info(self.h, i, buf)
            out.append(tuple(buf))
        return out

    def recorder_partition(self, r):
        return self.L.orc_recorder_partition(self.h, r)

    def step(self, n):
        out = np.zeros((n, max(self.nrec, 1)), dtype=np.float64)
        assert self.L.orc_step(self.h, n, _dp(out)) == 0
        return out[:, : self.nrec]

    def field(self, part, which=0):
        info = self.partitions()[part]
        w, h, d = info[6:9]
        out = np.zeros(w * h * d, dtype=np.float64)
        assert self.L.orc_get_field(self.h, part, which, _dp(out)) == 0
        return out.reshape(d, h, w)

    def set_field(self, part, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert self.L.orc_set_field(self.h, part, which, _dp(a)) == 0


def dct3d(x, inverse=False, double=True, fft=False):
    """Scaled ARD 3D DCT-II / DCT-III of x[d][h][w] (ARD/dct_volume.cpp:38-68),
    by the dense per-axis sums (fft=False) or by FFT (fft=True)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    d, h, w = x.shape
    out = np.zeros_like(x)
    lib().orc_dct3d(int(double), w, h, d, _dp(x), _dp(out), int(inverse), int(fft))
    return out


def mode_tables(dh, dt, c0, w, h, d, double=False):
    w2 = np.zeros(w * h * d)
    cwt = np.zeros(w * h * d)
    lib().orc_mode_tables(int(double), dh, dt, c0, w, h, d, _dp(w2), _dp(cwt))
    return w2.reshape(d, h, w), cwt.reshape(d, h, w)


def gaussian(dh, dt, c0, n, double=False):
    out = np.zeros(n)
    lib().orc_gaussian(int(double), dh, dt, c0, n, _dp(out))
    return out


def pml_profile(dh, c0, layers, ptype, w, h, d, double=False):
    out = np.zeros(3 * w * h * d)
    lib().orc_pml_profile(int(double), dh, c0, layers, ptype, w, h, d, _dp(out))
    return out.reshape(3, d, h, w)


def from_scene(scene, double=False, pml_scale=None, fft=True):
    """Build the oracle for an ard_scenes.Scene (fft: transforms by FFT or by dense sums)."""
    o = OracleSim(scene.h, scene.dt, scene.c0, scene.pml_layers, scene.absorption, double=double)
    o.set_fft_dct(fft)
    if pml_scale is not None:
        o.set_pml_scale(pml_scale)
    o.set_recorder_mode(scene.recorder_mode)
    for b in scene.boxes:
        o.add_dct(*b)
    for s in scene.sources:
        o.add_source(s.x, s.y, s.z, s.wave)
    for r in scene.recorders:
        o.add_recorder(*r)
    o.finalize()
    return o


def colour(p):
    """Pixel of one pressure value in Simulation::Update's visualisation
    (ARD/simulation.cpp:346-364), FP32 operations in the reference's order:
    norm = 0.5 * std::max(-1, std::min(1, 10 p)) + 0.5; norm >= 0.5 gives
    r = g = 255 - round(255 * 2 * (norm - 0.5)), b = 255, else r = 255,
    g = b = 255 - round(255 * (1 - 2 norm)); SDL_MapRGBA(RGBA8888, 255, r, g, b)
    = 0xFF000000 | r << 16 | g << 8 | b. round() is half away from zero."""
    f = np.float32
    v = np.asarray(p, np.float32) * f(10)
    v = np.where(v < f(1), v, f(1)).astype(np.float32)      # std::min(1.0f, v): (v < 1) ? v : 1
    v = np.where(f(-1) < v, v, f(-1)).astype(np.float32)    # std::max(-1.0f, v): (-1 < v) ? v : -1
    norm = (f(0.5) * v + f(0.5)).astype(np.float32)
    hi = norm >= f(0.5)

    def rnd(x):  # C round() of a non-negative float
        return np.floor(np.asarray(x, np.float64) + 0.5)

    a = 255 - rnd((f(255) * f(2)) * (norm - f(0.5)))
    b = 255 - rnd(f(255) * (f(1) - f(2) * norm))
    r = np.where(hi, a, 255).astype(np.uint32)
    g = np.where(hi, a, b).astype(np.uint32)
    bl = np.where(hi, 255, b).astype(np.uint32)
    return (np.uint32(0xFF000000) | (r << 16) | (g << 8) | bl).astype(np.uint32)


def render(parts, fields, view, coord, pixels=None):
    """The visualisation step of Simulation::Update (ARD/simulation.cpp:320-420).

    parts: partition info tuples (is_pml, type, parent, x0, y0, z0, w, h, d)
    in list order; fields: each partition's pressure [d][h][w]; view 0 draws
    the xy plane z = coord, view 1 the yz plane x = coord. A partition is drawn
    when start <= coord <= end along the view axis, reading its plane with
    coord as its local index (Partition::get_xy_plane / get_yz_plane,
    ARD/partition.cpp:58-78): a DCT partition through its flat array (cells
    past the array, undefined in the reference, read 0), a PML slab through
    GetIndex (out of range reads the zero sentinel, ARD/pml_partition.cpp:6-15).
    Pixels go to (y - y0) * size_x + (x - x0) (xy) or (z - z0) * size_y +
    (y - y0) (yz) of the persistent array `pixels` (zeros if None), sized
    max(size_x, size_z) * size_y. Returns (pixels, columns, rows)."""
    x0 = min(p[3] for p in parts); y0 = min(p[4] for p in parts); z0 = min(p[5] for p in parts)
    x1 = max(p[3] + p[6] for p in parts); y1 = max(p[4] + p[7] for p in parts)
    z1 = max(p[5] + p[8] for p in parts)
    sx, sy, sz = x1 - x0, y1 - y0, z1 - z0
    if pixels is None:
        pixels = np.zeros(max(sx, sz) * sy, np.uint32)
    for p, fld in zip(parts, fields):
        is_pml, _, _, px, py, pz, w, h, d = p
        lo, hi = (pz, pz + d) if view == 0 else (px, px + w)
        if lo > coord or hi < coord:
            continue
        flat = np.asarray(fld, np.float32).ravel()
        rows, cols = (h, w) if view == 0 else (d, h)
        i = np.arange(rows)[:, None]
        j = np.arange(cols)[None, :]
        idx = coord * h * w + i * w + j if view == 0 else i * h * w + j * w + coord
        ok = (idx >= 0) & (idx < flat.size)
        if is_pml:
            ok &= (0 <= coord) & (coord < (d if view == 0 else w))
        val = np.where(ok, flat[np.clip(idx, 0, flat.size - 1)], np.float32(0)).astype(np.float32)
        if view == 0:
            pos = (py - y0 + i) * sx + (px - x0 + j)
        else:
            pos = (pz - z0 + i) * sy + (py - y0 + j)
        pixels[pos.ravel()] = colour(val).ravel()
    return pixels, (sx, sy) if view == 0 else (sy, sz)
