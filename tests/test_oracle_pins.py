"""Pins of the CPU oracle against mathematics the reference's method fixes.

The reference cannot be built in this image (it needs Vulkan headers and the
FFTW and SDL libraries), so every oracle function is pinned here by an
independent route: a library routine (scipy), a closed form, an invariant or a
hand-derived expectation. Each test names the reference code it pins.
"""
import numpy as np
import pytest
import scipy.fft
import scipy.ndimage

import ard_oracle as O
import ard_scenes as S

C6 = np.array([2, -27, 270, -490, 270, -27, 2], dtype=np.float64)
C4 = np.array([1, -8, 0, 8, -1], dtype=np.float64)
DAMP = float(np.float32(0.999))  # the reference's literal 0.999f (ARD/dct_partition.cpp:68)


# --- ARD/dct_volume.cpp:38-68: scaled FFTW REDFT10 / REDFT01 ----------------

@pytest.mark.parametrize("shape", [(3, 4, 5), (6, 2, 7), (1, 1, 1), (8, 8, 8)])
def test_dct_matches_scipy(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape)
    n = x.size
    f = O.dct3d(x, inverse=False)
    ref = scipy.fft.dctn(x, type=2, norm=None) / (2 * np.sqrt(2 * n))
    np.testing.assert_allclose(f, ref, rtol=1e-12, atol=1e-12)
    b = O.dct3d(x, inverse=True)
    ref = scipy.fft.dctn(x, type=3, norm=None) / np.sqrt(2 * n)
    np.testing.assert_allclose(b, ref, rtol=1e-12, atol=1e-12)


def test_dct_roundtrip_gain_is_two():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 6, 7))
    np.testing.assert_allclose(O.dct3d(O.dct3d(x), inverse=True), 2 * x, rtol=1e-12, atol=1e-12)


def test_dct_bruteforce_tiny():
    rng = np.random.default_rng(2)
    d, h, w = 2, 3, 2
    x = rng.standard_normal((d, h, w))
    out = np.zeros_like(x)
    for kz in range(d):
        for ky in range(h):
            for kx in range(w):
                s = 0.0
                for z in range(d):
                    for y in range(h):
                        for xx in range(w):
                            s += 8 * x[z, y, xx] * np.cos(np.pi * kx * (xx + .5) / w) * \
                                np.cos(np.pi * ky * (y + .5) / h) * np.cos(np.pi * kz * (z + .5) / d)
                out[kz, ky, kx] = s / (2 * np.sqrt(2 * x.size))
    np.testing.assert_allclose(O.dct3d(x), out, rtol=1e-12, atol=1e-12)


def test_dct_fp32_close_to_fp64():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((6, 5, 4)).astype(np.float32).astype(np.float64)
    a = O.dct3d(x, double=False)
    b = O.dct3d(x, double=True)
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-6


# --- ARD/gaussian_source.cpp:16-20 -------------------------------------------

def test_gaussian_known_values():
    g = O.gaussian(0.1, 1 / 8000, 343.5, 40, double=False)
    assert abs(g[14] - 9.99861e8) / 9.99861e8 < 2e-6
    assert abs(g[10] - 4.10656e7) / 4.10656e7 < 2e-6
    # closed form in double
    t = np.arange(40)
    C = 343.5 / 8000 / 0.1
    ref = 1e9 * np.exp(-(np.pi * (2 * C * t / 6 - 2)) ** 2)
    np.testing.assert_allclose(g, ref, rtol=2e-5, atol=1e-3)
    # peak where 2*C*t/6 == 2
    assert abs(np.argmax(O.gaussian(0.1, 1 / 8000, 343.5, 40, double=True)) - 6 / C) <= 0.5


# --- ARD/dct_partition.cpp:20-36 mode tables (1-based wavenumbers) ------------

def test_mode_tables_closed_form():
    w, h, d, dh, dt, c0 = 4, 3, 5, 0.1, 1 / 8000, 343.0
    w2, cwt = O.mode_tables(dh, dt, c0, w, h, d, double=True)
    k = np.arange(1, w + 1)[None, None, :]
    j = np.arange(1, h + 1)[None, :, None]
    i = np.arange(1, d + 1)[:, None, None]
    om = c0 * np.pi * np.sqrt((i / (d * dh)) ** 2 + (j / (h * dh)) ** 2 + (k / (w * dh)) ** 2)
    np.testing.assert_allclose(w2, om ** 2, rtol=1e-12)
    np.testing.assert_allclose(cwt, np.cos(om * dt), rtol=1e-12)
    w2f, cwtf = O.mode_tables(dh, dt, c0, w, h, d, double=False)
    np.testing.assert_allclose(w2f, om ** 2, rtol=1e-6)
    np.testing.assert_allclose(cwtf, np.cos(om * dt), rtol=1e-6, atol=5e-7)


# --- ARD/dct_partition.cpp:66-70 modal recurrence ------------------------------

def _single_box(double=True, absorption=0.0, w=8, h=6, d=5, dh=0.1):
    o = O.OracleSim(dh, 1 / 8000, 343.0, 5, absorption, double=double)
    o.add_dct(0, 0, 0, w, h, d)
    o.finalize()
    return o


def test_free_mode_decays_as_sqrt_0999():
    """No forcing (absorption 0): each mode obeys M+ = .999(2cM - M-), |lambda|^2 = .999."""
    o = _single_box()
    w2, cwt = O.mode_tables(0.1, 1 / 8000, 343.0, 8, 6, 5, double=True)
    M0 = np.zeros((5, 6, 8))
    M0[1, 2, 3] = 1.0
    M0[0, 0, 0] = 0.5
    o.set_field(0, 2, M0)  # current modes
    o.set_field(0, 3, np.zeros_like(M0))  # previous modes
    n = 25
    o.step(n)
    got = o.field(0, 2)
    # closed form: roots of l^2 - 2*.999*c*l + .999 = 0
    for idx, a0 in (((1, 2, 3), 1.0), ((0, 0, 0), 0.5)):
        c = cwt[idx]
        r = np.sqrt(DAMP)
        th = np.arccos(DAMP * c / r)
        # M_0 = a0, M_{-1} = 0 -> M_n = a0 r^n (cos n th + A sin n th), M_1 = .999*2c*a0
        A = (DAMP * 2 * c / r - np.cos(th)) / np.sin(th)
        ref = a0 * r ** n * (np.cos(n * th) + A * np.sin(n * th))
        assert abs(got[idx] - ref) < 1e-12
    mask = np.ones_like(M0, bool)
    mask[1, 2, 3] = mask[0, 0, 0] = False
    assert np.all(got[mask] == 0)
    # pressure is the scaled DCT-III of the current modes
    np.testing.assert_allclose(o.field(0, 0), scipy.fft.dctn(got, type=3) / np.sqrt(2 * got.size), atol=1e-13)


def test_constant_force_fixed_point():
    """Persistent interior force F (never cleared: ARD/dct_partition.cpp:48-73) has the
    fixed point M* = .999*2Fh(1-c)/w^2 / (1 - .999(2c-1)); starting there it stays."""
    # F sits at a cell no boundary rewrites (at least 3 cells from every face)
    o = _single_box(w=8, h=8, d=8)
    w2, cwt = O.mode_tables(0.1, 1 / 8000, 343.0, 8, 8, 8, double=True)
    F = np.zeros((8, 8, 8))
    F[3, 4, 3] = 7.0
    Fh = scipy.fft.dctn(F, type=2) / (2 * np.sqrt(2 * F.size))
    Ms = DAMP * 2 * Fh * (1 - cwt) / w2 / (1 - DAMP * (2 * cwt - 1))
    o.set_field(0, 1, F)
    o.set_field(0, 2, Ms)
    o.set_field(0, 3, Ms)
    o.step(3)
    np.testing.assert_allclose(o.field(0, 2), Ms, rtol=1e-10, atol=1e-18)
    np.testing.assert_allclose(o.field(0, 1), F)  # the force persists
    np.testing.assert_allclose(o.field(0, 0), scipy.fft.dctn(Ms, type=3) / np.sqrt(2 * F.size), atol=1e-15)


# --- ARD/pml_partition.cpp:28-92 damping profile -------------------------------

@pytest.mark.parametrize("ptype", range(6))
def test_pml_profile_closed_form(ptype):
    dh, c0, L = 0.1, 343.0, 5
    shape = {0: (5, 4, 3), 1: (5, 4, 3), 2: (4, 5, 3), 3: (4, 5, 3), 4: (4, 3, 5), 5: (4, 3, 5)}[ptype]
    w, h, d = shape
    prof = O.pml_profile(dh, c0, L, ptype, w, h, d, double=True)
    T = L * dh
    z0 = c0 / T * np.log10(1 / np.float64(np.float32(1e-5)))

    def f(q):
        x = q * dh
        return z0 * (x / T - np.sin(2 * np.pi * x / T) / (2 * np.pi))

    axis = {0: 0, 1: 0, 2: 1, 3: 1, 4: 2, 5: 2}[ptype]
    n = shape[axis]
    q = np.arange(n)
    # increasing outward: RIGHT/BOTTOM/BACK grow with the index, LEFT/TOP/FRONT shrink
    depth = q + 1 if ptype in (1, 3, 5) else n - q
    expect = f(depth)
    comp = prof[axis]
    line = {0: comp[0, 0, :], 1: comp[0, :, 0], 2: comp[:, 0, 0]}[axis]
    np.testing.assert_allclose(line, expect, rtol=1e-12)
    for other in range(3):
        if other != axis:
            assert np.all(prof[other] == 0)
    assert f(0) == 0 and f(L) > f(1) > 0


# --- ARD/pml_partition.cpp:113-200 Grote-Sim update --------------------------

def _pml_sim(scale=1.0):
    o = O.OracleSim(0.1, 1 / 8000, 343.0, 5, 0.0, double=True)
    o.set_pml_scale([scale] * 6)
    o.add_dct(0, 0, 0, 12, 12, 12)
    o.finalize()
    return o


def _zero_pad_deriv(a, coefs, axis):
    # correlate along an axis with zero padding (the sentinel of GetIndex)
    return scipy.ndimage.correlate1d(a, coefs, axis=axis, mode="constant", cval=0.0)


def test_pml_undamped_is_leapfrog_with_6th_order_laplacian():
    """With zeta = 0 the update is p+ = 2p - p- + dt^2 (c^2 L6 p + f); L6 with zero padding."""
    o = _pml_sim(scale=0.0)
    parts = o.partitions()
    front = [i for i, p in enumerate(parts) if p[0] == 1 and p[1] == O.P_FRONT][0]
    w, h, d = parts[front][6:9]
    rng = np.random.default_rng(4)
    p = rng.standard_normal((d, h, w))
    pm = rng.standard_normal((d, h, w))
    f = rng.standard_normal((d, h, w))
    o.set_field(front, 0, p)
    o.set_field(front, 7, pm)
    o.set_field(front, 1, f)
    o.step(1)
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    lap = sum(_zero_pad_deriv(p, C6, ax) for ax in range(3)) / (180 * dh * dh)
    ref = 2 * p - pm + dt * dt * (c0 * c0 * lap + f)
    np.testing.assert_allclose(o.field(front, 0), ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(o.field(front, 7), p)
    # the force was cleared by the update and rewritten by the (absorption 0) boundary
    assert np.all(o.field(front, 1) == 0)


def test_pml_laplacian_symbol_plane_wave():
    """Plane wave cos(k x): interior cells see the stencil symbol (closed form)."""
    o = _pml_sim(scale=0.0)
    parts = o.partitions()
    front = [i for i, p in enumerate(parts) if p[0] == 1 and p[1] == O.P_FRONT][0]
    w, h, d = parts[front][6:9]
    th = 0.7
    x = np.arange(w)
    p = np.broadcast_to(np.cos(th * x)[None, None, :], (d, h, w)).copy()
    o.set_field(front, 0, p)
    o.set_field(front, 7, p)
    o.step(1)
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    sym = (-490 + 540 * np.cos(th) - 54 * np.cos(2 * th) + 4 * np.cos(3 * th)) / (180 * dh * dh)
    # the slab is 5 cells thick in z: there the constant profile meets the zero sentinel
    thin = _zero_pad_deriv(np.ones((d, h, w)), C6, 0)[2, 5, 3:-3] / (180 * dh * dh)
    got = o.field(front, 0)[2, 5, 3:-3]
    np.testing.assert_allclose(got, (1 + dt * dt * c0 * c0 * (sym + thin)) * p[2, 5, 3:-3], rtol=1e-12)


def test_pml_damping_terms_grote_sim():
    """Uniform p, zero phi: interior update reduces to the ODE part of Grote-Sim
    p_tt + zeta p_t = 0 (one nonzero zeta) and phi_t = -zeta phi + (-zeta) dp/dx."""
    o = _pml_sim(scale=1.0)
    parts = o.partitions()
    right = [i for i, p in enumerate(parts) if p[0] == 1 and p[1] == O.P_RIGHT][0]
    w, h, d = parts[right][6:9]
    prof = O.pml_profile(0.1, 343.0, 5, O.P_RIGHT, w, h, d, double=True)
    zx = prof[0]
    dt = 1 / 8000
    p = np.full((d, h, w), 2.0)
    pm = np.full((d, h, w), 1.5)
    o.set_field(right, 0, p)
    o.set_field(right, 7, pm)
    o.step(1)
    got = o.field(right, 0)
    # cells whose 7-point stencils stay inside the slab in y and z, and x fixed at the
    # centre column (x stencil still leaves the 5-wide slab; compare with zero padding)
    dh, c0 = 0.1, 343.0
    lap = sum(_zero_pad_deriv(p, C6, ax) for ax in range(3)) / (180 * dh * dh)
    ref = 2 * p - pm + dt * dt * (c0 * c0 * lap - zx * (p - pm) / dt)
    np.testing.assert_allclose(got, ref, rtol=1e-12)
    # phi_x update: phi+ = phi - dt zx phi + dt (0 + 0 - zx) dp/dx, phi = 0
    dpx = _zero_pad_deriv(p, C4, 2) / (12 * dh)
    np.testing.assert_allclose(o.field(right, 4), dt * (-zx) * dpx, rtol=1e-12, atol=1e-15)
    # phi_y: dt*(zx + 0 - 0)*dp/dy
    dpy = _zero_pad_deriv(p, C4, 1) / (12 * dh)
    np.testing.assert_allclose(o.field(right, 5), dt * zx * dpy, rtol=1e-12, atol=1e-15)


# --- ARD/boundary.cpp:21-173 interface forcing --------------------------------

def _neumann_minus_global(pl, pr, axis):
    """c^2/(180 h^2) * (global L6 - mirrored L6) at the 3 cells on each side.

    The modal solver of each box has even (Neumann) symmetry about its faces, so
    its implicit 6th-order operator at the face sees mirrored copies of its own
    cells; the forcing restores the neighbour's true values."""
    n_l = pl.shape[axis]
    full = np.concatenate([pl, pr], axis=axis)
    g = scipy.ndimage.correlate1d(full, C6, axis=axis, mode="constant")
    ml = scipy.ndimage.correlate1d(pl, C6, axis=axis, mode="reflect")  # reflect = half-sample mirror
    mr = scipy.ndimage.correlate1d(pr, C6, axis=axis, mode="reflect")
    take = lambda a, idx: np.take(a, idx, axis=axis)
    left = take(g, range(n_l - 3, n_l)) - take(ml, range(n_l - 3, n_l))
    right = take(g, range(n_l, n_l + 3)) - take(mr, range(0, 3))
    return left, right


def test_interface_forcing_is_global_minus_neumann_operator():
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    o = O.OracleSim(dh, dt, c0, 5, 0.0, double=True)  # PML boundaries write zeros
    W, H, D = 10, 11, 12
    o.add_dct(0, 0, 0, W, H, D)
    o.add_dct(W, 0, 0, W, H, D)
    o.finalize()
    rng = np.random.default_rng(5)
    # after one step with zero force and zero previous modes, p = IDCT(.999*2c*M)
    M = [rng.standard_normal((D, H, W)) for _ in range(2)]
    for i in range(2):
        o.set_field(i, 2, M[i])
        o.set_field(i, 3, np.zeros((D, H, W)))
    o.step(1)
    pl, pr = o.field(0, 0), o.field(1, 0)
    _, cwt = O.mode_tables(dh, dt, c0, W, H, D, double=True)
    for i, pp in enumerate((pl, pr)):
        np.testing.assert_allclose(pp, scipy.fft.dctn(DAMP * 2 * cwt * M[i], type=3) / np.sqrt(2 * W * H * D),
                                   atol=1e-12)
    left, right = _neumann_minus_global(pl, pr, axis=2)
    k = c0 * c0 / (180 * dh * dh)
    fl, fr = o.field(0, 1), o.field(1, 1)
    # cells not rewritten later by the PML boundaries (y, z shells are zeroed by them)
    sl = (slice(3, D - 3), slice(3, H - 3))
    np.testing.assert_allclose(fl[sl][:, :, W - 3:], k * left[sl], rtol=1e-10, atol=1e-6)
    np.testing.assert_allclose(fr[sl][:, :, :3], k * right[sl], rtol=1e-10, atol=1e-6)
    # last writer: the PML boundaries (absorption 0) come after the DCT-DCT one
    assert np.all(fl[:3, :, W - 3:] == 0) and np.all(fl[:, :3, W - 3:] == 0)
    # forcing rows sum to zero: a constant field gives no forcing
    assert np.allclose(np.sum(np.array([[0, 0, -2, 2, 0, 0], [0, -2, 27, -27, 2, 0],
                                        [-2, 27, -270, 270, -27, 2]]), axis=1), 0)


def test_pml_side_forcing_uses_neighbour_only():
    """include_self_terms_ = false for PML (ARD/pml_partition.cpp:20): its forcing is
    absorption * c^2/(180h^2) * (rows of the matrix applied to the DCT cells only)."""
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    o = O.OracleSim(dh, dt, c0, 5, 0.7, double=True)
    W, H, D = 10, 10, 10
    o.add_dct(0, 0, 0, W, H, D)
    o.finalize()
    rng = np.random.default_rng(6)
    M = rng.standard_normal((D, H, W))
    o.set_field(0, 2, M)
    o.set_field(0, 3, np.zeros_like(M))
    o.step(1)
    p = o.field(0, 0)
    parts = o.partitions()
    right = [i for i, q in enumerate(parts) if q[0] == 1 and q[1] == O.P_RIGHT][0]
    f = o.field(right, 1)  # PML force written by the boundary after the update
    k = 0.7 * c0 * c0 / (180 * dh * dh)
    rows = np.array([[2, -27, 270], [0, 2, -27], [0, 0, 2]], dtype=float)  # m = 0,1,2; n = 0..2 (left cells)
    for m in range(3):
        expect = k * sum(rows[m, n] * p[:, :, W - 3 + n] for n in range(3))
        np.testing.assert_allclose(f[:, :, m], expect, rtol=1e-12, atol=1e-9)
    assert np.all(f[:, :, 3:] == 0)


# --- ARD/simulation.cpp:14-266, ARD/boundary.cpp:289-334 topology -------------

def test_topology_two_rooms_hand_derived():
    o = O.from_scene(S.two_rooms())
    parts = o.partitions()
    expect = [
        (0, -1, -1, 0, 0, 0, 30, 30, 30), (0, -1, -1, 30, 0, 0, 30, 30, 30),
        (1, 0, 0, -5, 0, 0, 5, 30, 30), (1, 2, 0, 0, -5, 0, 30, 5, 30), (1, 3, 0, 0, 30, 0, 30, 5, 30),
        (1, 4, 0, 0, 0, -5, 30, 30, 5), (1, 5, 0, 0, 0, 30, 30, 30, 5),
        (1, 1, 1, 60, 0, 0, 5, 30, 30), (1, 2, 1, 30, -5, 0, 30, 5, 30), (1, 3, 1, 30, 30, 0, 30, 5, 30),
        (1, 4, 1, 30, 0, -5, 30, 30, 5), (1, 5, 1, 30, 0, 30, 30, 30, 5)]
    assert parts == expect
    b = o.boundaries()
    assert b[0] == (0, 0, 1, 27, 33, 0, 30, 0, 30)
    assert b[1] == (0, 2, 0, -3, 3, 0, 30, 0, 30)
    assert b[5] == (2, 6, 0, 0, 30, 0, 30, 27, 33)
    assert len(b) == 11
    assert sum(p[6] * p[7] * p[8] for p in parts if p[0]) == 45000


def test_topology_scene3_counts_and_first_box():
    o = O.from_scene(S.scene3())
    parts = o.partitions()
    assert sum(1 for p in parts if p[0] == 0) == 5
    assert sum(1 for p in parts if p[0] == 1) == 26
    assert len(o.boundaries()) == 30
    assert sum(p[6] * p[7] * p[8] for p in parts if p[0] == 0) == 240000
    assert sum(p[6] * p[7] * p[8] for p in parts if p[0] == 1) == 140000
    pml0 = [p[1:] for p in parts if p[0] == 1 and p[2] == 0]
    assert pml0 == [
        (0, 0, -5, 0, 0, 5, 60, 40),      # left: whole height free
        (1, 0, 60, 40, 0, 5, 20, 40),     # right: y 40..60 free (box 3 covers 0..40)
        (2, 0, 0, -5, 0, 60, 5, 40),      # top
        (3, 0, 40, 60, 0, 20, 5, 40),     # bottom: x 40..60 free (box 2 covers 0..40)
        (4, 0, 0, 0, -5, 60, 60, 5),
        (5, 0, 0, 0, 40, 60, 60, 5)]


def test_topology_run_detection_quirks():
    """Run-length rules of the PML loop (ARD/simulation.cpp:62-98)."""
    def right_pmls(boxes):
        o = O.OracleSim(0.1, 1 / 8000, 343.0, 5, 1.0)
        for b in boxes:
            o.add_dct(*b)
        o.finalize()
        return [p[4:8] for p in o.partitions() if p[0] == 1 and p[1] == O.P_RIGHT and p[2] == 0]
    A = (0, 0, 0, 10, 10, 10)
    # free only at the last row -> no PML there
    assert right_pmls([A, (10, 0, 0, 10, 9, 10)]) == []
    # free only at row 0 -> a PML of height 1
    assert right_pmls([A, (10, 1, 0, 10, 9, 10)]) == [(0, 0, 5, 1)]
    # free only at row 8 of 10 -> the run ends on the last row without the decrement: height 2
    assert right_pmls([A, (10, 0, 0, 10, 8, 10), (10, 9, 0, 10, 1, 10)]) == [(8, 0, 5, 2)]
    # free rows 3..5 -> height 3
    assert right_pmls([A, (10, 0, 0, 10, 3, 10), (10, 6, 0, 10, 4, 10)]) == [(3, 0, 5, 3)]


def test_recorder_reference_addressing():
    """Recorder samples index the owner's array with global coordinates."""
    sc = S.two_rooms()
    o = O.from_scene(sc)
    assert o.recorder_partition(0) == 1
    o.step(30)
    tr = o.step(1)[0, 0]
    p1 = o.field(1, 0)
    flat = 15 * 900 + 20 * 30 + 40
    assert tr == p1.ravel()[flat] == p1[15, 21, 10]


# --- coupled solver invariants -----------------------------------------------

def test_linearity_in_source_amplitude():
    sc = S.two_rooms()
    g = O.gaussian(sc.h, sc.dt, sc.c0, 120, double=True)
    traces = []
    for a in (1.0, 2.5):
        sc2 = S.two_rooms()
        sc2.sources = [S.Source(10, 10, 10, wave=a * g)]
        traces.append(O.from_scene(sc2, double=True).step(120))
    np.testing.assert_allclose(traces[1], 2.5 * traces[0], rtol=1e-9, atol=1e-12 * np.abs(traces[0]).max())


def test_mirror_symmetry_two_rooms():
    """x -> 59 - x maps the scene to itself with the boxes swapped."""
    def run(src, rec):
        sc = S.two_rooms()
        sc.sources = [S.Source(*src)]
        sc.recorders = [rec]
        sc.recorder_mode = S.REC_LOCAL
        return O.from_scene(sc, double=True).step(150)[:, 0]
    a = run((10, 10, 10), (40, 20, 15))
    b = run((49, 10, 10), (19, 20, 15))
    assert np.abs(a).max() > 0
    np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10 * np.abs(a).max())


def test_direct_arrival_time():
    """The pulse reaches a recorder r metres away r/c seconds after its peak."""
    sc = S.two_rooms()
    sc.recorders = [(40, 10, 10)]
    sc.recorder_mode = S.REC_LOCAL
    tr = O.from_scene(sc).step(260)[:, 0]
    C = sc.c0 * sc.dt / sc.h
    t_peak = 6 / C
    r = 30 * sc.h
    expect = t_peak + r / (sc.c0 * sc.dt)
    # first arrival (before wall reflections); the 1-based mode numbering of the
    # reference (ARD/dct_partition.cpp:24-31) speeds up the lowest modes slightly
    assert abs(np.argmax(np.abs(tr[:130])) - expect) < 6


# --- ARD/pml_partition.cpp:113-200 with nonzero phi (all terms at once) ------

def _pml_slab(ptype, layers=5, scale=1.0, absorption=0.0, box=(12, 12, 12)):
    o = O.OracleSim(0.1, 1 / 8000, 343.0, layers, absorption, double=True)
    o.set_pml_scale([scale] * 6)
    o.add_dct(0, 0, 0, *box)
    o.finalize()
    parts = o.partitions()
    i = [k for k, p in enumerate(parts) if p[0] == 1 and p[1] == ptype][0]
    return o, i, parts[i][6:9]


@pytest.mark.parametrize("ptype", range(6))
def test_pml_update_random_fields_all_terms(ptype):
    """Random p, p-, phi_x, phi_y, phi_z and f on a damped slab of each face type:
    p+ = 2p - p- + dt^2 (c^2 L6 p - sum(zeta) (p - p-)/dt - (zx zy + zy zz + zx zz) p
    + D4x phi_x + D4y phi_y + D4z phi_z + f) and
    phi_a+ = phi_a - dt zeta_a phi_a + dt (sum of the other zetas - zeta_a) D4a p
    (ARD/pml_partition.cpp:158-186), every stencil applied by scipy with zero
    padding (the GetIndex sentinel, ARD/pml_partition.cpp:6-15)."""
    o, i, (w, h, d) = _pml_slab(ptype)
    rng = np.random.default_rng(40 + ptype)
    p, pm, phx, phy, phz, f = (rng.standard_normal((d, h, w)) for _ in range(6))
    for which, a in ((0, p), (7, pm), (4, phx), (5, phy), (6, phz), (1, f)):
        o.set_field(i, which, a)
    o.step(1)
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    zx, zy, zz = O.pml_profile(dh, c0, 5, ptype, w, h, d, double=True)
    D4 = [_zero_pad_deriv(a, C4, ax) / (12 * dh) for a, ax in ((phx, 2), (phy, 1), (phz, 0))]
    lap = sum(_zero_pad_deriv(p, C6, ax) for ax in range(3)) / (180 * dh * dh)
    ref = 2 * p - pm + dt * dt * (c0 * c0 * lap - (zx + zy + zz) * (p - pm) / dt
                                  - (zx * zy + zy * zz + zx * zz) * p + D4[0] + D4[1] + D4[2] + f)
    np.testing.assert_allclose(o.field(i, 0), ref, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(o.field(i, 7), p)
    dpx, dpy, dpz = (_zero_pad_deriv(p, C4, ax) / (12 * dh) for ax in (2, 1, 0))
    np.testing.assert_allclose(o.field(i, 4), phx - dt * zx * phx + dt * (zy + zz - zx) * dpx, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(o.field(i, 5), phy - dt * zy * phy + dt * (zx + zz - zy) * dpy, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(o.field(i, 6), phz - dt * zz * phz + dt * (zx + zy - zz) * dpz, rtol=1e-11, atol=1e-13)


def test_pml_update_polynomial_fields_analytic_derivatives():
    """Independent of any stencil convention: on polynomial fields the 4th-order
    first differences and the 6th-order second differences are exact, so at
    cells whose stencils stay inside the slab the update must equal the
    Grote-Sim right-hand side with analytic derivatives. An 8-layer P_RIGHT
    slab (every zeta nonzero only along x) leaves x = 3, 4 interior."""
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    o, i, (w, h, d) = _pml_slab(O.P_RIGHT, layers=8)
    z, y, x = np.meshgrid(np.arange(d) * dh, np.arange(h) * dh, np.arange(w) * dh, indexing="ij")
    # p: quadratic + cross terms; phi: quadratics; derivatives by hand
    p = 1.0 + 0.3 * x - 0.7 * y + 0.2 * z + 0.9 * x * x - 0.4 * y * y + 0.6 * z * z + 0.5 * x * y - 0.8 * y * z
    pm = 0.4 - 0.1 * x + 0.25 * y * z
    phx = 0.2 + 1.5 * x * x - 0.3 * y          # d/dx = 3 x
    phy = -0.6 * y * y + 0.7 * x * z           # d/dy = -1.2 y
    phz = 0.9 * z - 0.35 * z * z + x * y       # d/dz = 0.9 - 0.7 z
    f = 0.05 * x - 0.02
    for which, a in ((0, p), (7, pm), (4, phx), (5, phy), (6, phz), (1, f)):
        o.set_field(i, which, a)
    o.step(1)
    zx, zy, zz = O.pml_profile(dh, c0, 8, O.P_RIGHT, w, h, d, double=True)
    lap_p = 2 * 0.9 + 2 * (-0.4) + 2 * 0.6
    div_phi = 3 * x + (-1.2 * y) + (0.9 - 0.7 * z)
    ref = 2 * p - pm + dt * dt * (c0 * c0 * lap_p - (zx + zy + zz) * (p - pm) / dt
                                  - (zx * zy + zy * zz + zx * zz) * p + div_phi + f)
    dpx = 0.3 + 1.8 * x + 0.5 * y
    dpy = -0.7 - 0.8 * y + 0.5 * x - 0.8 * z
    dpz = 0.2 + 1.2 * z - 0.8 * y
    inner = (slice(3, d - 3), slice(3, h - 3), slice(3, w - 3))
    assert w == 8 and p[inner].size > 0
    np.testing.assert_allclose(o.field(i, 0)[inner], ref[inner], rtol=1e-10)
    np.testing.assert_allclose(o.field(i, 4)[inner], (phx - dt * zx * phx + dt * (zy + zz - zx) * dpx)[inner],
                               rtol=1e-10)
    np.testing.assert_allclose(o.field(i, 5)[inner], (phy - dt * zy * phy + dt * (zx + zz - zy) * dpy)[inner],
                               rtol=1e-10)
    np.testing.assert_allclose(o.field(i, 6)[inner], (phz - dt * zz * phz + dt * (zx + zy - zz) * dpz)[inner],
                               rtol=1e-10)


# --- ARD/boundary.cpp:78-173: Y and Z branches, PML sides of every face ------

def test_y_interface_forcing_is_global_minus_neumann_operator():
    """Boxes stacked in y (ARD/boundary.cpp:78-125): same identity as the x pin."""
    dh, dt, c0 = 0.1, 1 / 8000, 343.0
    o = O.OracleSim(dh, dt, c0, 5, 0.0, double=True)
    W, H, D = 11, 10, 12
    o.add_dct(0, 0, 0, W, H, D)
    o.add_dct(0, H, 0, W, H, D)
    o.finalize()
    assert o.boundaries()[0] == (1, 0, 1, 0, W, H - 3, H + 3, 0, D)
    rng = np.random.default_rng(7)
    for i in range(2):
        o.set_field(i, 2, rng.standard_normal((D, H, W)))
        o.set_field(i, 3, rng.standard_normal((D, H, W)))
    o.step(1)
    pt, pb = o.field(0, 0), o.field(1, 0)
    top, bot = _neumann_minus_global(pt, pb, axis=1)
    k = c0 * c0 / (180 * dh * dh)
    ft, fb = o.field(0, 1), o.field(1, 1)
    # cells no later (PML) boundary rewrites: x and z away from the shells
    sl = (slice(3, D - 3), slice(None), slice(3, W - 3))
    np.testing.assert_allclose(ft[:, H - 3:, :][sl], k * top[sl], rtol=1e-10, atol=1e-6)
    np.testing.assert_allclose(fb[:, :3, :][sl], k * bot[sl], rtol=1e-10, atol=1e-6)
    # and the PML boundaries, registered later, won the edge cells (absorption 0)
    assert np.all(ft[:, H - 3:, :3] == 0) and np.all(fb[:3, :3, :] == 0)


# (type, axis, PML before the DCT box along the axis?)
_FACE = {O.P_LEFT: (2, True), O.P_RIGHT: (2, False), O.P_TOP: (1, True), O.P_BOTTOM: (1, False),
         O.P_FRONT: (0, True), O.P_BACK: (0, False)}


@pytest.mark.parametrize("ptype", range(6))
def test_pml_boundary_both_sides_every_face(ptype):
    """A box with absorption 0.7 and random fields in it and in its PML slabs.
    After one step, across the face of each slab type (X: ARD/boundary.cpp:30-77,
    Y: 78-125, Z: 126-173):
    * the PML cells carry 0.7 c^2/(180h^2) x the 6th-order operator's part that
      reaches into the box (include_self_terms_ = false, ARD/pml_partition.cpp:20);
    * the box cells carry 0.7 c^2/(180h^2) x (global operator over box + slab
      minus the box's mirrored (Neumann) operator).
    Both computed with scipy correlate1d; the Z branch writes last, so its box
    cells hold its values everywhere on the face; X/Y faces are checked away
    from the z shells (rewritten by the Z boundaries)."""
    dh, dt, c0, ab = 0.1, 1 / 8000, 343.0, 0.7
    o = O.OracleSim(dh, dt, c0, 5, ab, double=True)
    W, H, D = 12, 11, 10
    o.add_dct(0, 0, 0, W, H, D)
    o.finalize()
    parts = o.partitions()
    rng = np.random.default_rng(60 + ptype)
    o.set_field(0, 2, rng.standard_normal((D, H, W)))
    o.set_field(0, 3, rng.standard_normal((D, H, W)))
    for q, info in enumerate(parts):
        if info[0] == 1:
            w, h, d = info[6:9]
            o.set_field(q, 0, rng.standard_normal((d, h, w)))
            o.set_field(q, 7, rng.standard_normal((d, h, w)))
    o.step(1)
    i = [q for q, p in enumerate(parts) if p[0] == 1 and p[1] == ptype][0]
    axis, before = _FACE[ptype]
    pd, pp = o.field(0, 0), o.field(i, 0)
    fd, fp = o.field(0, 1), o.field(i, 1)
    k = ab * c0 * c0 / (180 * dh * dh)
    n_d, n_p = pd.shape[axis], pp.shape[axis]
    take = lambda a, idx: np.take(a, list(idx), axis=axis)
    full = np.concatenate([pp, pd] if before else [pd, pp], axis=axis)
    only_box = np.concatenate([np.zeros_like(pp), pd] if before else [pd, np.zeros_like(pp)], axis=axis)
    g = scipy.ndimage.correlate1d(full, C6, axis=axis, mode="constant")
    gb = scipy.ndimage.correlate1d(only_box, C6, axis=axis, mode="constant")
    mirror = scipy.ndimage.correlate1d(pd, C6, axis=axis, mode="reflect")
    if before:   # slab cells n_p-3..n_p-1, box cells 0..2
        pml_cells, box_full, box_own = range(n_p - 3, n_p), range(n_p, n_p + 3), range(0, 3)
    else:        # box cells n_d-3..n_d-1, slab cells 0..2
        pml_cells, box_full, box_own = range(n_d, n_d + 3), range(n_d - 3, n_d), range(n_d - 3, n_d)
    exp_pml = k * take(gb, pml_cells)
    exp_box = k * (take(g, box_full) - take(mirror, box_own))
    got_pml = take(fp, range(n_p - 3, n_p) if before else range(0, 3))
    got_box = take(fd, box_own)
    np.testing.assert_allclose(got_pml, exp_pml, rtol=1e-10, atol=1e-6)
    if axis == 0:
        np.testing.assert_allclose(got_box, exp_box, rtol=1e-10, atol=1e-6)
    else:
        inner = [slice(3, D - 3), slice(None), slice(None)]
        # along the face the other tangential axis is also away from the other slabs' shells
        other = 1 if axis == 2 else 2
        inner[other] = slice(3, (H if other == 1 else W) - 3)
        np.testing.assert_allclose(got_box[tuple(inner)], exp_box[tuple(inner)], rtol=1e-10, atol=1e-6)
    assert np.abs(exp_pml).max() > 1 and np.abs(exp_box).max() > 1


def test_mirror_symmetry_y_stacked_rooms():
    """y -> 2H-1-y maps two y-stacked boxes onto each other (Y-boundary branch,
    PML top/bottom slabs and their last-writer order all map consistently)."""
    def run(src, rec):
        sc = S.Scene("y_pair", 0.1, 1 / 8000, 343.0, [(0, 0, 0, 24, 20, 22), (0, 20, 0, 24, 20, 22)],
                     [S.Source(*src)], [rec], 150, recorder_mode=S.REC_LOCAL)
        return O.from_scene(sc, double=True).step(150)[:, 0]
    a = run((9, 6, 10), (12, 28, 11))
    b = run((9, 33, 10), (12, 11, 11))
    assert np.abs(a).max() > 0
    np.testing.assert_allclose(a, b, rtol=1e-8, atol=1e-10 * np.abs(a).max())


# --- the differentiation target's damping (Grote-Sim PML near its limit) ------

def test_pml_damping_scale_1p2_diverges_scale_1_does_not():
    """The reference's PML update with the damping profile scaled by 1.2 grows
    without bound on the 1/4-scale differentiation room, while scale 1 stays
    bounded: the reason the gradient workload's target uses damping / 1.2."""
    sc = S.diff_room(scale=4, steps=1200)
    peaks = {}
    for scale in (1.0, 1.2):
        o = O.from_scene(sc, double=True, pml_scale=[scale] * 6)
        tr = o.step(1200)
        early = np.abs(tr[:400]).max()
        late = np.abs(tr[800:]).max() if np.all(np.isfinite(tr[800:])) else np.inf
        peaks[scale] = (early, late)
    e1, l1 = peaks[1.0]
    e2, l2 = peaks[1.2]
    assert np.isfinite(l1) and l1 < 10 * e1
    assert not np.isfinite(l2) or l2 > 1e6 * e2


# --- the O(n log n) transforms (REDFT10/REDFT01 by FFT) ------------------------

@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 1, 1), (2, 1, 3), (13, 37, 40), (7, 11, 49), (20, 40, 60)])
def test_fft_dct_matches_scipy_and_dense(shape):
    rng = np.random.default_rng(8)
    x = rng.standard_normal(shape)
    n = x.size
    for inv, t, sc in ((False, 2, 2 * np.sqrt(2 * n)), (True, 3, np.sqrt(2 * n))):
        f = O.dct3d(x, inverse=inv, fft=True)
        np.testing.assert_allclose(f, scipy.fft.dctn(x, type=t, norm=None) / sc, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(f, O.dct3d(x, inverse=inv, fft=False), rtol=1e-11, atol=1e-12)


def test_fft_oracle_reproduces_dense_oracle_goldens():
    """The two-room golden traces were written by the dense-sum oracle; the FFT
    path reproduces them to rounding (FP64) and to FP32 rounding (FP32)."""
    import os
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sc = S.two_rooms()
    for double, tol in ((True, 1e-11), (False, 1e-5)):
        ref = np.load(os.path.join(gdir, "two_rooms_%s.npy" % ("fp64" if double else "fp32")))
        got = O.from_scene(sc, double=double, fft=True).step(sc.steps)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


def test_hall_goldens_fp32_within_1e4_of_fp64():
    """The hall's two golden traces (2000 steps, 16 recorders, both by the FFT
    path): the FP32 oracle, in the reference's operation order, lies within
    the metric's 1e-4 of the FP64 oracle; the traces carry the pulse (the
    direct sound reaches every recorder within the window)."""
    import os
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    r64 = np.load(os.path.join(gdir, "hall_fp64.npy"))
    r32 = np.load(os.path.join(gdir, "hall_fp32.npy")).astype(np.float64)
    sc = S.hall()
    assert r64.shape == r32.shape == (sc.steps, len(sc.recorders))
    assert np.linalg.norm(r32 - r64) / np.linalg.norm(r64) < 1e-4
    assert np.all(np.abs(r64).max(axis=0) > 1e-3 * np.abs(r64).max())
