"""The discrete adjoint on the differentiation workload at its configured size:
a 10 x 8 x 4 m room in 3 partitions at h = 0.05 m (200 x 160 x 80 cells, 2.56M
DCT cells), one source of seeded band-limited noise (4000 samples), 8
recorders, 4000 steps, checkpoints every ceil(sqrt(T)) = 64 steps. Loss =
squared L2 between the traces and a target made by the same solver with the
waveform x 1.05 and every face's PML damping divided by 1.2 (damping x 1.2
makes the reference's own PML update diverge:
tests/test_oracle_pins.py::test_pml_damping_scale_1p2_diverges_scale_1_does_not).

Checked: the gradients are finite and nonzero; checkpoint strides sqrt(T) and
2 sqrt(T) give the same gradients (up to the order of atomic additions); two
damping scales and two waveform samples agree with central finite
differences within 1e-3 relative. Time and checkpoint memory are printed."""
import time

import numpy as np
import pytest

import ard_scenes as S
import ard_sim as A

pytestmark = pytest.mark.gpu

STEPS = 4000
TOL = 1e-3


def traces(sc, wave, theta):
    s2 = S.Scene(**{**sc.__dict__})
    s2.sources = [S.Source(sc.sources[0].x, sc.sources[0].y, sc.sources[0].z, wave=np.asarray(wave, np.float32))]
    sim = A.Simulation(s2, form=A.FORM_PLAIN, pml_scale=np.asarray(theta, np.float32))
    sim.update(STEPS)
    return sim, sim.recorders(0, STEPS).astype(np.float64)


def loss(sc, w, th, target):
    sim, tr = traces(sc, w, th)
    sim.close()
    return float(np.sum((tr - target) ** 2))


@pytest.fixture(scope="module")
def full():
    sc = S.diff_room(scale=1, steps=STEPS)
    assert sc.dct_cells() == 200 * 160 * 80 and len(sc.boxes) == 3 and len(sc.recorders) == 8
    w0 = sc.sources[0].wave.astype(np.float32).astype(np.float64)
    print("case differentiation room, full size, %d steps" % STEPS, flush=True)
    s, target = traces(sc, w0 * 1.05, np.full(6, 1 / 1.2))
    s.close()
    sim, tr = traces(sc, w0, np.ones(6))
    g = np.ascontiguousarray(2 * (tr - target), dtype=np.float32)
    t0 = time.perf_counter()
    dw, ds = sim.adjoint(g)  # stride ceil(sqrt(T)) = 64
    t1 = time.perf_counter()
    dw2, ds2 = sim.adjoint(g, checkpoint_stride=128)
    # checkpoint memory of the call: (T/K + 1) states of all DCT and PML partitions
    cells = sum(p[6] * p[7] * p[8] for p in sim.partitions())
    print("adjoint of %d steps (%d cells) in %.2f s; forward step state ~%.0f MB, %d checkpoints" % (
        STEPS, cells, t1 - t0, cells * 4 * 4 / 2 ** 20, STEPS // 64 + 1), flush=True)
    sim.close()
    return sc, w0, target, tr, dw, ds, dw2, ds2


def test_full_size_gradients_finite_and_stride_independent(full):
    sc, w0, target, tr, dw, ds, dw2, ds2 = full
    assert np.all(np.isfinite(tr)) and np.abs(tr).max() > 0
    assert np.all(np.isfinite(dw)) and np.all(np.isfinite(ds))
    assert np.abs(dw).max() > 0 and np.all(ds != 0)
    assert np.linalg.norm(dw2 - dw) <= 1e-5 * np.linalg.norm(dw)
    assert np.linalg.norm(ds2 - ds) <= 1e-5 * np.linalg.norm(ds)


def test_full_size_damping_gradients_match_finite_differences(full):
    sc, w0, target, tr, dw, ds, dw2, ds2 = full
    for f in (0, 4):  # an x face and a z face
        eps = 1e-2
        tp, tm = np.ones(6), np.ones(6)
        tp[f] += eps
        tm[f] -= eps
        fd = (loss(sc, w0, tp, target) - loss(sc, w0, tm, target)) / (2 * eps)
        print("face %d adjoint %.6e fd %.6e" % (f, ds[f], fd), flush=True)
        assert abs(ds[f] - fd) <= TOL * abs(fd)


def test_full_size_waveform_gradients_match_finite_differences(full):
    sc, w0, target, tr, dw, ds, dw2, ds2 = full
    # The traces are linear in the waveform, so the loss is quadratic in it and
    # a central difference is exact for any step: a large one (20 against
    # samples of unit variance) keeps the difference far above the FP32
    # rounding of the 4000-step runs (a step of 0.5 moved the loss by too
    # little of itself at this size: 2.6 % error at t = 500).
    eps = 20.0
    l0 = float(np.sum((tr - target) ** 2))
    for t in (500, 2900):
        wp, wm = w0.copy(), w0.copy()
        wp[t] += eps
        wm[t] -= eps
        lp, lm = loss(sc, wp, np.ones(6), target), loss(sc, wm, np.ones(6), target)
        fd = (lp - lm) / (2 * eps)
        print("wave[%d] adjoint %.6e fd %.6e (loss %.4e, change %.2e of it)" % (
            t, dw[0, t], fd, l0, abs(lp - lm) / l0), flush=True)
        assert abs(dw[0, t] - fd) <= TOL * abs(fd)
