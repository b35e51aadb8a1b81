"""The discrete adjoint (ard_adjoint) on the differentiation workload at 1/4
scale: the 10 x 8 x 4 m room of 3 partitions scaled to 2.5 x 2 x 1 m (h =
0.05 m, one source of seeded band-limited noise, 8 recorders). Loss = squared
L2 between the traces and a target made by the same solver with the
waveform x 1.05 and the damping of every PML face divided by 1.2 (the
baseline's damping is 1.2 x the target's; scaling the damping up by 1.2
makes the reference's own PML update diverge, as the oracle shows for this
room). Gradients with respect to the waveform and the six damping scales are
checked against central finite differences, against autograd through the
naive torch restatement (tests/torch_ref.py), and through the
torch.autograd wrapper."""
import numpy as np
import pytest
import torch

import ard_autograd
import ard_scenes as S
import ard_sim as A
import torch_ref

pytestmark = pytest.mark.gpu

TOL = 1e-3  # adjoint vs finite differences, relative, per gradient entry


def quarter_room(steps):
    sc = S.diff_room(scale=4, steps=steps)
    sc.sources[0].wave = sc.sources[0].wave[:steps].astype(np.float32)
    return sc


def traces(sc, wave, theta, steps):
    s2 = S.Scene(**{**sc.__dict__})
    s2.sources = [S.Source(sc.sources[0].x, sc.sources[0].y, sc.sources[0].z, wave=np.asarray(wave, np.float32))]
    sim = A.Simulation(s2, form=A.FORM_PLAIN, pml_scale=np.asarray(theta, np.float32))
    sim.update(steps)
    return sim, sim.recorders(0, steps).astype(np.float64)


@pytest.fixture(scope="module")
def problem():
    steps = 4000  # the differentiation workload's length
    sc = quarter_room(steps)
    w0 = sc.sources[0].wave.astype(np.float64)
    s, target = traces(sc, w0 * 1.05, np.full(6, 1 / 1.2), steps)
    s.close()
    sim, tr = traces(sc, w0, np.ones(6), steps)
    dw, ds = sim.adjoint(2 * (tr - target))
    sim.close()
    return sc, steps, w0, target, tr, dw, ds


def loss(sc, w, th, target, steps):
    sim, tr = traces(sc, w, th, steps)
    sim.close()
    return float(np.sum((tr - target) ** 2))


def test_adjoint_damping_scales_match_finite_differences(problem):
    sc, steps, w0, target, tr, dw, ds = problem
    assert np.all(np.isfinite(tr)) and np.abs(tr).max() > 0
    for f in range(6):
        eps = 1e-2  # central differences: O(eps^2) truncation, well above FP32 noise of the loss
        tp, tm = np.ones(6), np.ones(6)
        tp[f] += eps
        tm[f] -= eps
        fd = (loss(sc, w0, tp, target, steps) - loss(sc, w0, tm, target, steps)) / (2 * eps)
        print("face %d adjoint %.6e fd %.6e" % (f, ds[f], fd))
        assert abs(ds[f] - fd) <= TOL * abs(fd)


def test_adjoint_waveform_matches_finite_differences(problem):
    sc, steps, w0, target, tr, dw, ds = problem
    for t in (3, 150, 1420, 2777, 3960):
        wp, wm = w0.copy(), w0.copy()
        wp[t] += 0.5
        wm[t] -= 0.5
        fd = (loss(sc, wp, np.ones(6), target, steps) - loss(sc, wm, np.ones(6), target, steps)) / 1.0
        print("wave[%d] adjoint %.6e fd %.6e" % (t, dw[0, t], fd))
        assert abs(dw[0, t] - fd) <= TOL * abs(fd)


def test_adjoint_matches_naive_autograd():
    """Autograd through the torch restatement (FP64, on the GPU) of the same
    loss: both gradients agree to FP32 rounding."""
    steps = 300
    sc = quarter_room(steps)
    w0 = sc.sources[0].wave.astype(np.float64)
    _, target = traces(sc, w0 * 1.05, np.full(6, 1 / 1.2), steps)
    sim, tr = traces(sc, w0, np.ones(6), steps)
    dw, ds = sim.adjoint(2 * (tr - target))
    sim.close()
    R = torch_ref.TorchARD(sc, device="cuda", dtype=torch.float64)
    w = torch.tensor(w0, dtype=torch.float64, device="cuda", requires_grad=True)
    th = torch.ones(6, dtype=torch.float64, device="cuda", requires_grad=True)
    tt = torch.tensor(target, dtype=torch.float64, device="cuda")
    out = R.run([w], th, steps)
    assert np.linalg.norm(out.detach().cpu().numpy() - tr) / np.linalg.norm(tr) < 1e-4
    ((out - tt) ** 2).sum().backward()
    gw, gs = w.grad.cpu().numpy(), th.grad.cpu().numpy()
    print("dscale adjoint", ds, "autograd", gs)
    assert np.linalg.norm(ds - gs) / np.linalg.norm(gs) < 1e-3
    assert np.linalg.norm(dw[0] - gw) / np.linalg.norm(gw) < 1e-3


def test_checkpoint_stride_does_not_change_gradients():
    steps = 200
    sc = quarter_room(steps)
    w0 = sc.sources[0].wave
    sim, tr = traces(sc, w0, np.ones(6), steps)
    g = np.ascontiguousarray(2 * tr, dtype=np.float32)
    ref = sim.adjoint(g, checkpoint_stride=1)
    for k in (0, 7, 200):
        dw, ds = sim.adjoint(g, checkpoint_stride=k)
        # equal up to the order of atomic additions (interface transpose, damping sums)
        assert np.linalg.norm(dw - ref[0]) <= 1e-5 * np.linalg.norm(ref[0])
        assert np.linalg.norm(ds - ref[1]) <= 1e-5 * np.linalg.norm(ref[1])
    # the context is left in its forward state: same traces as the first run
    assert np.array_equal(sim.recorders(0, steps).astype(np.float64), tr)
    sim.close()


@pytest.mark.parametrize("form", [A.FORM_PLAIN, A.FORM_BATCHED])
def test_torch_autograd_wrapper(form):
    steps = 200
    sc = quarter_room(steps)
    w = torch.tensor(sc.sources[0].wave[None, :].astype(np.float64), requires_grad=True)
    th = torch.ones(6, dtype=torch.float64, requires_grad=True)
    tr = ard_autograd.recorder_traces(sc, w, th, steps, form=form)
    assert tr.shape == (steps, len(sc.recorders))
    (tr.double() ** 2).sum().backward()
    sim, t2 = traces(sc, sc.sources[0].wave, np.ones(6), steps)
    dw, ds = sim.adjoint(2 * t2)
    sim.close()
    # equal up to the order of atomic additions in the reverse sweep (and,
    # for the batched forward, the FP32 rounding of its traces in the loss)
    assert np.linalg.norm(w.grad.numpy() - dw) <= 1e-5 * np.linalg.norm(dw)
    assert np.linalg.norm(th.grad.numpy() - ds) <= 1e-5 * np.linalg.norm(ds)


def batched_traces(sc, wave, theta, steps):
    s2 = S.Scene(**{**sc.__dict__})
    s2.sources = [S.Source(sc.sources[0].x, sc.sources[0].y, sc.sources[0].z, wave=np.asarray(wave, np.float32))]
    sim = A.Simulation(s2, form=A.FORM_BATCHED, pml_scale=np.asarray(theta, np.float32))
    sim.update(steps)
    return sim, sim.recorders(0, steps).astype(np.float64)


def test_batched_context_adjoint():
    """A batched context's adjoint (run on its plain-form twin): the plain
    context's gradients for the same loss, the batched context's own state
    untouched, and the damping-scale gradients against central finite
    differences of the batched form's own forward loss."""
    steps = 400
    sc = quarter_room(steps)
    w0 = sc.sources[0].wave.astype(np.float64)
    s, target = batched_traces(sc, w0 * 1.05, np.full(6, 1 / 1.2), steps)
    s.close()
    simb, trb = batched_traces(sc, w0, np.ones(6), steps)
    g = 2 * (trb - target)
    dwb, dsb = simb.adjoint(g)
    assert simb.time_step == steps
    assert np.array_equal(simb.recorders(0, steps).astype(np.float64), trb)
    simb.close()
    simp, trp = traces(sc, w0, np.ones(6), steps)
    dwp, dsp = simp.adjoint(g)
    simp.close()
    print("batched vs plain forward rel-L2 %.2e" % (np.linalg.norm(trb - trp) / np.linalg.norm(trp)))
    assert np.linalg.norm(dwb - dwp) <= 1e-5 * np.linalg.norm(dwp)
    assert np.linalg.norm(dsb - dsp) <= 1e-5 * np.linalg.norm(dsp)

    def bloss(th):
        sim, tr = batched_traces(sc, w0, th, steps)
        sim.close()
        return float(np.sum((tr - target) ** 2))

    for f in (0, 3, 5):
        eps = 1e-2
        tp, tm = np.ones(6), np.ones(6)
        tp[f] += eps
        tm[f] -= eps
        fd = (bloss(tp) - bloss(tm)) / (2 * eps)
        print("face %d batched adjoint %.6e fd %.6e" % (f, dsb[f], fd))
        assert abs(dsb[f] - fd) <= TOL * abs(fd)
