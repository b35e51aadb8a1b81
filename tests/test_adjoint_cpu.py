"""The naive autograd reference (tests/torch_ref.py) against the oracle, and
the adjoint entry point's host-side contract. The adjoint itself runs in
test_gpu_adjoint.py."""
import numpy as np
import pytest
import torch

import ard_oracle as O
import ard_scenes as S
import ard_sim as A
import torch_ref


def quarter_room(steps):
    sc = S.diff_room(scale=4, steps=steps)
    sc.sources[0].wave = sc.sources[0].wave[:steps].astype(np.float32)
    return sc


def test_torch_reference_matches_oracle():
    """The differentiable restatement reproduces the FP64 oracle's traces
    (waveform source, 3 partitions, PML slabs on every exterior face)."""
    steps = 40
    sc = quarter_room(steps)
    R = torch_ref.TorchARD(sc)
    w = torch.tensor(sc.sources[0].wave, dtype=torch.float64)
    tr = R.run([w], torch.ones(6, dtype=torch.float64), steps).numpy()
    ot = O.from_scene(sc, double=True).step(steps)
    assert np.abs(ot).max() > 0
    assert np.linalg.norm(tr - ot) / np.linalg.norm(ot) < 1e-9


def test_torch_reference_damping_scale_matches_oracle():
    steps = 30
    sc = quarter_room(steps)
    th = np.array([0.7, 1.3, 0.9, 1.1, 0.8, 1.2])
    R = torch_ref.TorchARD(sc)
    w = torch.tensor(sc.sources[0].wave, dtype=torch.float64)
    tr = R.run([w], torch.tensor(th), steps).numpy()
    ot = O.from_scene(sc, double=True, pml_scale=th).step(steps)
    assert np.linalg.norm(tr - ot) / np.linalg.norm(ot) < 1e-9


def test_torch_reference_gradient_matches_finite_difference():
    """Autograd through the reference's loop agrees with central differences
    of its own loss (pins the naive reference's differentiability)."""
    steps = 25
    sc = quarter_room(steps)
    R = torch_ref.TorchARD(sc)
    w0 = torch.tensor(sc.sources[0].wave, dtype=torch.float64)
    tgt = R.run([w0 * 1.05], torch.full((6,), 1 / 1.2, dtype=torch.float64), steps).detach()

    def loss(w, th):
        return ((R.run([w], th, steps) - tgt) ** 2).sum()

    th = torch.ones(6, dtype=torch.float64, requires_grad=True)
    w = w0.clone().requires_grad_(True)
    loss(w, th).backward()
    for f in (0, 3):
        e = 1e-4
        tp, tm = torch.ones(6, dtype=torch.float64), torch.ones(6, dtype=torch.float64)
        tp[f] += e
        tm[f] -= e
        fd = (loss(w0, tp) - loss(w0, tm)).item() / (2 * e)
        assert abs(th.grad[f].item() - fd) <= 1e-5 * abs(fd) + 1e-40
    t = 7
    wp, wm = w0.clone(), w0.clone()
    wp[t] += 0.1
    wm[t] -= 0.1
    fd = (loss(wp, torch.ones(6, dtype=torch.float64)) - loss(wm, torch.ones(6, dtype=torch.float64))).item() / 0.2
    assert abs(w.grad[t].item() - fd) <= 1e-6 * abs(fd) + 1e-40


def test_adjoint_rejects_bad_contexts():
    sc = quarter_room(10)
    host = A.Simulation(sc, device=-1, finalize=False)
    g = np.zeros((10, len(sc.recorders)), np.float32)
    with pytest.raises(A.ArdError):
        host.adjoint(g)  # not finalized
    with pytest.raises(ValueError):
        host.adjoint(np.zeros((10, 3), np.float32))  # wrong recorder count
    host.close()


@pytest.mark.parametrize("n", [5, 8, 13])
def test_fft_style_dct_matches_the_matrix_transforms(n):
    """The FFT-style transform of the naive reference against the definitions."""
    rng = np.random.default_rng(n)
    v = torch.tensor(rng.standard_normal((3, 4, n)), dtype=torch.float64)
    c2, c3 = torch_ref._cos_mats(n, torch.float64, "cpu")
    assert torch.allclose(torch_ref._dct_fft(v, 2, False), v @ c2.t(), rtol=0, atol=1e-12)
    assert torch.allclose(torch_ref._dct_fft(v, 2, True), v @ c3.t(), rtol=0, atol=1e-12)
