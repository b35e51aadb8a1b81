"""Precision of the transform kernels on the batch-of-rooms workload: 64
independent rooms (sides U[2,6] m, 1-4 boxes each, h = 0.05 m) advanced in one
batched context for 1500 steps, with the per-axis transforms in FP32 on
vector FMA, FP32 on matrix cores (plain and even/odd folded) and BF16 split
precision on matrix cores;
and, as the FFT-style point of comparison, the naive torch restatement with
FFT-based DCTs in FP32. Each is compared with the FP64 torch restatement
(on the GPU) on the first rooms; the errors are printed per result."""
import numpy as np
import pytest
import torch

import ard_scenes as S
import ard_sim as A
import torch_ref

pytestmark = pytest.mark.gpu

STEPS = 1500
CHECK_ROOMS = 1  # the FP64 restatement steps one room in about a minute


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def study():
    rooms = S.split_rooms(S.rooms64(steps=STEPS))
    ref = []
    for sc in rooms[:CHECK_ROOMS]:
        R = torch_ref.TorchARD(sc, device="cuda", dtype=torch.float64)
        with torch.no_grad():
            ref.append(R.run([None] * len(sc.sources), torch.ones(6, dtype=torch.float64, device="cuda"), STEPS)
                       .cpu().numpy())
    out = {}
    for v, name in ((0, "fp32 vector"), (1, "fp32 mfma"), (2, "bf16x3 mfma"), (3, "fp32 mfma folded")):
        A.set_transform_variant(v)
        b = A.Batch(rooms)
        b.update(STEPS)
        tr = [b.recorders(r, 0, STEPS).astype(np.float64) for r in range(CHECK_ROOMS)]
        b.close()
        out[name] = rel(np.concatenate(tr), np.concatenate(ref))
    A.set_transform_variant(1)
    fft = []
    for sc in rooms[:CHECK_ROOMS]:
        R = torch_ref.TorchARD(sc, device="cuda", dtype=torch.float32, dct="fft")
        with torch.no_grad():
            fft.append(R.run([None] * len(sc.sources), torch.ones(6, dtype=torch.float32, device="cuda"), STEPS)
                       .double().cpu().numpy())
    out["fp32 fft-style (torch)"] = rel(np.concatenate(fft), np.concatenate(ref))
    for k, e in out.items():
        print("recorder rel-L2 vs FP64, %d of %d rooms, %d steps: %-24s %.3e" % (CHECK_ROOMS, len(rooms), STEPS, k, e))
    return out


def test_fp32_transforms_within_tolerance(study):
    assert study["fp32 vector"] <= 1e-4
    assert study["fp32 mfma"] <= 1e-4
    assert study["fp32 mfma folded"] <= 1e-4


def test_bf16_split_error_is_bounded(study):
    # BF16 split products carry about 2^-16 relative error per transform pass
    assert study["bf16x3 mfma"] <= 5e-5


def test_fft_style_reference_point(study):
    assert study["fp32 fft-style (torch)"] <= 1e-4
