"""Precision of the transform kernels on the batch-of-rooms workload: 64
independent rooms (sides U[2,6] m, 1-4 boxes each, h = 0.05 m) advanced in one
batched context for 1500 steps, with the per-axis transforms in FP32 on
vector FMA, FP32 on matrix cores (plain and even/odd folded), BF16 split
precision on matrix cores (folded, FP32 accumulation) and the FFT-style DCT
(FP32). Every room's recorder trace is compared with the FP64 oracle's
(tests/golden/rooms64_fp64.npy: the 64 rooms as one oracle scene, the rooms
far enough apart never to touch); the errors are printed per result, overall
and per room (median, worst, and the smallest and largest room)."""
import os

import numpy as np
import pytest

import ard_scenes as S
import ard_sim as A

pytestmark = pytest.mark.gpu

STEPS = 1500
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rooms64_fp64.npy")
VARIANTS = ((0, "fp32 vector"), (1, "fp32 mfma"), (3, "fp32 mfma folded"), (2, "bf16x3 mfma folded"),
            (4, "fp32 fft-style"))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fft_lengths_ok(sc):
    """Every side of every box factors into the FFT kernel's radices."""
    def ok(n):
        for p in (2, 3, 5, 7, 11, 13):
            while n % p == 0:
                n //= p
        return n == 1
    return all(ok(b[k]) for b in sc.boxes for k in (3, 4, 5))


@pytest.fixture(scope="module")
def study():
    rooms = S.split_rooms(S.rooms64(steps=STEPS))
    ref = np.load(GOLD)
    assert ref.shape == (STEPS, len(rooms)) and all(len(sc.recorders) == 1 for sc in rooms)
    cells = np.array([sc.dct_cells() for sc in rooms])
    small, large = int(np.argmin(cells)), int(np.argmax(cells))
    out, traces = {}, {}
    try:
        for v, name in VARIANTS:
            A.set_transform_variant(v)
            idx = list(range(len(rooms)))
            if v == 4:  # rooms whose sides all factor (the others would fall back to the matrix cores)
                idx = [i for i, sc in enumerate(rooms) if fft_lengths_ok(sc)]
            b = A.Batch([rooms[i] for i in idx])
            assert b.batched_config()[2] == v, (name, b.batched_config())
            b.update(STEPS)
            tr = np.stack([b.recorders(k, 0, STEPS)[:, 0].astype(np.float64) for k in range(len(idx))], axis=1)
            b.close()
            traces[name] = (idx, tr)
            per = np.array([rel(tr[:, k], ref[:, i]) for k, i in enumerate(idx)])
            out[name] = {"all": rel(tr, ref[:, idx]), "rooms": len(idx), "median": float(np.median(per)),
                         "worst": float(per.max()),
                         "smallest": float(per[idx.index(small)]) if small in idx else None,
                         "largest": float(per[idx.index(large)]) if large in idx else None}
    finally:
        A.set_transform_variant(1)
    # the variants' distance from the FP32 matrix-core traces (their own
    # rounding, far below the step's other FP32 rounding)
    idx1, tr1 = traces["fp32 mfma"]
    for name, (idx, tr) in traces.items():
        out[name]["vs_fp32_mfma"] = rel(tr, tr1[:, idx])
    print("recorder rel-L2 vs FP64 oracle, %d steps; rooms %d (%d cells) .. %d (%d cells)"
          % (STEPS, small, cells[small], large, cells[large]))
    for k, e in out.items():
        print("  %-20s %2d rooms: all %.3e  median %.3e  worst %.3e  smallest room %s  largest room %s"
              "  (vs fp32 mfma %.2e)"
              % (k, e["rooms"], e["all"], e["median"], e["worst"],
                 "%.3e" % e["smallest"] if e["smallest"] is not None else "-",
                 "%.3e" % e["largest"] if e["largest"] is not None else "-", e["vs_fp32_mfma"]))
    return out


def test_fp32_transforms_within_tolerance(study):
    for name in ("fp32 vector", "fp32 mfma", "fp32 mfma folded"):
        assert study[name]["rooms"] == 64 and study[name]["all"] <= 1e-4, name
        assert study[name]["worst"] <= 1e-4, name


def test_bf16_split_error_is_bounded(study):
    # BF16 split products carry about 2^-16 relative error per transform pass
    assert study["bf16x3 mfma folded"]["rooms"] == 64 and study["bf16x3 mfma folded"]["all"] <= 5e-5


def test_fft_style_within_tolerance(study):
    assert study["fp32 fft-style"]["rooms"] >= 8 and study["fp32 fft-style"]["all"] <= 1e-4
