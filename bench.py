"""Throughput of the ARD time step on one MI355X.

Metric: Gcell-steps/s = (DCT cells + PML cells) * steps / seconds / 1e9 over
the timed steps, with the whole scene resident in HBM. Workload: the
L-shaped 40 x 25 x 10 m hall at h = 0.05 m (20 boxes from seeded guillotine
cuts, 74M DCT + 6.3M PML cells, 8 sources, 16 recorders), FP32. The batched
form (hipGraph per step) gives `value`; the plain form (the reference's
structure: dense per-axis transforms of every partition) is timed beside it
on the same scene, with its even/odd folded passes (`plain`, the ratio's
denominator) and with its dense passes (`plain_dense`).

The metric's condition is checked on the hall's own run length, whatever
--steps is: after the timed steps both forms continue (untimed) to the
workload's 2000 steps, and their recorder traces over steps 0..1999 are
compared with the FP64 oracle's (tests/golden/hall_fp64.npy, written by
tests/golden/make_golden.py --hall): `recorder_rel_l2_vs_fp64` carries the
value, the window, the 1e-4 tolerance and whether it is met.

    python bench.py [--steps K] [--warmup W] [--workload hall|scene3|two_rooms]
                    [--no-plain] [--no-cpu-baseline] [--no-rooms64]

Beside the hall, the 64-room batch (batched plus precision) is timed with its
face transforms in FP32 and in BF16 split precision (`rooms64_precision`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "ard-simulator_amd")]

import ard_scenes as S  # noqa: E402
import ard_sim as A  # noqa: E402

METRIC = "Gcell-steps/s on 1 MI355X, batched vs plain GPU at recorder rel-L2 ≤1e-4"
HBM_PEAK_GBS = 8000.0
# transform kernels by variant number (ard_set_transform_variant)
TRANSFORM_NAMES = ["vector-fma", "mfma-fp32", "mfma-bf16x3", "mfma-fp32-folded", "fft", "mfma-fp32-folded-direct"]
FP32_PEAK_TFLOPS = 157.3
# algorithmic bytes per unit of work (dense formulation, FP32)
B_DCT, B_PML, B_FACE = 20, 44, 48


def scene_by_name(name):
    return {"hall": S.hall, "scene3": S.scene3, "two_rooms": S.two_rooms}[name]()


def counts(sim):
    parts = sim.partitions()
    n_dct = sum(p[6] * p[7] * p[8] for p in parts if p[0] == 0)
    n_pml = sum(p[6] * p[7] * p[8] for p in parts if p[0] == 1)
    faces = 0
    for b in sim.boundaries():
        t, xs, xe, ys, ye, zs, ze = b[0], b[3], b[4], b[5], b[6], b[7], b[8]
        faces += {0: (ye - ys) * (ze - zs), 1: (xe - xs) * (ze - zs), 2: (xe - xs) * (ye - ys)}[t]
    return n_dct, n_pml, faces


def face_gemm_flops(sim):
    """2*M*N*K of the four face-transform GEMM launches of one batched step."""
    fl = 0
    for p in sim.partitions():
        if p[0]:
            continue
        W, H, D = p[6:9]
        # forward and inverse: x faces (H then D), y faces (W then D), z faces (W then H)
        fl += 2 * 2 * 6 * (D * H * H + H * D * D + D * W * W + W * D * D + H * W * W + W * H * H)
    return fl


REL_L2_TOL = 1e-4


def fp64_reference(sc):
    """The FP64 oracle's recorder trace of the workload ([steps][recorders]), or None."""
    path = os.path.join(ROOT, "tests", "golden", "%s_fp64.npy" % sc.name)
    if not os.path.exists(path):
        return None
    return np.load(path)


def rel_l2_window(sim, done, ref):
    """Steps the simulation on (untimed) to len(ref) steps and returns the
    recorder rel-L2 against ref over all of them."""
    n = len(ref)
    if done < n:
        sim.update(n - done)
    tr = sim.recorders(0, n).astype(np.float64)
    err = float(np.linalg.norm(tr - ref) / np.linalg.norm(ref))
    return {"value": err, "steps": n, "tolerance": REL_L2_TOL, "met": bool(err <= REL_L2_TOL)}


def run_form(sc, form, steps, warmup, timing_steps, direct=False, autotune=False, ref=None):
    sim = A.Simulation(sc, form=form, autotune=autotune)
    if direct:
        sim.set_kernel_timing(True)
    sim.update(warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.update(steps)
    sim.synchronize()
    secs = time.perf_counter() - t0
    phases = {}
    if timing_steps:
        sim.set_kernel_timing(True)
        sim.update(timing_steps)
        sim.synchronize()
        phases = sim.kernel_times()
        sim.set_kernel_timing(False)
    done = warmup + steps + timing_steps
    parity = rel_l2_window(sim, done, ref) if ref is not None else None
    return sim, secs, phases, parity


TRAFFIC_SOURCE = ("profiles/traffic.json: HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                  "(calibrated by a device copy), recorded by a profiling run, not measured in this run")


def traffic_from_profile(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    recorded in profiles/traffic.json by the profiling run; None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("kernels", {}).get(kernel)


def rooms64_precision(steps=1500, timed=300):
    """The batched-plus-precision workload: 64 independent rooms (sides U[2,6] m,
    1-4 boxes each, h = 0.05 m) in one batched context, with the per-axis face
    transforms in FP32 (folded, matrix cores) and in BF16 split precision (three
    BF16 products, FP32 accumulate). Throughput of each over `timed` steps and
    the recorder rel-L2 of BF16 split against FP32 over `steps` steps (the
    errors against the FP64 restatement are the GPU precision tests')."""
    rooms = S.split_rooms(S.rooms64(steps=steps))
    out, traces = {}, {}
    for v, name in ((3, "fp32_folded"), (2, "bf16x3")):
        A.set_transform_variant(v)
        try:
            b = A.Batch(rooms)
        finally:
            A.set_transform_variant(3)
        parts = b.partitions()
        cells = sum(p[6] * p[7] * p[8] for p in parts)
        b.update(20)
        b.synchronize()
        t0 = time.perf_counter()
        b.update(timed)
        b.synchronize()
        secs = time.perf_counter() - t0
        b.update(steps - 20 - timed)
        traces[name] = np.concatenate([b.recorders(r, 0, steps).astype(np.float64).ravel()
                                       for r in range(len(rooms))])
        b.close()
        out[name] = {"value": round(cells * timed / secs / 1e9, 4), "unit": "Gcell-steps/s",
                     "ms_per_step": round(secs * 1e3 / timed, 4)}
    ref = traces["fp32_folded"]
    out["bf16x3"]["recorder_rel_l2_vs_fp32"] = float(np.linalg.norm(traces["bf16x3"] - ref) / np.linalg.norm(ref))
    out["rooms"] = len(rooms)
    out["cells"] = int(cells)
    return out


def cpu_baseline(sc, budget_s=15.0):
    """The CPU oracle timed on this host's cores: the FP32 restatement of the
    reference's step with its DCTs by FFT (O(n log n) REDFT10/REDFT01, the
    algorithm class of the reference's FFTW path), OpenMP over lines, cells
    and face elements. Steps run until about budget_s seconds have passed
    (at least 2); setup and the first step (plans, first touch) are excluded."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ard_oracle  # the checker, used here only as the timed CPU baseline

    cores = min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    o = ard_oracle.from_scene(sc, fft=True)
    n_dct = sc.dct_cells()
    parts = o.partitions()
    n_pml = sum(p[6] * p[7] * p[8] for p in parts if p[0] == 1)
    o.step(1)
    t0 = time.perf_counter()
    n = 0
    while n < 2 or time.perf_counter() - t0 < budget_s:
        o.step(1)
        n += 1
    secs = time.perf_counter() - t0
    del o
    return {"value": round((n_dct + n_pml) * n / secs / 1e9, 6), "unit": "Gcell-steps/s", "cores": cores,
            "kind": "port", "sample": "%d steps of the %s scene (%d cells) in FP32, DCTs by FFT, setup and first "
                                      "step excluded, %.1f s" % (n, sc.name, n_dct + n_pml, secs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--workload", default="hall")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rooms64", action="store_true",
                    help="skip the 64-room batch in FP32 and BF16 split precision")
    ap.add_argument("--transform-variant", type=int, default=1,
                    help="0 vector FMA, 1 matrix cores FP32, 2 matrix cores BF16 split, 3 matrix cores FP32 folded")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--direct", action="store_true",
                    help="launch each phase directly instead of replaying the step graph (counter collection)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this size (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this ARD path runs on a single MI355X; "
              "multi-GPU sharding is a design note only", file=sys.stderr)
        sys.exit(2)

    A.set_transform_variant(args.transform_variant)
    if args.copy_calibration_mib:
        import torch

        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sc = scene_by_name(args.workload)
    ref = fp64_reference(sc)
    sim, secs, phases, parity_b = run_form(sc, A.FORM_BATCHED, args.steps, args.warmup, args.timing_steps,
                                           args.direct, autotune=not args.no_autotune, ref=ref)
    tiling = sim.batched_config()
    n_dct, n_pml, faces = counts(sim)
    gflop_face = face_gemm_flops(sim)
    sim.close()
    cells = n_dct + n_pml
    value = cells * args.steps / secs / 1e9
    ms = secs * 1e3 / args.steps

    # roofline of the dominant kernel (the phase with the largest time per step)
    per_step = {k: v[0] * v[1] / max(args.timing_steps, 1) for k, v in phases.items()}
    dom = max(per_step, key=per_step.get) if per_step else None
    alg = {"modal": (B_DCT * n_dct, "hbm"), "pml": (B_PML * n_pml, "hbm"), "interface": (B_FACE * faces, "hbm")}
    roof = None
    if dom:
        avg_ms, launches = phases[dom]
        if dom in alg:
            # modal launches run once per rows-per-wave group; the DCT cells are split among them
            per_launch = alg[dom][0] * args.timing_steps / launches
            achieved = per_launch / (avg_ms * 1e-3) / 1e9
            roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profile(dom), "avg_ms": round(avg_ms, 4),
                    "traffic_source": TRAFFIC_SOURCE}
        else:
            per_launch = gflop_face / 4 * args.timing_steps / launches
            achieved = per_launch / (avg_ms * 1e-3) / 1e12
            roof = {"kernel": dom, "bound": "mfma", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                    "traffic": traffic_from_profile(dom), "avg_ms": round(avg_ms, 4),
                    "traffic_source": TRAFFIC_SOURCE}

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s: batched face-shell form, hipGraph per step" % sc.name,
                   "tiling": {"rows_per_wave": tiling[0], "kz_per_block": tiling[1],
                              "transform": TRANSFORM_NAMES[tiling[2]], "autotune_candidates": tiling[3]}, "dct_cells": n_dct,
                   "pml_cells": n_pml, "face_elements": faces, "boxes": len(sc.boxes), "h": sc.h, "dt": sc.dt,
                   "sources": len(sc.sources), "recorders": len(sc.recorders)},
        "roofline": roof,
        "step_algorithmic_GBs": round((B_DCT * n_dct + B_PML * n_pml + B_FACE * faces) / (ms * 1e-3) / 1e9, 1),
        "phase_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items(), key=lambda kv: -kv[1])},
    }
    if parity_b is not None:
        out["recorder_rel_l2_vs_fp64"] = parity_b
        if not parity_b["met"]:
            print("bench.py: batched recorder rel-L2 vs FP64 %.3g exceeds %.0e over %d steps" % (
                parity_b["value"], REL_L2_TOL, parity_b["steps"]), file=sys.stderr)
    if not args.no_plain:
        # the plain form (the reference's structure: per-partition dense
        # per-axis transforms, modal update, PML, interface forcing) in its
        # fastest configuration, the even/odd folded passes with pass j of
        # every partition in one launch (transform variant 3), which the
        # ratio is taken against; and with the dense passes (variant 1)
        plain = {}
        for v, key in ((3, "plain"), (1, "plain_dense")):
            A.set_transform_variant(v)
            try:
                simp, secs_p, _, parity_p = run_form(sc, A.FORM_PLAIN, args.steps, args.warmup, 0,
                                                     ref=ref if key == "plain" else None)
            finally:
                A.set_transform_variant(args.transform_variant)
            simp.close()
            n_p = args.steps
            plain[key] = cells * n_p / secs_p / 1e9
            out[key] = {"value": round(plain[key], 4), "ms_per_step": round(secs_p * 1e3 / n_p, 3), "steps": n_p,
                        "transform": TRANSFORM_NAMES[v]}
            if parity_p is not None:
                out[key]["recorder_rel_l2_vs_fp64"] = parity_p
        out["batched_over_plain"] = round(value / plain["plain"], 2)
        out["batched_over_plain_dense"] = round(value / plain["plain_dense"], 2)
    if not args.no_rooms64:
        out["rooms64_precision"] = rooms64_precision()
        A.set_transform_variant(args.transform_variant)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
