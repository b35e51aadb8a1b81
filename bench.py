"""Throughput of the ARD time step on one MI355X.

Metric: Gcell-steps/s = (DCT cells + PML cells) * steps / seconds / 1e9 over
the timed steps, with the scene resident in HBM. Default workload: the
L-shaped 40 x 25 x 10 m hall at h = 0.05 m (20 boxes, 74M DCT cells), FP32,
batched form; the plain per-partition form is timed beside it.

    python bench.py [--steps K] [--warmup W] [--form batched|plain]
                    [--workload hall|scene3|two_rooms] [--no-cpu-baseline]
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
# algorithmic bytes per unit of work (dense formulation, FP32)
B_DCT, B_PML, B_FACE = 20, 44, 48


def scene_by_name(name):
    return {"hall": S.hall, "scene3": S.scene3, "two_rooms": S.two_rooms}[name]()


def cell_counts(sim):
    parts = sim.partitions()
    n_dct = sum(p[6] * p[7] * p[8] for p in parts if p[0] == 0)
    n_pml = sum(p[6] * p[7] * p[8] for p in parts if p[0] == 1)
    faces = 0
    for b in sim.boundaries():
        t, xs, xe, ys, ye, zs, ze = b[0], b[3], b[4], b[5], b[6], b[7], b[8]
        faces += {0: (ye - ys) * (ze - zs), 1: (xe - xs) * (ze - zs), 2: (xe - xs) * (ye - ys)}[t]
    return n_dct, n_pml, faces


def time_form(sc, form, steps, warmup):
    sim = A.Simulation(sc, form=form)
    sim.update(warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.update(steps)
    sim.synchronize()
    dt = time.perf_counter() - t0
    return sim, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--form", default="plain", choices=["batched", "plain"])
    ap.add_argument("--workload", default="hall")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the ARD step here runs on a single MI355X "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    sc = scene_by_name(args.workload)
    form = A.FORM_BATCHED if args.form == "batched" else A.FORM_PLAIN
    sim, secs = time_form(sc, form, args.steps, args.warmup)
    n_dct, n_pml, faces = cell_counts(sim)
    cells = n_dct + n_pml
    value = cells * args.steps / secs / 1e9
    ms = secs * 1e3 / args.steps
    alg_bytes = B_DCT * n_dct + B_PML * n_pml + B_FACE * faces
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
        "config": {"workload": "%s (%s form)" % (sc.name, args.form), "dct_cells": n_dct, "pml_cells": n_pml,
                   "face_elements": faces, "h": sc.h, "dt": sc.dt, "boxes": len(sc.boxes)},
        "step_roofline": {"alg_bytes_per_step": alg_bytes,
                          "achieved_GBs": round(alg_bytes / (ms * 1e-3) / 1e9, 1)},
    }
    sim.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
