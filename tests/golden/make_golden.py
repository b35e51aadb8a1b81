"""Writes the golden recorder traces and topology fixtures used by the parity tests.

The traces come from the CPU oracle (oracle/ard_oracle.cpp) in FP32 (the
reference's arithmetic) and FP64; the oracle is pinned by
tests/test_oracle_pins.py. Run from the repository root:

    python tests/golden/make_golden.py            # two rooms, scene-3, hall cell files
    python tests/golden/make_golden.py --hall     # hall traces (2000 steps; ~30 min per precision)
    python tests/golden/make_golden.py --rooms64  # 64-room batch, FP64 (1500 steps; ~15 min)

The hall traces (hall_fp64.npy, hall_fp32.npy: [2000][16] recorder samples of
the seeded L-shaped hall, 74M DCT + 6.3M PML cells) are computed with the
oracle's O(n log n) transforms; they state the FP32 forms' recorder error
against FP64 on the full-size batched workload.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ard-simulator_amd")]
import ard_oracle as O  # noqa: E402
import ard_scenes as S  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def hall_traces():
    sc = S.hall()
    for double in (True, False):
        o = O.from_scene(sc, double=double)
        chunks = []
        for t0 in range(0, sc.steps, 100):
            chunks.append(o.step(min(100, sc.steps - t0)))
            print("hall %s: %d steps" % ("fp64" if double else "fp32", t0 + len(chunks[-1])), flush=True)
        tr = np.concatenate(chunks)
        name = "hall_%s.npy" % ("fp64" if double else "fp32")
        np.save(os.path.join(OUT, name), tr.astype(np.float64 if double else np.float32))
        print(name, tr.shape, float(np.abs(tr).max()), flush=True)
        del o


def rooms64_trace():
    """The 64 independent rooms (sides U[2,6] m, 1-4 boxes each) in one oracle
    scene: rooms 2 m apart never touch, so each evolves as alone. FP64,
    [1500][64] (one recorder per room, room order)."""
    sc = S.rooms64()
    o = O.from_scene(sc, double=True)
    chunks = []
    for t0 in range(0, sc.steps, 100):
        chunks.append(o.step(min(100, sc.steps - t0)))
        print("rooms64 fp64: %d steps" % (t0 + len(chunks[-1])), flush=True)
    tr = np.concatenate(chunks)
    np.save(os.path.join(OUT, "rooms64_fp64.npy"), tr.astype(np.float64))
    print("rooms64_fp64.npy", tr.shape, float(np.abs(tr).max()), flush=True)


def main():
    if "--hall" in sys.argv[1:]:
        hall_traces()
        return
    if "--rooms64" in sys.argv[1:]:
        rooms64_trace()
        return
    # the seeded hall in cell units (ard_run --cells)
    S.write_cell_files(S.hall(), os.path.join(OUT, "hall_cells"))
    for sc in (S.two_rooms(), S.scene3()):
        topo = None
        for double in (False, True):
            o = O.from_scene(sc, double=double)
            if topo is None:
                topo = {"partitions": o.partitions(), "boundaries": o.boundaries(),
                        "recorders": sc.recorders, "recorder_partitions": [o.recorder_partition(i) for i in range(len(sc.recorders))]}
            tr = o.step(sc.steps)
            name = "%s_%s.npy" % (sc.name, "fp64" if double else "fp32")
            np.save(os.path.join(OUT, name), tr.astype(np.float64))
            print(name, tr.shape, float(np.abs(tr).max()), flush=True)
        with open(os.path.join(OUT, "%s_topology.json" % sc.name), "w") as f:
            json.dump(topo, f)


if __name__ == "__main__":
    main()
