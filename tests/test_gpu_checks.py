"""GPU checks of the failure detection and schedule options of the library:
ard_check_finite on a clean state and after a NaN force is injected from the
host (the condition is forced from the host; kernels run normally on NaN
data), and the batched form's schedule option (the LDS ring of the modal
pass), which must leave the recorder traces bit-identical."""
import numpy as np
import pytest

import ard_scenes as S
import ard_sim as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [A.FORM_PLAIN, A.FORM_BATCHED])
def test_check_finite_reports_nan_state(form):
    sc = S.two_rooms()
    print("case check_finite form", form, flush=True)
    sim = A.Simulation(sc, form=form)
    sim.update(40)
    assert sim.check_finite() == -1
    # a NaN force in the second room's interior (partition 1): after one step its
    # modes and pressure hold NaN, and the interface carries it to partition 0
    sim.set_force(1, 12, 14, 15, float("nan"))
    sim.update(1)
    bad = sim.check_finite()
    print("first partition with non-finite state:", bad)
    assert bad in (0, 1)
    sim.close()


@pytest.mark.parametrize("form", [A.FORM_PLAIN, A.FORM_BATCHED])
def test_check_finite_nan_in_pml(form):
    sc = S.two_rooms()
    sim = A.Simulation(sc, form=form)
    parts = sim.partitions()
    pml = [i for i, p in enumerate(parts) if p[0] == 1]
    sim.update(5)
    assert sim.check_finite() == -1
    # the last PML slab: a NaN force enters its pressure in the next update
    last = pml[-1]
    sim.set_force(last, 1, 1, 1, float("inf"))
    sim.update(1)
    bad = sim.check_finite()
    print("PML slab", last, "-> first non-finite partition", bad)
    assert 0 <= bad <= last
    sim.close()


def test_batched_schedule_options_bit_identical():
    sc = S.scene3()
    steps = 300
    traces = {}
    for name, opts in (("default", {}), ("ring2", {A.OPT_RING: 2}), ("pml_shape1", {A.OPT_PML_SHAPE: 1}),
                       ("both", {A.OPT_RING: 2, A.OPT_PML_SHAPE: 1}), ("split100", {A.OPT_PML_SPLIT: 100}),
                       ("split50", {A.OPT_PML_SPLIT: 50}), ("split33_shape1", {A.OPT_PML_SPLIT: 33, A.OPT_PML_SHAPE: 1})):
        print("case options", name, flush=True)
        sim = A.Simulation(sc, form=A.FORM_BATCHED)
        sim.set_batched_tiling(4, 48)
        for k, v in opts.items():
            sim.set_batched_option(k, v)
            assert sim.batched_option(k) == v
        sim.update(steps)
        traces[name] = sim.recorders(0, steps)
        sim.close()
    for name, tr in traces.items():
        assert np.array_equal(tr, traces["default"]), name


@pytest.mark.parametrize("layers", [3, 5, 8])
def test_pml_tile_shapes_bit_identical_fields(layers):
    """The tiled PML update in both tile shapes: the same arithmetic per cell,
    so every partition's field (PML slabs included) is bit-identical."""
    sc = S.two_rooms()
    sc.boxes = [(0, 0, 0, 16, 12, 16), (16, 0, 0, 12, 20, 16), (0, 12, 0, 16, 8, 16)]
    sc.sources = [S.Source(5, 1, 4), S.Source(21, 10, 8)]
    sc.recorders = [(22, 10, 8)]
    sc.recorder_mode = S.REC_LOCAL
    sc.pml_layers = layers
    fields = {}
    for sh in (0, 1):
        sim = A.Simulation(sc, form=A.FORM_BATCHED)
        sim.set_batched_option(A.OPT_PML_SHAPE, sh)
        sim.update(40)
        fields[sh] = [sim.pressure_field(i) for i in range(len(sim.partitions()))]
        sim.close()
    for a, b in zip(fields[0], fields[1]):
        assert np.array_equal(a, b)


def test_autotuned_context_starts_from_the_zero_state():
    """The autotuner times its candidates on the step-0 state; afterwards the
    context holds that state again: an autotuned context and one pinned to
    the same choices give bit-identical traces, also with an interior
    waveform source whose step-0 sample is large (the modal pass's timed
    launches leave a residue of it in the previous modes) and PML slabs (the
    timed PML launches write their next buffers)."""
    sc = S.two_rooms()
    sc.boxes = [(0, 0, 0, 24, 20, 24), (24, 0, 0, 16, 28, 24), (0, 20, 0, 24, 8, 24)]
    wave = S.band_limited_noise(300, seed=4)
    wave[0] = 5.0 * np.abs(wave).max()
    sc.sources = [S.Source(10, 9, 11, wave=wave), S.Source(30, 12, 9)]
    sc.recorders = [(12, 10, 12), (30, 14, 10)]
    sc.recorder_mode = S.REC_LOCAL
    n = 120
    tuned = A.Simulation(sc, form=A.FORM_BATCHED, autotune=True)
    rpw, kzc, variant, _ = tuned.batched_config()
    ring, shape = tuned.batched_option(A.OPT_RING), tuned.batched_option(A.OPT_PML_SHAPE)
    tuned.update(n)
    got = tuned.recorders()
    tuned.close()
    A.set_transform_variant(variant)
    try:
        pinned = A.Simulation(sc, form=A.FORM_BATCHED, autotune=False)
    finally:
        A.set_transform_variant(1)
    pinned.set_batched_tiling(rpw, kzc)
    pinned.set_batched_option(A.OPT_RING, ring)
    pinned.set_batched_option(A.OPT_PML_SHAPE, shape)
    pinned.update(n)
    ref = pinned.recorders()
    pinned.close()
    print("autotuned", (rpw, kzc, variant, ring, shape), "max |diff|", np.abs(got - ref).max())
    assert np.abs(ref).max() > 0 and np.array_equal(got, ref)


def test_batched_option_values_checked():
    sim = A.Simulation(S.two_rooms(), form=A.FORM_BATCHED)
    with pytest.raises(A.ArdError):
        sim.set_batched_option(A.OPT_RING, 3)
    with pytest.raises(A.ArdError):
        sim.set_batched_option(A.OPT_PML_SHAPE, 2)
    for bad in (-1, 101):
        with pytest.raises(A.ArdError):
            sim.set_batched_option(A.OPT_PML_SPLIT, bad)
    assert sim.batched_option(A.OPT_PML_SPLIT) == 0
    with pytest.raises(A.ArdError):
        sim.set_batched_option(99, 1)
    sim.close()
    plain = A.Simulation(S.two_rooms(), form=A.FORM_PLAIN)
    with pytest.raises(A.ArdError):
        plain.set_batched_option(A.OPT_RING, 0)
    plain.close()


@pytest.mark.parametrize("form", [A.FORM_PLAIN, A.FORM_BATCHED])
def test_ard_run_info_and_finite_checks(tmp_path, form):
    """ard_run (the C++ mirror's main loop) with --info (Partition::Info and
    Boundary::Info of every partition and boundary, the boundaries being the
    mirror's Boundary handles) and --check-every (ard_check_finite between
    steps): the boundary lines match the topology, and a clean run passes."""
    import os
    import re
    import subprocess

    sc = S.scene3(steps=60)
    pkg = os.path.dirname(os.path.abspath(A.__file__))
    S.write_cell_files(sc, str(tmp_path / "scene"))
    cmd = [os.path.join(pkg, "ard_run"), "--cells", "--no-output", "--info", "--check-every", "20"]
    if form == A.FORM_PLAIN:
        cmd.append("--plain")
    cmd += [str(tmp_path / "scene" / n) for n in ("scene.txt", "sources.txt", "recorders.txt")]
    cmd += ["60", repr(sc.h)]
    print("case ard_run --info --check-every form", form, flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    num = r"(-?\d+)"  # PML boundaries reach outside the scene: negative coordinates
    lines = re.findall(r"^Boundary (\d+): ([XYZ]) (\d+),(\d+) \| %s,%s,%s->%s,%s,%s$" % ((num,) * 6), r.stdout, re.M)
    sim = A.Simulation(sc, device=-1, finalize=False)
    bnds = sim.boundaries()
    assert len(lines) == len(bnds)
    for ln, b in zip(lines, bnds):
        typ = "XYZ"[b[0]]
        # Info prints start "xs,ys,zs" and end "xe,ye,ze"
        assert (ln[1], int(ln[2]), int(ln[3])) == (typ, b[1], b[2])
        assert tuple(int(v) for v in ln[4:]) == (b[3], b[5], b[7], b[4], b[6], b[8])


def test_pml_placements_bit_identical_traces():
    """The PML's launch placements (beside the forward transforms, beside the
    unfold and inverse transforms, or split between them) change no
    arithmetic: bit-identical traces and fields."""
    sc = S.scene3()
    steps = 200
    out = {}
    A.set_transform_variant(3)
    try:
        for name, opts in (("split0", {}), ("split100", {A.OPT_PML_SPLIT: 100}),
                           ("split40", {A.OPT_PML_SPLIT: 40}), ("split70_shape1", {A.OPT_PML_SPLIT: 70, A.OPT_PML_SHAPE: 1})):
            print("case", name, flush=True)
            sim = A.Simulation(sc, form=A.FORM_BATCHED)
            for k, v in opts.items():
                sim.set_batched_option(k, v)
                assert sim.batched_option(k) == v
            sim.update(steps)
            out[name] = (sim.recorders(0, steps), [sim.pressure_field(i) for i in range(len(sim.partitions()))])
            sim.close()
    finally:
        A.set_transform_variant(1)
    for name, (tr, fields) in out.items():
        assert np.array_equal(tr, out["split0"][0]), name
        for a, b in zip(fields, out["split0"][1]):
            assert np.array_equal(a, b), name
