"""Batch of independent rooms (ard_batch_*) on the GPU: each room's recorder
traces in the batch equal the room run on its own."""
import numpy as np
import pytest

import ard_scenes as S
import ard_sim as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form,variant", [(A.FORM_BATCHED, 1), (A.FORM_PLAIN, 1), (A.FORM_PLAIN, 3)])
def test_batch_rooms_match_rooms_alone(form, variant):
    """Rooms in one batched or plain context against each room alone (the
    plain form with its dense passes, and with its folded passes grouped
    across the partitions of all rooms)."""
    steps = 300
    rooms = S.split_rooms(S.rooms64(n_rooms=4, steps=steps))
    A.set_transform_variant(variant)
    try:
        b = A.Batch(rooms, form=form)
    finally:
        A.set_transform_variant(1)
    b.update(steps)
    for r, sc in enumerate(rooms):
        A.set_transform_variant(variant)
        try:
            sim = A.Simulation(sc, form=form)
        finally:
            A.set_transform_variant(1)
        sim.update(steps)
        alone = sim.recorders(0, steps)
        sim.close()
        got = b.recorders(r, 0, steps)
        assert np.abs(alone).max() > 0
        err = np.linalg.norm(got - alone) / np.linalg.norm(alone)
        print("room %d rel-L2 %.2e" % (r, err))
        assert err <= 1e-6
    b.close()


def many_points_scene():
    """small three-box scene whose first box holds 7 interior point sources
    (Gaussian and waveform) and 9 recorders, past the 4 of each that the
    batched form's modal pass carries; one more recorder in box 1."""
    sc = S.two_rooms()
    sc.boxes = [(0, 0, 0, 24, 20, 24), (24, 0, 0, 16, 28, 24), (0, 20, 0, 24, 8, 24)]
    sc.sources = [S.Source(6 + i, 6 + i % 4, 7 + i,
                           wave=None if i % 3 else S.band_limited_noise(200, seed=i))
                  for i in range(7)]
    # recorders more than 5 cells inside their box (Recorder::FindPartition's margin)
    sc.recorders = [(6 + i, 6 + (i * 5) % 8, 6 + (i * 7) % 12) for i in range(9)] + [(30, 12, 10)]
    sc.recorder_mode = S.REC_LOCAL
    return sc


@pytest.mark.parametrize("form", [A.FORM_BATCHED, A.FORM_PLAIN])
def test_many_points_per_partition_match_oracle(form):
    """Point sources and recorders per partition are not limited (as in the
    reference, ARD/simulation.cpp:37-48, ARD/recorder.cpp:26-38): in the
    batched form the modal pass carries 4 of each and the rest go through
    the extra passes before and after it. Recorder traces and every
    partition's pressure against the FP32 oracle."""
    import ard_oracle as O

    sc = many_points_scene()
    sim = A.Simulation(sc, form=form)
    o = O.from_scene(sc)
    assert sim.partitions() == o.partitions()
    n = 60
    sim.update(n)
    ref_tr = o.step(n)
    got = sim.recorders()
    assert got.shape == ref_tr.shape and np.abs(ref_tr).max() > 0
    per_rec = [np.linalg.norm(got[:, r] - ref_tr[:, r]) / max(np.linalg.norm(ref_tr[:, r]), 1e-30) for r in range(got.shape[1])]
    print("form %d per-recorder rel-L2 %s" % (form, " ".join("%.1e" % e for e in per_rec)))
    assert max(per_rec) < 1e-4
    for i, info in enumerate(sim.partitions()):
        f = sim.pressure_field(i)
        ref = o.field(i, 0)
        err = np.abs(f - ref).max() / np.abs(ref).max()
        assert err < 1e-4, (i, info, err)
    sim.close()


def test_many_points_under_modal_tilings():
    """The extra passes under modal tilings of 2, 4 and 8 rows per wave (the
    modal pass's blocks and partials change, the extra passes' do not): the
    traces agree."""
    sc = many_points_scene()
    n = 60
    base = None
    for rpw, kzc in [(2, 16), (4, 48), (8, 96)]:
        sim = A.Simulation(sc, form=A.FORM_BATCHED, autotune=False)
        sim.set_batched_tiling(rpw, kzc)
        sim.update(n)
        tr = sim.recorders()
        sim.close()
        if base is None:
            base = tr
        err = np.linalg.norm(tr - base) / np.linalg.norm(base)
        print("tiling %d/%d rel-L2 vs first %.2e" % (rpw, kzc, err))
        assert err < 1e-5


def test_many_points_graph_replays_match_direct_launches():
    """The extra point passes inside the captured step graphs (one-step and
    8-step graphs, from even and odd steps) against direct launches:
    identical recorder samples and fields."""
    sc = many_points_scene()
    g = A.Simulation(sc, form=A.FORM_BATCHED)
    d = A.Simulation(sc, form=A.FORM_BATCHED)
    d.set_kernel_timing(True)
    t = 0
    for chunk, n in enumerate((7, 8, 17, 3, 16, 9)):
        g.update(n)
        d.update(n)
        t += n
        assert np.array_equal(g.recorders(0, t), d.recorders(0, t)), chunk
    for i in range(len(g.partitions())):
        assert np.array_equal(g.pressure_field(i), d.pressure_field(i)), i
    g.close()
    d.close()
