"""Batch of independent rooms (ard_batch_*) on the GPU: each room's recorder
traces in the batch equal the room run on its own."""
import numpy as np
import pytest

import ard_scenes as S
import ard_sim as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [A.FORM_BATCHED, A.FORM_PLAIN])
def test_batch_rooms_match_rooms_alone(form):
    steps = 300
    rooms = S.split_rooms(S.rooms64(n_rooms=4, steps=steps))
    b = A.Batch(rooms, form=form)
    b.update(steps)
    for r, sc in enumerate(rooms):
        sim = A.Simulation(sc, form=form)
        sim.update(steps)
        alone = sim.recorders(0, steps)
        sim.close()
        got = b.recorders(r, 0, steps)
        assert np.abs(alone).max() > 0
        err = np.linalg.norm(got - alone) / np.linalg.norm(alone)
        print("room %d rel-L2 %.2e" % (r, err))
        assert err <= 1e-6
    b.close()


@pytest.mark.parametrize("what", ["sources", "recorders"])
def test_batched_refuses_more_than_four_points_per_partition(what):
    """The batched form carries at most 4 interior point sources and 4
    recorders per DCT partition in its modal pass (the plain form has no such
    limit, like the reference); a fifth is refused at finalize with
    ARD_ERR_UNSUPPORTED, not dropped."""
    sc = S.two_rooms()
    sc.recorder_mode = S.REC_LOCAL
    pts = [(8 + 3 * i, 12, 12) for i in range(5)]  # interior cells of box 0 (no face slab)
    if what == "sources":
        sc.sources = [S.Source(*p) for p in pts]
    else:
        sc.recorders = pts
    with pytest.raises(A.ArdError, match="error -5"):
        A.Simulation(sc, form=A.FORM_BATCHED)
    A.Simulation(sc, form=A.FORM_PLAIN).close()  # the plain form takes five
    if what == "sources":  # and the batched form four
        sc.sources = sc.sources[:4]
    else:
        sc.recorders = sc.recorders[:4]
    A.Simulation(sc, form=A.FORM_BATCHED).close()
