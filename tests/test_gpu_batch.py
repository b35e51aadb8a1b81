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
