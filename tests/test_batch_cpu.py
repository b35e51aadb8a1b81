"""Batch of independent rooms (ard_batch_create), host side: the merged
context holds each room's own topology (partitions, PML slabs, boundaries in
order), shifted along x."""
import ard_scenes as S
import ard_sim as A


def room_topology(sc):
    sim = A.Simulation(sc, device=-1, finalize=False)
    parts, bnds = sim.partitions(), sim.boundaries()
    sim.close()
    return parts, bnds


def test_batch_keeps_each_rooms_topology():
    rooms = S.split_rooms(S.rooms64(n_rooms=5))
    assert len(rooms) == 5
    b = A.Batch(rooms, device=-1)
    mparts, mbnds = b.partitions(), b.boundaries()
    seen_parts = 0
    for r, sc in enumerate(rooms):
        parts, bnds = room_topology(sc)
        seen_parts += len(parts)
        first = b.room_partition(r, 0)
        shift = mparts[first][3] - parts[0][3]
        dct = [q for q in parts if not q[0]]
        x0, x1 = min(q[3] for q in dct) + shift, max(q[3] + q[6] for q in dct) + shift
        # the room's partitions, shifted, are exactly those of the merged context
        # from its left PML (x0 - pml layers) to its right PML (starting at x1)
        mine = [(p[0], p[1], p[3] - shift, p[4], p[5], p[6], p[7], p[8]) for p in mparts
                if x0 - sc.pml_layers <= p[3] <= x1]
        assert sorted(mine) == sorted((p[0], p[1], p[3], p[4], p[5], p[6], p[7], p[8]) for p in parts)
        # boundaries of the room, in the merged list order, are the room's own list
        lo, hi = min(q[3] for q in parts) + shift, max(q[3] + q[6] for q in parts) + shift
        mb = [(t[0], t[3] - shift, t[4] - shift, t[5], t[6], t[7], t[8]) for t in mbnds if lo <= t[3] and t[4] <= hi]
        assert mb == [(t[0], t[3], t[4], t[5], t[6], t[7], t[8]) for t in bnds]
    assert seen_parts == len(mparts)
    b.close()
