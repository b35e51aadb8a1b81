"""CPU checks of libard.so: it loads, exports the C-ABI of include/ard.h, and its
host-side topology (the Simulation constructor's rules) matches the oracle and
the committed topology fixtures. No compute calls: these run without a GPU."""
import json
import os
import re

import pytest

import ard_oracle as O
import ard_scenes as S
import ard_sim as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    with open(os.path.join(ROOT, "include", "ard.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ard_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = A.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(A.EXPORTED) == syms


@pytest.mark.parametrize("name", ["two_rooms", "scene3"])
def test_topology_matches_fixture(name):
    sc = getattr(S, name)()
    sim = A.Simulation(sc, device=-1, finalize=False)
    with open(os.path.join(GOLD, name + "_topology.json")) as f:
        topo = json.load(f)
    assert [list(p) for p in sim.partitions()] == topo["partitions"]
    assert [list(b) for b in sim.boundaries()] == topo["boundaries"]


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_hall_topology_matches_oracle(seed):
    sc = S.hall(seed=seed)
    sim = A.Simulation(sc, device=-1, finalize=False)
    o = O.from_scene(sc)
    assert sim.partitions() == o.partitions()
    assert sim.boundaries() == o.boundaries()


def test_recorder_outside_margin_is_an_error():
    sc = S.two_rooms()
    sc.recorders = [(2, 20, 15)]  # 2 cells from the x = 0 face: no owner (ARD/recorder.cpp:26-38)
    sim = A.Simulation(sc, device=-1, finalize=False)
    with pytest.raises(A.ArdError, match="not 5 cells inside"):
        A._check(sim.L.ard_finalize(sim.ctx))


def test_reference_addressing_out_of_array_is_an_error():
    sc = S.two_rooms()
    # a 20^3 box at y = 200: global coordinates (10, 210, 14) address
    # 14*400 + 210*20 + 10 = 9810 >= 8000 cells of its array
    sc.boxes = [(0, 200, 0, 20, 20, 20)]
    sc.sources = [S.Source(10, 205, 10)]
    sc.recorders = [(10, 210, 14)]
    sim = A.Simulation(sc, device=-1, finalize=False)
    with pytest.raises(A.ArdError, match="outside its partition"):
        A._check(sim.L.ard_finalize(sim.ctx))


def test_bad_arguments_rejected():
    sc = S.two_rooms()
    sim = A.Simulation(sc, device=-1, finalize=False)
    with pytest.raises(A.ArdError):
        A._check(sim.L.ard_add_dct_partition(sim.ctx, 0, 0, 0, 30, 30, 30))  # topology already built
    with pytest.raises(A.ArdError):
        A._check(sim.L.ard_step(sim.ctx, 1))  # not finalized


def test_hall_cell_fixture_matches_generator():
    """The committed hall scene files (cell units) are the seeded generator's output."""
    sc = S.hall()
    boxes, srcs, recs = S.read_cell_files(os.path.join(os.path.dirname(__file__), "golden", "hall_cells"))
    assert boxes == [tuple(b) for b in sc.boxes]
    assert srcs == [(s.x, s.y, s.z) for s in sc.sources]
    assert recs == [tuple(r) for r in sc.recorders]
    assert sum(b[3] * b[4] * b[5] for b in boxes) == 74_000_000


@pytest.mark.parametrize("boxes", [
    ((0, 0, 0, 30, 30, 30), (30, 0, 0, 30, 30, 30)),      # x neighbours, b to the right
    ((30, 0, 0, 30, 30, 30), (0, 0, 0, 30, 30, 30)),      # x neighbours, b to the left
    ((0, 0, 0, 40, 20, 16), (10, 20, 0, 50, 30, 16)),     # y neighbours, partial overlap
    ((10, 20, 0, 50, 30, 16), (0, 0, 0, 40, 20, 16)),     # y neighbours, b below
    ((0, 0, 0, 20, 20, 20), (40, 0, 0, 20, 20, 20)),      # apart
    ((0, 0, 0, 20, 20, 20), (20, 20, 0, 20, 20, 20)),     # corner contact only
])
def test_find_boundary_matches_oracle_topology(boxes):
    """ard_find_boundary (Boundary::FindBoundary, ARD/boundary.cpp:289-334, needing
    no context) against the oracle's boundary between the same two DCT boxes: the
    first boundary of a two-box oracle scene, whose partitions 0 and 1 are the boxes."""
    a, b = boxes
    o = O.OracleSim(0.1, 1.25e-4, 343.0)
    o.add_dct(*a)
    o.add_dct(*b)
    o.finalize()
    dct_pairs = [bd for bd in o.boundaries() if {bd[1], bd[2]} == {0, 1}]
    got = A.find_boundary(a, b)
    if not dct_pairs:
        assert got is None
    else:
        bd = dct_pairs[0]
        assert got == (bd[0],) + tuple(bd[3:9])
