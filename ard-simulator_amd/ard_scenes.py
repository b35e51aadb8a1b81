"""Scene files and the seeded workloads the solver is measured and checked on.

The scene format is the reference's: one box per line, ``x y z w h d`` in
integer metres (ARD/partition.cpp:151-176); sources and recorders are
``x y z`` in integer metres (ARD/sound_source.cpp:24-48,
ARD/recorder.cpp:67-89). Metres become cells as ``(int)(v / h)`` in FP32,
exactly as the reference converts them.

Workloads (all cell coordinates are global):
  * ``two_rooms``   -- two 3 m cubes side by side, h = 0.1 m, 1000 steps.
  * ``scene3``      -- the five-box 2.5D room of assets/scene-3.txt, h = 0.05 m,
                       4 seeded recorders, 4000 steps.
  * ``hall``        -- an L-shaped 40 x 25 x 10 m hall cut into 20 boxes by
                       seeded guillotine cuts, h = 0.05 m, 8 sources,
                       16 recorders, 2000 steps.
  * ``diff_room``   -- a 10 x 8 x 4 m room in 3 boxes driven by a seeded
                       band-limited noise waveform (gradient checks).
  * ``rooms64``     -- 64 independent boxes with seeded sides, 1-4 boxes each.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional, Tuple

import numpy as np

# (h, dt) pairs hard-coded by the reference (ARD/main.cpp:32-42).
DT_FOR_H = {0.05: 0.625e-4, 0.1: 1.0 / 8000.0, 0.2: 2e-4}

# Recorder addressing. "reference": the recorder's global cell coordinates
# index the owning partition's array directly (ARD/recorder.cpp:44-55 passes
# global x,y,z to get_pressure, which does not subtract the partition origin,
# ARD/dct_volume.cpp:70-73). "local": the physical cell at those coordinates.
REC_REFERENCE = 0
REC_LOCAL = 1

SCENE3_TEXT = """0 0 0 3 3 2
0 3 0 2 1 2
3 0 0 1 2 2
4 0 0 1 1 2
0 4 0 1 1 2
"""


@dataclasses.dataclass
class Source:
    x: int
    y: int
    z: int
    wave: Optional[np.ndarray] = None  # None = Gaussian pulse


@dataclasses.dataclass
class Scene:
    name: str
    h: float
    dt: float
    c0: float
    boxes: List[Tuple[int, int, int, int, int, int]]  # cells: xs ys zs w h d
    sources: List[Source]
    recorders: List[Tuple[int, int, int]]
    steps: int
    recorder_mode: int = REC_REFERENCE
    pml_layers: int = 5
    absorption: float = 1.0
    room_of_box: Optional[List[int]] = None  # independent-room id per box (batches)

    def dct_cells(self):
        return sum(b[3] * b[4] * b[5] for b in self.boxes)


def m2cell(v, h):
    """(int)(v / m_dh) in FP32 (ARD/partition.cpp:165-172)."""
    return int(np.float32(v) / np.float32(h))


def parse_boxes(text, h):
    boxes = []
    vals = [int(t) for t in text.split()]
    for i in range(0, len(vals) - len(vals) % 6, 6):
        boxes.append(tuple(m2cell(v, h) for v in vals[i : i + 6]))
    return boxes


def parse_points(text, h):
    vals = [int(t) for t in text.split()]
    return [tuple(m2cell(v, h) for v in vals[i : i + 3]) for i in range(0, len(vals) - len(vals) % 3, 3)]


def load_scene_file(path, h):
    with open(path) as f:
        return parse_boxes(f.read(), h)


def recorder_owner(boxes, r):
    """Recorder::FindPartition (ARD/recorder.cpp:26-38): first box with a 5-cell margin."""
    x, y, z = r
    for i, (xs, ys, zs, w, hh, d) in enumerate(boxes):
        if xs < x - 5 and xs + w > x + 4 and ys < y - 5 and ys + hh > y + 4 and zs < z - 5 and zs + d > z + 4:
            return i
    return -1


def recorder_valid(boxes, r, mode):
    i = recorder_owner(boxes, r)
    if i < 0:
        return False
    if mode == REC_REFERENCE:
        xs, ys, zs, w, hh, d = boxes[i]
        flat = r[2] * hh * w + r[1] * w + r[0]
        return flat < w * hh * d
    return True


def two_rooms():
    h = 0.1
    boxes = [(0, 0, 0, 30, 30, 30), (30, 0, 0, 30, 30, 30)]
    # source at (1,1,1) m; recorder at (4,2,1.5) m given by cell index since the
    # scene format only holds integer metres.
    return Scene("two_rooms", h, 1.0 / 8000.0, 343.0, boxes, [Source(10, 10, 10)], [(40, 20, 15)], 1000)


def _seeded_recorders(boxes, n, rng, mode, exclude_shell=5):
    recs = []
    tries = 0
    while len(recs) < n:
        tries += 1
        assert tries < 100000
        i = int(rng.integers(len(boxes)))
        xs, ys, zs, w, hh, d = boxes[i]
        r = (
            int(rng.integers(xs + 6, xs + w - 5)),
            int(rng.integers(ys + 6, ys + hh - 5)),
            int(rng.integers(zs + 6, zs + d - 5)),
        )
        if recorder_valid(boxes, r, mode) and r not in recs:
            recs.append(r)
    return recs


def scene3(seed=3, steps=4000):
    h = 0.05
    boxes = parse_boxes(SCENE3_TEXT, h)
    rng = np.random.default_rng(seed)
    recs = _seeded_recorders(boxes, 4, rng, REC_REFERENCE)
    return Scene("scene3", h, DT_FOR_H[0.05], 343.0, boxes, [Source(20, 20, 20)], recs, steps)


def _guillotine(rects, n, rng, min_side):
    rects = list(rects)
    while len(rects) < n:
        order = sorted(range(len(rects)), key=lambda i: -(rects[i][2] * rects[i][3]))
        for i in order:
            x, y, w, hh = rects[i]
            axes = [a for a, L in ((0, w), (1, hh)) if L >= 2 * min_side]
            if not axes:
                continue
            a = axes[int(rng.integers(len(axes)))] if len(axes) > 1 else axes[0]
            L = w if a == 0 else hh
            c = int(rng.integers(min_side, L - min_side + 1))
            if a == 0:
                new = [(x, y, c, hh), (x + c, y, w - c, hh)]
            else:
                new = [(x, y, w, c), (x, y + c, w, hh - c)]
            rects = rects[:i] + new + rects[i + 1 :]
            break
        else:
            raise ValueError("cannot cut further")
    return rects


def hall(seed=7, steps=2000, h=0.05, n_boxes=20):
    """L-shaped 40 x 25 x 10 m hall: [0,40]x[0,20] plus [0,25]x[20,25] (925 m^2)."""
    rng = np.random.default_rng(seed)
    rects = _guillotine([(0, 0, 40, 20), (0, 20, 25, 5)], n_boxes, rng, 2)
    k = m2cell(1, h)
    depth = m2cell(10, h)
    boxes = [(x * k, y * k, 0, w * k, hh * k, depth) for (x, y, w, hh) in rects]
    srcs = []
    while len(srcs) < 8:
        i = int(rng.integers(len(boxes)))
        xs, ys, zs, w, hh, d = boxes[i]
        s = (int(rng.integers(xs + 3, xs + w - 3)), int(rng.integers(ys + 3, ys + hh - 3)), int(rng.integers(3, d - 3)))
        if s not in [(q.x, q.y, q.z) for q in srcs]:
            srcs.append(Source(*s))
    recs = _seeded_recorders(boxes, 16, rng, REC_LOCAL)
    return Scene("hall", h, DT_FOR_H[0.05], 343.0, boxes, srcs, recs, steps, recorder_mode=REC_LOCAL)


def band_limited_noise(n, seed, fs=16000.0, cutoff=1000.0, taps=63):
    """Seeded N(0,1) noise low-passed by a Hamming-windowed sinc FIR."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n + taps)
    m = np.arange(taps) - (taps - 1) / 2
    fir = 2 * cutoff / fs * np.sinc(2 * cutoff / fs * m) * np.hamming(taps)
    y = np.convolve(x, fir, mode="full")[taps : taps + n]
    return y / np.std(y)


def diff_room(seed=11, steps=4000, h=0.05, scale=1):
    """10 x 8 x 4 m room in 3 boxes; scale=4 gives the 1/4-scale variant (2.5 x 2 x 1 m)."""
    k = m2cell(1, h)
    if scale == 1:
        rects = [(0, 0, 4, 8), (4, 0, 3, 8), (7, 0, 3, 8)]
        depth = 4 * k
    else:  # quarter scale in metres is not integral; use cells directly
        k4 = k // 4
        rects = [(0, 0, 4, 8), (4, 0, 3, 8), (7, 0, 3, 8)]
        depth = 4 * k4
        k = k4
    boxes = [(x * k, y * k, 0, w * k, hh * k, depth) for (x, y, w, hh) in rects]
    rng = np.random.default_rng(seed)
    wave = band_limited_noise(steps, seed)
    xs, ys, zs, w, hh, d = boxes[0]
    src = Source(xs + w // 2, ys + hh // 3, d // 2, wave=wave)
    recs = _seeded_recorders(boxes, 8, rng, REC_LOCAL)
    return Scene("diff_room", h, DT_FOR_H[0.05], 343.0, boxes, [src], recs, steps, recorder_mode=REC_LOCAL)


def rooms64(seed=5, steps=1500, h=0.05, n_rooms=64):
    """64 independent rooms: sides U[2,6] m (integers), each cut into 1-4 boxes.

    Rooms are laid out far apart along x so that no two rooms touch; each
    box records its room id so the batched form can treat them as separate.
    """
    rng = np.random.default_rng(seed)
    k = m2cell(1, h)
    boxes, room_of, srcs, recs = [], [], [], []
    x0 = 0
    for r in range(n_rooms):
        W, H, D = (int(v) for v in rng.integers(2, 7, size=3))
        nb = int(rng.integers(1, 5))
        rects = [(0, 0, W, H)]
        while len(rects) < nb:
            try:
                rects = _guillotine(rects, len(rects) + 1, rng, 1)
            except ValueError:
                break
        rb = [((x0 + x) * k, y * k, 0, w * k, hh * k, D * k) for (x, y, w, hh) in rects]
        boxes += rb
        room_of += [r] * len(rb)
        b = rb[0]
        srcs.append(Source(b[0] + b[3] // 2, b[1] + b[4] // 2, b[5] // 2))
        recs += _seeded_recorders(rb, 1, rng, REC_LOCAL)
        x0 += W + 2  # 2 m gap: PML slabs (5 cells) never touch the next room
    return Scene("rooms64", h, DT_FOR_H[0.05], 343.0, boxes, srcs, recs, steps, recorder_mode=REC_LOCAL,
                 room_of_box=room_of)


def split_rooms(scene):
    """The independent rooms of a scene with ``room_of_box`` as separate
    scenes (boxes, sources and recorders by room), in room order."""
    rooms = sorted(set(scene.room_of_box))
    out = []
    for r in rooms:
        boxes = [b for b, q in zip(scene.boxes, scene.room_of_box) if q == r]

        def inside(pt):
            return any(b[0] <= pt[0] < b[0] + b[3] and b[1] <= pt[1] < b[1] + b[4] and b[2] <= pt[2] < b[2] + b[5]
                       for b in boxes)

        srcs = [s for s in scene.sources if inside((s.x, s.y, s.z))]
        recs = [q for q in scene.recorders if inside(q)]
        out.append(dataclasses.replace(scene, name="%s_room%d" % (scene.name, r), boxes=boxes, sources=srcs,
                                       recorders=recs, room_of_box=None))
    return out


def write_cell_files(scene, directory):
    """Writes scene.txt / sources.txt / recorders.txt in the reference's
    whitespace format (ARD/partition.cpp:151-176, ARD/sound_source.cpp:24-48,
    ARD/recorder.cpp:67-89) with values in cells, as read by
    ``ard_run --cells``. Gaussian sources only."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "scene.txt"), "w") as f:
        for b in scene.boxes:
            f.write(" ".join(str(int(v)) for v in b) + "\n")
    with open(os.path.join(directory, "sources.txt"), "w") as f:
        for s in scene.sources:
            if s.wave is not None:
                raise ValueError("waveform sources have no file form")
            f.write("%d %d %d\n" % (s.x, s.y, s.z))
    with open(os.path.join(directory, "recorders.txt"), "w") as f:
        for r in scene.recorders:
            f.write("%d %d %d\n" % tuple(r))


def read_cell_files(directory):
    """Boxes, source points and recorder points of write_cell_files' format."""
    def rows(name, n):
        with open(os.path.join(directory, name)) as f:
            v = [int(t) for t in f.read().split()]
        return [tuple(v[i : i + n]) for i in range(0, len(v) - len(v) % n, n)]

    return rows("scene.txt", 6), rows("sources.txt", 3), rows("recorders.txt", 3)
