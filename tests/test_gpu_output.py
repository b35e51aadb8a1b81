"""GPU checks of the recorder cubes (the out_<id>.txt rows of
Recorder::RecordField, ARD/recorder.cpp:44-54) and of the output files.

* plain form: cubes against the oracle's pressure field at the same cells,
  every step; the cube's centre is the point sample, bit for bit;
* batched form (cubes evaluated from the modes on matrix cores): against the
  plain form's cubes, including a cube whose rows wrap in the partition's
  array under the reference's addressing;
* ard_run (the C++ mirror's main loop): source_<id>.txt, response_<id>.txt
  and out_<id>.txt byte-identical to the Python writers over the same form,
  and its final frame equal to the Python mirror's render;
* the visualisation plane (ard_render_view) against the oracle's colouring of
  the same fields.

Tolerance: rel-L2 1e-4 over all cube values (the bar for this path)."""
import os
import subprocess

import numpy as np
import pytest

import ard_files as F
import ard_oracle as O
import ard_scenes as S
import ard_sim as A

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ard-simulator_amd")
TOL = 1e-4


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cube_flat(box, r, mode):
    """Flat indices of the 1000 cube cells in the owner's array."""
    xs, ys, zs, w, h, d = box
    if mode == S.REC_LOCAL:
        f0 = ((r[2] - zs) * h + (r[1] - ys)) * w + (r[0] - xs)
    else:
        f0 = (r[2] * h + r[1]) * w + r[0]
    o = np.arange(-5, 5)
    return (f0 + (o[:, None, None] * h + o[None, :, None]) * w + o[None, None, :]).ravel()


def wrap_scene():
    """A recorder at local x 7 of a partition whose array it addresses with
    global x 27 (x0 = 20): its cube rows run past x = W - 1 into the next row."""
    boxes = [(0, 0, 0, 20, 30, 30), (20, 0, 0, 30, 30, 30)]
    return S.Scene("wrap", 0.1, 1.0 / 8000.0, 343.0, boxes, [S.Source(10, 12, 14)], [(27, 15, 15), (35, 16, 14)], 200)


@pytest.mark.parametrize("name", ["two_rooms", "wrap"])
def test_plain_cubes_match_oracle(name):
    sc = S.two_rooms() if name == "two_rooms" else wrap_scene()
    steps = 120
    print("case plain cubes", name, flush=True)
    sim = A.Simulation(sc, form=A.FORM_PLAIN, cubes=True)
    sim.update(steps)
    cubes = sim.recorder_cubes()
    tr = sim.recorders()
    o = O.from_scene(sc)
    want = np.zeros((steps, len(sc.recorders), 1000), np.float32)
    flats = [(S.recorder_owner(sc.boxes, r), cube_flat(sc.boxes[S.recorder_owner(sc.boxes, r)], r, sc.recorder_mode))
             for r in sc.recorders]
    for t in range(steps):
        o.step(1)
        for k, (part, fl) in enumerate(flats):
            want[t, k] = o.field(part).ravel()[fl]
    e = rel_l2(cubes.reshape(steps, -1, 1000), want)
    print("%s plain cubes rel-L2 vs oracle %.3e" % (name, e))
    assert e <= TOL
    # the centre cell (offsets 0, 0, 0) is the point sample
    np.testing.assert_array_equal(cubes[:, :, 5, 5, 5], tr)


@pytest.mark.parametrize("name", ["scene3", "wrap"])
def test_batched_cubes_match_plain(name):
    sc = S.scene3(steps=300) if name == "scene3" else wrap_scene()
    steps = 300 if name == "scene3" else 200
    out = {}
    for form in (A.FORM_PLAIN, A.FORM_BATCHED):
        print("case cubes", name, "form", form, flush=True)
        sim = A.Simulation(sc, form=form, cubes=True)
        sim.update(steps)
        out[form] = (sim.recorder_cubes(), sim.recorders())
        sim.close()
    e = rel_l2(out[A.FORM_BATCHED][0], out[A.FORM_PLAIN][0])
    ec = rel_l2(out[A.FORM_BATCHED][0][:, :, 5, 5, 5], out[A.FORM_BATCHED][1])
    print("%s batched vs plain cubes rel-L2 %.3e; batched centre vs point %.3e" % (name, e, ec))
    assert e <= TOL and ec <= TOL


def test_cubes_across_trace_growth():
    """Cube rows survive the trace buffer's growth past its first 1024 steps."""
    sc = S.two_rooms()
    sim = A.Simulation(sc, form=A.FORM_BATCHED, cubes=True)
    sim.update(1000)
    head = sim.recorder_cubes(0, 50)
    sim.update(100)
    assert sim.time_step == 1100
    np.testing.assert_array_equal(sim.recorder_cubes(0, 50), head)
    np.testing.assert_array_equal(sim.recorder_cubes()[:, :, 5, 5, 5].shape, (1100, 1))


@pytest.mark.parametrize("form", [A.FORM_PLAIN, A.FORM_BATCHED])
def test_ard_run_files_match_python_writers(tmp_path, form):
    sc = S.scene3(steps=80)
    steps = 80
    exe = os.path.join(PKG, "ard_run")
    S.write_cell_files(sc, str(tmp_path / "scene"))
    root = tmp_path / "output"
    root.mkdir()
    dur = steps * np.float32(sc.dt)
    cmd = [exe, "--cells", "--c0", repr(sc.c0), "--duration", repr(float(dur)), "--frame", str(tmp_path / "f.ppm")]
    if form == A.FORM_PLAIN:
        cmd.append("--plain")
    cmd += [str(tmp_path / "scene" / n) for n in ("scene.txt", "sources.txt", "recorders.txt")]
    cmd += [str(steps), repr(sc.h), str(root)]
    print("case ard_run form", form, " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    d = F.output_dir(str(root), sc.h, sc.absorption)
    n_rows = int(np.float32(dur) / np.float32(sc.dt))
    assert n_rows in (steps - 1, steps)
    sim = A.Simulation(sc, form=form, cubes=True)
    sim.update(steps)
    mine = tmp_path / "mine"
    F.write_recorder_files(sim, str(mine), t0=0, nt=n_rows)
    md = F.output_dir(str(mine), sc.h, sc.absorption)
    for k in range(len(sc.recorders)):
        for f in ("response_%d.txt" % k, "out_%d.txt" % k):
            a = open(os.path.join(d, f), "rb").read()
            b = open(os.path.join(md, f), "rb").read()
            assert a == b, f
    # the final frame (PPM of the xy view at the source's z) against the Python mirror's
    ppm = (tmp_path / "f.ppm").read_bytes()
    px = sim.render_view(0)
    header = b"P6\n%d %d\n255\n" % (px.shape[1], px.shape[0])
    rgb = np.stack([(px >> 16) & 0xFF, (px >> 8) & 0xFF, px & 0xFF], axis=-1).astype(np.uint8)
    assert ppm == header + rgb.tobytes()
    n_src = sum(1 for t in range(steps + 2) if np.float32(t) < np.float32(dur) / np.float32(sc.dt))
    F.write_source(str(tmp_path / "src.txt"), O.gaussian(sc.h, sc.dt, sc.c0, n_src).astype(np.float32))
    assert open(os.path.join(d, "source_0.txt"), "rb").read() == (tmp_path / "src.txt").read_bytes()


# --- visualisation plane (ard_render_view)


def _oracle_views(sim, views):
    """The oracle's render of the simulation's own fields, views drawn in
    order into one persistent array, as the device does."""
    parts = sim.partitions()
    fields = [sim.pressure_field(k) for k in range(len(parts))]
    pix, out = None, []
    for view, coord in views:
        pix, (cols, rows) = O.render(parts, fields, view, coord, pixels=pix)
        out.append(pix[: cols * rows].reshape(rows, cols).copy())
    return out


def _channel_diff(a, b):
    d = 0
    for s in (16, 8, 0):
        d = np.maximum(d, np.abs(((a >> s) & 0xFF).astype(int) - ((b >> s) & 0xFF).astype(int)))
    return d


@pytest.mark.parametrize("name", ["scene3", "wrap"])
@pytest.mark.parametrize("form", [A.FORM_PLAIN, A.FORM_BATCHED])
def test_render_matches_oracle(name, form):
    """xy view at the source's z, then the yz view at the source's x (one
    pixel array, as the reference's), then an xy view at z = 0 where the PML
    slabs below the scene are drawn over the boxes. Plain form: every pixel
    equal to the oracle's colouring of the same fields. Batched form (planes
    evaluated from the modes, the fields by the full inverse transform):
    channels within 1 of it, at under 0.5 % of the pixels."""
    sc = S.scene3(steps=150) if name == "scene3" else wrap_scene()
    print("case render", name, "form", form, flush=True)
    sim = A.Simulation(sc, form=form)
    sim.update(150)
    src = sc.sources[0]
    # x = 35 in the wrap scene: the box at x0 = 20, W = 30 reads local x 35 >= W
    views = [(0, src.z), (1, src.x), (0, 0), (1, 35)]
    got = [sim.render_view(0), sim.render_view(1), sim.render_view(0, 0), sim.render_view(1, 35)]
    want = _oracle_views(sim, views)
    for (view, coord), g, w in zip(views, got, want):
        assert g.shape == w.shape
        n_white = int(np.sum(g == 0xFFFFFFFF))
        d = _channel_diff(g, w)
        frac = float(np.mean(d > 0))
        print("view %d at %d: %dx%d pixels, %d white, differing %.2e, max channel diff %d"
              % (view, coord, g.shape[1], g.shape[0], n_white, frac, int(d.max())))
        assert n_white < g.size  # the wave is visible
        if form == A.FORM_PLAIN:
            np.testing.assert_array_equal(g, w)
        else:
            assert d.max() <= 1 and frac < 5e-3
