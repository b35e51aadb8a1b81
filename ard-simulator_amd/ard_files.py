"""The reference's text output files (ARD/output/README.md), written from the
device traces.

Per run the reference writes into ``./output/<h>_<absorption>/`` (folder name
from ``std::to_string`` of the two floats, ARD/recorder.cpp:12-16):

* ``source_<id>.txt``: one Gaussian sample per step of the duration
  (ARD/sound_source.cpp:50-57);
* ``response_<id>.txt``: the recorder's point sample per step
  (ARD/recorder.cpp:55);
* ``out_<id>.txt``: per step one row of the 10 x 10 x 10 cells around the
  recorder, z offset outermost and x fastest, every value followed by a space
  (ARD/recorder.cpp:44-54).

Values are formatted as ``std::ostream << float`` does by default (printf
``%g``, six significant digits), so files from these writers and from the C++
mirror (``Recorder::RecordField``) are byte-identical for identical values.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np


def format_value(v) -> str:
    """``operator<<(std::ostream&, float)`` with default flags: ``%g`` of the
    value as a double; a NaN with its sign bit set prints as ``-nan`` (glibc)."""
    v = float(np.float32(v))
    if math.isnan(v):
        return "-nan" if struct.pack(">d", v)[0] & 0x80 else "nan"
    return "%g" % v


def output_dir(root, h, absorption):
    """``root/<to_string(h)>_<to_string(absorption)>`` (ARD/recorder.cpp:12-13)."""
    return os.path.join(root, "%f_%f" % (float(np.float32(h)), float(np.float32(absorption))))


def gaussian_samples(h, dt, c0, duration):
    """GaussianSource::SampleValue(t) (ARD/gaussian_source.cpp:16-20) for every
    t with (float)t < duration / dt (ARD/sound_source.cpp:52), in FP32 with
    numpy's pow and exp (the C++ mirror uses the C library's powf and expf)."""
    f = np.float32
    n_f = f(duration) / f(dt)
    n = int(math.ceil(float(n_f)))
    t = np.arange(n, dtype=np.float32)
    t = t[t < n_f]
    courant = f(c0) * f(dt) / f(h)
    arg = np.power(f(np.pi) * ((f(2) * courant * t) / f(6) - f(2)), f(2))
    return (f(1e9) * np.exp(-arg)).astype(np.float32)


def _lines(rows):
    return "".join(r + "\n" for r in rows)


def write_source(path, samples):
    with open(path, "w") as f:
        f.write(_lines(format_value(v) for v in np.asarray(samples, np.float32).ravel()))


def write_response(path, trace):
    """One recorder's point samples, one per line."""
    write_source(path, trace)


def write_cube(path, cubes):
    """One recorder's cube rows: ``cubes`` is [nt][1000] or [nt][10][10][10]."""
    c = np.asarray(cubes, np.float32).reshape(-1, 1000)
    with open(path, "w") as f:
        for row in c:
            f.write("".join(format_value(v) + " " for v in row) + "\n")


def write_recorder_files(sim, root, first_id=0, t0=0, nt=None, cubes=True):
    """response_<id>.txt (and out_<id>.txt if the simulation samples cubes) for
    every recorder of an ard_sim.Simulation, ids first_id, first_id + 1, ...
    Returns the folder written."""
    d = output_dir(root, sim.scene.h, sim.scene.absorption)
    os.makedirs(d, exist_ok=True)
    tr = sim.recorders(t0, nt)
    cb = sim.recorder_cubes(t0, nt) if cubes else None
    for r in range(sim.nrec):
        write_response(os.path.join(d, "response_%d.txt" % (first_id + r)), tr[:, r])
        if cb is not None:
            write_cube(os.path.join(d, "out_%d.txt" % (first_id + r)), cb[:, r])
    return d
