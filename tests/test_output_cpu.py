"""CPU checks of the reference's output files (ARD/output/README.md): the text
format of the Python writers against C++ ``operator<<`` on float, the C++
mirror's source_<id>.txt against the oracle's Gaussian samples, the output
folder name, and the cube range check of ard_finalize."""
import os
import subprocess

import numpy as np
import pytest

import ard_files as F
import ard_oracle as O
import ard_scenes as S
import ard_sim as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ard-simulator_amd")

FORMAT_CPP = r"""
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>
int main(int argc, char** argv) {
  std::ifstream in(argv[1], std::ios::binary);
  std::vector<float> v;
  float x;
  while (in.read(reinterpret_cast<char*>(&x), 4)) v.push_back(x);
  std::ofstream out(argv[2]);
  for (size_t i = 0; i < v.size(); ++i) {
    out << v[i] << " ";
    if (i % 1000 == 999) out << std::endl;
  }
  return 0;
}
"""

SOURCES_CPP = r"""
#include <cstdlib>
#include <iostream>
#include "ard_host.h"
using namespace ard::host;
int main(int argc, char** argv) {
  Simulation::m_dh = float(std::atof(argv[3]));
  Simulation::m_dt = float(std::atof(argv[4]));
  Simulation::m_c0 = float(std::atof(argv[5]));
  Simulation::m_duration = float(std::atof(argv[6]));
  Simulation::m_output_root = argv[1];
  std::cout << Simulation::OutputDir() << std::endl;
  auto s = SoundSource::ImportSources(argv[2], true);
  return s.size() == 2 ? 0 : 1;
}
"""


def _compile(tmp_path, name, src, extra=()):
    cpp = tmp_path / (name + ".cpp")
    cpp.write_text(src)
    exe = tmp_path / name
    subprocess.run(["g++", "-O1", "-std=c++17", str(cpp), "-o", str(exe)] + list(extra), check=True)
    return str(exe)


def test_format_matches_ostream(tmp_path):
    rng = np.random.default_rng(0)
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 1e-5, 1e-4, 123456.0, 1234565.0, 999999.5, 1e9, 3.4e38,
                        1.4e-45, 1.17e-38, np.inf, -np.inf, np.nan, 0.1, 2.5, 1e-40], dtype=np.float32)
    mant = rng.standard_normal(4980).astype(np.float32)
    expo = rng.integers(-40, 39, 4980).astype(np.float32)
    vals = np.concatenate([special, (mant * np.float32(10.0) ** expo).astype(np.float32)])
    vals[np.isinf(vals) & (np.abs(np.arange(len(vals))) > 20)] = 1.0
    assert len(vals) == 5000
    exe = _compile(tmp_path, "fmt", FORMAT_CPP)
    (tmp_path / "v.bin").write_bytes(vals.tobytes())
    subprocess.run([exe, str(tmp_path / "v.bin"), str(tmp_path / "cpp.txt")], check=True)
    F.write_cube(str(tmp_path / "py.txt"), vals.reshape(5, 1000))
    assert (tmp_path / "py.txt").read_bytes() == (tmp_path / "cpp.txt").read_bytes()


def test_format_negative_nan():
    neg = np.frombuffer(np.uint32(0xFFC00000).tobytes(), np.float32)[0]
    assert F.format_value(neg) == "-nan"
    assert F.format_value(np.float32(np.nan)) == "nan"


def test_output_dir_name():
    assert F.output_dir("out", 0.1, 1.0) == os.path.join("out", "0.100000_1.000000")
    assert F.output_dir("out", 0.05, 0.5) == os.path.join("out", "0.050000_0.500000")


@pytest.mark.parametrize("h,dt,c0,duration", [(0.1, 1.0 / 8000.0, 343.5, 0.05), (0.05, 0.625e-4, 343.0, 0.031)])
def test_cpp_mirror_source_files(tmp_path, h, dt, c0, duration):
    """SoundSource::ImportSources writes source_<id>.txt for each source
    (ARD/sound_source.cpp:24-57): the oracle's FP32 Gaussian, one sample per
    step while (float)t < duration / dt, in the ostream format."""
    if not os.path.exists(os.path.join(PKG, "libard.so")):
        pytest.skip("libard.so not built")
    exe = _compile(tmp_path, "src", SOURCES_CPP,
                   ["-I" + os.path.join(PKG, "host"), "-L" + PKG, "-lard", "-Wl,-rpath," + PKG])
    root = tmp_path / "output"
    d = root / ("%f_%f" % (np.float32(h), np.float32(1.0)))
    d.mkdir(parents=True)
    (tmp_path / "sources.txt").write_text("10 10 10\n20 12 7\n")
    r = subprocess.run([exe, str(root) + "/", str(tmp_path / "sources.txt"), repr(h), repr(dt), repr(c0),
                        repr(duration)], check=True, stdout=subprocess.PIPE, text=True)
    assert r.stdout.strip() == str(d) + "/"
    n = int(np.ceil(np.float32(duration) / np.float32(dt)))
    n = n if np.float32(n - 1) < np.float32(duration) / np.float32(dt) else n - 1
    want = O.gaussian(h, dt, c0, n).astype(np.float32)
    F.write_source(str(tmp_path / "want.txt"), want)
    files = sorted(p.name for p in d.iterdir())
    assert len(files) == 2 and all(f.startswith("source_") for f in files)
    for f in files:
        assert (d / f).read_bytes() == (tmp_path / "want.txt").read_bytes()
    py = F.gaussian_samples(h, dt, c0, duration)
    assert len(py) == n
    np.testing.assert_allclose(py, want, rtol=2e-6, atol=1e-3)


def test_cube_outside_the_array_is_an_error():
    """With the reference's addressing a recorder whose point is inside its
    partition's array can have cube cells past its end (undefined in the
    reference): ard_finalize rejects it when cubes are on."""
    sc = S.two_rooms()
    sc.recorders = [(40, 25, 25)]  # flat 23290 of 27000; the cube reaches 27014
    sim = A.Simulation(sc, device=-1, finalize=False, cubes=True)
    with pytest.raises(A.ArdError, match="cube addresses cells outside"):
        A._check(sim.L.ard_finalize(sim.ctx))
    sim = A.Simulation(sc, device=-1, finalize=False)
    with pytest.raises(A.ArdError, match="host-only"):
        A._check(sim.L.ard_finalize(sim.ctx))
    sc.recorders = [(40, 20, 15)]
    sim = A.Simulation(sc, device=-1, finalize=False, cubes=True)
    with pytest.raises(A.ArdError, match="host-only"):
        A._check(sim.L.ard_finalize(sim.ctx))
