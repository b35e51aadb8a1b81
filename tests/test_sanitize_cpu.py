"""AddressSanitizer and UBSan over the CPU-side code:

* the library's host-side scene set-up (topology, interface descriptors,
  last-writer masks: ard_topology.cpp, ard_masks.cpp) through the driver
  tests/sanitize/host_check.cpp on fixed, seeded and hall scenes;
* the oracle (oracle/ard_oracle.cpp), built with the sanitizers into a
  temporary folder and driven from Python (the sanitizer runtimes preloaded):
  both transform paths, tables, profiles, PML and boundary set-ups and short
  runs of small scenes in FP32 and FP64.

Any sanitizer report fails the run (-fno-sanitize-recover=all)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def test_host_scene_setup_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_check")
    subprocess.run(["g++", "-std=c++17"] + SAN + [os.path.join(ROOT, "tests", "sanitize", "host_check.cpp"),
                    os.path.join(ROOT, "ard-simulator_amd", "csrc", "ard_topology.cpp"),
                    os.path.join(ROOT, "ard-simulator_amd", "csrc", "ard_masks.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "hall_cells", "scene.txt")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-3000:]


ORACLE_RUN = r"""
import sys
sys.path[:0] = [%(oracle)r, %(pkg)r]
import numpy as np
import ard_oracle as O
import ard_scenes as S
O._LIB_PATH = %(lib)r
rng = np.random.default_rng(0)
for shape in [(3, 4, 5), (1, 1, 1), (7, 11, 13), (20, 12, 9)]:
    x = rng.standard_normal(shape)
    for inv in (False, True):
        a, b = O.dct3d(x, inv, fft=True), O.dct3d(x, inv, fft=False)
        assert np.abs(a - b).max() < 1e-9
for dbl in (False, True):
    O.mode_tables(0.1, 1 / 8000, 343.0, 5, 4, 3, double=dbl)
    O.gaussian(0.1, 1 / 8000, 343.5, 30, double=dbl)
    for t in range(6):
        O.pml_profile(0.1, 343.0, 5, t, 6, 5, 7, double=dbl)
    for sc in (S.two_rooms(), S.scene3(steps=40)):
        o = O.from_scene(sc, double=dbl, fft=True)
        o.step(25)
        for i in range(len(o.partitions())):
            o.field(i, 0)
    sc = S.two_rooms()
    sc.boxes = [(0, 0, 0, 17, 13, 15), (17, 0, 0, 11, 19, 15), (0, 13, 0, 17, 9, 15)]
    sc.sources = [S.Source(5, 1, 4)]
    sc.recorders = [(23, 9, 7)]
    sc.recorder_mode = S.REC_LOCAL
    sc.pml_layers = 3
    O.from_scene(sc, double=dbl, fft=False).step(20)
print("ORACLE OK")
"""


def test_oracle_under_asan_ubsan(tmp_path):
    lib = str(tmp_path / "libard_oracle_san.so")
    subprocess.run(["g++", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off"] + SAN +
                   [os.path.join(ROOT, "oracle", "ard_oracle.cpp"), "-o", lib], check=True)
    pre = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["g++", "-print-file-name=" + name], stdout=subprocess.PIPE, text=True).stdout.strip()
        if os.path.isabs(p) and os.path.exists(p):
            pre.append(p)
    assert pre, "sanitizer runtimes not found"
    env = dict(os.environ, LD_PRELOAD=":".join(pre), ASAN_OPTIONS="detect_leaks=0", OMP_NUM_THREADS="2")
    code = ORACLE_RUN % {"oracle": os.path.join(ROOT, "oracle"), "pkg": os.path.join(ROOT, "ard-simulator_amd"),
                         "lib": lib}
    r = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=env, timeout=900)
    assert r.returncode == 0 and "ORACLE OK" in r.stdout, r.stdout[-4000:]
