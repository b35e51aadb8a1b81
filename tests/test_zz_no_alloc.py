"""ard_dct3d_noalloc, the transform entry that allocates nothing (its
workspace is a static device array), through the C-ABI against ard_dct3d,
the allocating entry it varies: bit-identical results (same kernels, same
order) on two calls in a row, the second on a smaller volume, and a volume too
large for the workspace refused with ARD_ERR_RANGE."""
import ctypes

import numpy as np
import pytest
import scipy.fft

import ard_sim as A

pytestmark = pytest.mark.gpu


def _lib():
    L = A.lib()
    vp = ctypes.c_void_p
    L.ard_dct3d_noalloc.argtypes = [ctypes.c_int] * 5 + [vp, vp, ctypes.c_float, vp]
    return L


def _run(entry, x, dct_type, scale, stream):
    import torch

    d, h, w = x.shape
    out = torch.full_like(x, float("nan"))
    r = getattr(_lib(), entry)(0, dct_type, w, h, d, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                               ctypes.c_float(scale), ctypes.c_void_p(stream))
    assert r == 0, A.lib().ard_last_error().decode()
    return out


def test_noalloc_matches_allocating_entry_twice():
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    for shape, dct_type in (((40, 60, 200), 2), ((7, 13, 9), 3)):  # second call: smaller volume
        rng = np.random.default_rng(shape[0] + dct_type)
        x = torch.tensor(rng.standard_normal(shape).astype(np.float32), device="cuda")
        n = x.numel()
        scale = 1 / (2 * np.sqrt(2 * n)) if dct_type == 2 else 1 / np.sqrt(2 * n)
        ref = _run("ard_dct3d", x, dct_type, scale, stream)
        free0 = torch.cuda.mem_get_info()[0]
        got = _run("ard_dct3d_noalloc", x, dct_type, scale, stream)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert abs(free1 - free0) < (4 << 20), (free0, free1)  # nothing allocated by the call
        g, rf = got.cpu().numpy(), ref.cpu().numpy()
        assert np.array_equal(g, rf), np.abs(g - rf).max()
        exact = scipy.fft.dctn(x.cpu().numpy().astype(np.float64), type=dct_type, norm=None) * scale
        assert np.abs(g - exact).max() / np.abs(exact).max() < 1e-5


def test_noalloc_refuses_volumes_beyond_its_workspace():
    import torch

    # 256^3 cells need 64 MiB for the work volume alone, plus the tables
    x = torch.empty((256, 256, 256), device="cuda")
    out = torch.empty_like(x)
    r = _lib().ard_dct3d_noalloc(0, 2, 256, 256, 256, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_float(1.0), None)
    assert r == -4 and "workspace" in A.lib().ard_last_error().decode()
