"""The dense transform kernels (vector FMA, matrix-core FP32 and matrix-core
BF16 split precision) against numpy, including ragged sizes that leave
partial 64x64 tiles."""
import numpy as np
import pytest

import ard_sim as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("b_layout", [0, 1])
@pytest.mark.parametrize("shape", [(64, 64, 16), (40, 1200, 40), (200, 300, 200), (37, 129, 53), (540, 64, 540)])
def test_transform_kernel_matches_numpy(variant, b_layout, shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N + K)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    bb = b if b_layout == 0 else np.ascontiguousarray(b.T)
    c = A.selftest_transform(a, bb, variant=variant, b_layout=b_layout)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    err = np.abs(c - ref).max() / np.abs(ref).max()
    # BF16 split: products to about 2^-16 relative (three BF16 products, FP32 sums)
    assert err < (1e-5 if variant < 2 else 1e-4), err
