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


# L and its folded K = ceil(L/2): odd lengths (the halved middle column, the
# mirrored 16-byte loads, the skipped middle write), K tails of at most 8
# columns after whole 16-column chunks (K = 3, 4, 17, 33, 136) and of more than 8
# (K = 9, 100), even lengths (6, the smallest partition side, and the hall's 40
# and 200)
FOLD_L = [6, 7, 9, 11, 17, 33, 40, 65, 199, 200, 271]


@pytest.mark.parametrize("entry", ["ard_selftest_fold", "ard_selftest_fold_direct"])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("b_layout", [0, 1])
@pytest.mark.parametrize("L", FOLD_L)
@pytest.mark.parametrize("N", [1, 37, 130])
def test_folded_transform_matches_numpy(entry, inverse, b_layout, L, N):
    """The even/odd folded matrix-core transforms against scipy's REDFT10 /
    REDFT01 (FFTW's definitions): the LDS-staged kernel (variant 3) and the
    one that loads each wave's operands straight into registers (variant 5),
    on odd and even lengths, K tails and partial tiles."""
    import scipy.fft

    rng = np.random.default_rng(L * 131 + N * 7 + 2 * inverse + b_layout)
    x = rng.standard_normal((L, N)).astype(np.float32)
    ref = scipy.fft.dct(x.astype(np.float64), type=3 if inverse else 2, axis=0, norm=None)
    xin = x if b_layout == 0 else np.ascontiguousarray(x.T)
    y = A.selftest_fold(xin, inverse=inverse, b_layout=b_layout, entry=entry)
    if b_layout == 1:
        y = y.T
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


def test_folded_transform_refuses_lengths_below_six():
    with pytest.raises(A.ArdError, match="error -1"):
        A.selftest_fold(np.ones((5, 4), np.float32))


@pytest.mark.parametrize("shape", [(40, 60, 200), (7, 13, 9), (1, 1, 1), (30, 30, 30), (64, 3, 129)])
@pytest.mark.parametrize("dct_type", [2, 3])
def test_dct3d_matches_scipy(shape, dct_type):
    """ard_dct3d (the VkFFT_DCT seam plus DctVolume's scaling) on device
    volumes against scipy's dctn, with the reference's scales."""
    import scipy.fft
    import torch

    rng = np.random.default_rng(sum(shape) + dct_type)
    x = rng.standard_normal(shape).astype(np.float32)
    n = x.size
    scale = 1 / (2 * np.sqrt(2 * n)) if dct_type == 2 else 1 / np.sqrt(2 * n)
    ref = scipy.fft.dctn(x.astype(np.float64), type=dct_type, norm=None) * scale
    xi = torch.tensor(x, device="cuda")
    out = torch.empty_like(xi)
    A.dct3d(xi, out, dct_type=dct_type, scale=scale)
    got = out.cpu().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5
    assert np.array_equal(xi.cpu().numpy(), x)  # input untouched


def test_dct3d_refuses_overlap_and_bad_type():
    import torch

    x = torch.zeros((4, 5, 6), device="cuda")
    with pytest.raises(A.ArdError, match="error -1"):
        A.dct3d(x, x)
    with pytest.raises(A.ArdError, match="error -1"):
        A.dct3d(x, torch.empty_like(x), dct_type=1)


FFT_L = [2, 3, 5, 6, 7, 13, 20, 40, 60, 98, 121, 140, 169, 200, 220, 260, 540, 1024]


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("b_layout", [0, 1])
@pytest.mark.parametrize("L", FFT_L)
@pytest.mark.parametrize("N", [1, 37, 130])
def test_fft_transform_matches_numpy(inverse, b_layout, L, N):
    """The FFT-style transform (variant 4: Makhoul's reordering, two real lines
    per complex FFT, mixed-radix stages 4/2/3/5/7/11/13 in LDS) against scipy's
    REDFT10 / REDFT01, on every radix, odd lengths and partial tiles."""
    import scipy.fft

    rng = np.random.default_rng(L * 17 + N + 2 * inverse + b_layout)
    x = rng.standard_normal((L, N)).astype(np.float32)
    ref = scipy.fft.dct(x.astype(np.float64), type=3 if inverse else 2, axis=0, norm=None)
    xin = x if b_layout == 0 else np.ascontiguousarray(x.T)
    y = A.selftest_fold(xin, inverse=inverse, b_layout=b_layout, entry="ard_selftest_fft")
    if b_layout == 1:
        y = y.T
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err


def test_fft_transform_refuses_large_prime_factors():
    with pytest.raises(A.ArdError, match="error -5"):
        A.selftest_fold(np.ones((17, 4), np.float32), entry="ard_selftest_fft")


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("b_layout", [0, 1])
@pytest.mark.parametrize("L", FOLD_L)
@pytest.mark.parametrize("N", [1, 37, 130])
def test_folded_bf16_transform_matches_numpy(inverse, b_layout, L, N):
    """The folded transform in BF16 split precision (variant 2's face passes:
    three BF16 products per 16 folded columns on v_mfma_f32_32x32x16_bf16,
    FP32 accumulate) against scipy's REDFT10 / REDFT01; a product carries
    about 2^-16 relative error."""
    import scipy.fft

    rng = np.random.default_rng(L * 131 + N * 7 + 2 * inverse + b_layout + 1)
    x = rng.standard_normal((L, N)).astype(np.float32)
    ref = scipy.fft.dct(x.astype(np.float64), type=3 if inverse else 2, axis=0, norm=None)
    xin = x if b_layout == 0 else np.ascontiguousarray(x.T)
    y = A.selftest_fold(xin, inverse=inverse, b_layout=b_layout, entry="ard_selftest_fold_bf16")
    if b_layout == 1:
        y = y.T
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err < 1e-4, err


@pytest.mark.parametrize("shape", [(64, 200, 300), (130, 270, 77), (40, 40, 1000)])
def test_vector_and_matrix_core_fp32_are_bit_identical(shape):
    """The FP32 vector-FMA kernel and the FP32 matrix-core kernel accumulate
    each output's products in K order with one rounding per multiply-add: the
    same bits, as the recorder traces of the precision study show."""
    M, K, N = shape
    rng = np.random.default_rng(M + K + N)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    c0 = A.selftest_transform(a, b, variant=0)
    c1 = A.selftest_transform(a, b, variant=1)
    assert np.array_equal(c0, c1)
    assert np.abs(c1 - a.astype(np.float64) @ b).max() <= 1e-5 * np.abs(a.astype(np.float64) @ b).max()
