"""The 16-bit error bounds the GPU tests use (conftest.storage_tol) rest on
round_storage; pin it against independent conversions and against the
rounding error that a p-bit significand fixes for Gaussian data."""
import numpy as np
import pytest

from conftest import ROUNDING_FACTOR_16, rel_l2, round_storage, storage_tol


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.random.default_rng(1).standard_normal(1 << 16) * np.exp(np.random.default_rng(2).uniform(-20, 20, 1 << 16))
    want = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(round_storage(x, "bf16"), want)


def test_f16_and_f32_rounding_match_numpy_casts():
    x = np.random.default_rng(3).standard_normal(1 << 16)
    assert np.array_equal(round_storage(x, "f16"), x.astype(np.float16).astype(np.float64))
    assert np.array_equal(round_storage(x, "f32"), x.astype(np.float32).astype(np.float64))
    assert np.array_equal(round_storage(x, "f64"), x)


@pytest.mark.parametrize("prec,bits", [("f16", 11), ("bf16", 8)])
def test_rounding_error_of_gaussian_data(prec, bits):
    """Rounding N(0,1) data to a p-bit significand leaves a relative L2 error
    of about 0.43 x 2^-p (an error uniform over half an ulp either side,
    averaged over the binades the data spans); a wrong significand width
    would be off by a factor 2 per bit."""
    x = np.random.default_rng(4).standard_normal(1 << 20)
    e = rel_l2(round_storage(x, prec), x)
    assert 0.3 * 2.0 ** -bits < e < 0.55 * 2.0 ** -bits
    assert storage_tol(prec, x) == pytest.approx(ROUNDING_FACTOR_16 * e)
