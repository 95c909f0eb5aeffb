"""Device FP32 functions against the host, bit for bit, over a strided sweep
of all 2^32 float bit patterns plus special values."""
import numpy as np
import pytest

import oracle
from kinc_amd import _native

pytestmark = pytest.mark.gpu


def _inputs():
    sweep = (np.arange(0, 1 << 32, 241, dtype=np.uint64) & 0xFFFFFFFF).astype(np.uint32)
    x = sweep.view(np.float32)
    special = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45,
                        1.17549435e-38, 88.7, 88.72284, 89.0, -103.9, -104.0, -87.33655],
                       np.float32)
    rng = np.random.default_rng(0)
    em = rng.uniform(-110, 0.5, 2_000_000).astype(np.float32)  # EM exp arguments
    return np.concatenate([x, special, em])


@pytest.mark.parametrize("fn,name", [(0, "exp"), (1, "log"), (2, "sqrt"), (3, "recip")])
def test_device_math_bit_exact(fn, name):
    x = _inputs()
    print("math", name, "inputs", x.size, flush=True)
    dev = _native.device_math(fn, x)
    host = oracle.math_array(fn, x)
    same = (dev.view(np.uint32) == host.view(np.uint32)) | (np.isnan(dev) & np.isnan(host))
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, "%s differs at %d inputs, e.g. x=%r dev=%r host=%r" % (
        name, bad.size, x[bad[:3]], dev[bad[:3]], host[bad[:3]])


def test_device_log_ge1_equals_log():
    """The E step's log of a log-sum-exp sum (at least 1, or NaN) leaves out
    the special-value cases; for such inputs it must give the general log's
    bits (NaN for NaN)."""
    rng = np.random.default_rng(1)
    x = np.concatenate([
        np.linspace(1.0, 8.0, 3_000_001, dtype=np.float32),
        rng.uniform(1.0, 8.0, 1_000_000).astype(np.float32),
        np.nextafter(np.float32(1.0), np.float32(2.0)) + np.zeros(1, np.float32),
        np.array([1.0, 2.0, 4.0, 8.0, np.nan], np.float32)])
    dev = _native.device_math(4, x)
    host = oracle.math_array(1, x)
    same = (dev.view(np.uint32) == host.view(np.uint32)) | (np.isnan(dev) & np.isnan(host))
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, "log_ge1 differs at %d inputs, e.g. x=%r dev=%r host=%r" % (
        bad.size, x[bad[:3]], dev[bad[:3]], host[bad[:3]])


def test_device_exp_le0_equals_exp():
    """The E step's exponentials of v - max(v) (at most 0, or NaN) clamp from
    below only; for such inputs they must give the general exp's bits."""
    rng = np.random.default_rng(2)
    x = np.concatenate([
        -np.linspace(0.0, 120.0, 3_000_001, dtype=np.float32),
        rng.uniform(-110.0, 0.0, 1_000_000).astype(np.float32),
        -(np.arange(1, 1 << 16, dtype=np.uint32)).view(np.float32),   # tiny negatives
        np.array([0.0, -0.0, -103.9, -104.0, -104.1, -87.33655, -np.inf, np.nan], np.float32)])
    dev = _native.device_math(5, x)
    host = oracle.math_array(0, x)
    same = (dev.view(np.uint32) == host.view(np.uint32)) | (np.isnan(dev) & np.isnan(host))
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, "exp_le0 differs at %d inputs, e.g. x=%r dev=%r host=%r" % (
        bad.size, x[bad[:3]], dev[bad[:3]], host[bad[:3]])
