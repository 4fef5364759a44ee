"""The device's scalar routines against the host C library, on the GPU:
the glibc-identical float functions on the BP/CNM domains, double sqrt on
every integer the motion costs can take, and the double exp reconstruction
uses (dvc_exp), bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIBM = ctypes.CDLL("libm.so.6")


def host(fn, x):
    f = getattr(LIBM, fn)
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float]
    return np.array([f(float(v)) for v in x], np.float32)


def _bits_sample(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    b = rng.integers(lo, hi, n, dtype=np.int64).astype(np.uint32)
    return b.view(np.float32)


@pytest.mark.parametrize("func,name,lo,hi", [(0, "logf", 0x00000001, 0x7f800000), (1, "expf", 0x00000000, 0xffffffff),
                                             (2, "tanhf", 0x00000000, 0xffffffff)])
def test_float_routines_match_host(ctx, func, name, lo, hi):
    x = _bits_sample(lo, hi, 200000, func)
    x = x[np.isfinite(x)]
    dev = ctx.selftest(func, x)
    ref = host(name, x)
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("func", [3, 6, 7])
def test_bp_phi_full_grid_of_messages(ctx, func):
    # every float in [0, 64) on a stride, plus the small magnitudes bit by bit;
    # func 3 is the branching form, 6 the branch-free one, 7 its shortened form (the BP kernel's)
    b = np.concatenate([np.arange(0, 0x42800000, 4099, dtype=np.uint32), np.arange(0, 1 << 20, dtype=np.uint32),
                        np.arange(0x42800000, 0x7f800001, 65537, dtype=np.uint32)])
    x = b.view(np.float32)
    dev = ctx.selftest(func, x)
    logf = LIBM.logf; logf.restype = ctypes.c_float; logf.argtypes = [ctypes.c_float]
    tanhf = LIBM.tanhf; tanhf.restype = ctypes.c_float; tanhf.argtypes = [ctypes.c_float]
    eps = np.float32(0.000000001)
    ref = np.array([-logf(tanhf(float((np.float32(v) + eps) / np.float32(2))))
                    for v in x], np.float32)
    assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))


def test_double_sqrt_exact_on_integer_costs(ctx):
    # motion-vector lengths: x^2 + y^2 for half-pel components up to 2*352
    n = np.arange(0, 2 * 704 * 704 + 1, dtype=np.int64)
    n = n[n < 2 ** 31].astype(np.int32)
    dev = ctx.selftest(4, n)
    assert np.array_equal(dev, np.sqrt(n.astype(np.float64)))


def test_double_exp_equals_host_libm(ctx):
    """dvc_exp on the device (reconstruction's exp) against the host C
    library's exp, bit for bit: uniform arguments over [-700, 700] and
    [-20, 20], and reconstruction's own forms +-alpha*k (alpha a float,
    k an integer step or distance)."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.exp.restype = ctypes.c_double
    libm.exp.argtypes = [ctypes.c_double]
    rng = np.random.default_rng(5)
    alpha = rng.uniform(1e-4, 64.0, 100000).astype(np.float32).astype(np.float64)
    k = rng.integers(0, 4097, 100000).astype(np.float64)
    x = np.concatenate([rng.uniform(-700, 700, 100000), rng.uniform(-20, 20, 100000), alpha * k, -alpha * k,
                        np.array([0.0, -0.0, 1e-300, 512.0, -512.0, 709.7, -745.0, 1023.0, -1075.0])])
    dev = ctx.selftest(5, x)
    ref = np.array([libm.exp(float(v)) for v in x])
    bad = np.flatnonzero(dev.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, [(x[i], dev[i], ref[i]) for i in bad[:5]]


def test_bp_phi_kernel_form_equals_restatement_on_every_argument(ctx):
    """The BP kernel's phi (reciprocal-based quotients) against the direct
    restatement of the reference's expression on the device, for every
    non-negative float and -0: zero mismatches."""
    n, bad, first = ctx.phi_sweep(0x00000000, 0x7f800000)
    assert n == 0x7f800001 and bad == 0, hex(first)
    n, bad, first = ctx.phi_sweep(0x80000000, 0x80000000)
    assert n == 1 and bad == 0
