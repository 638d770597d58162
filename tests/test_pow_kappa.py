"""pow_kappa, the device's x**kappa for pressures (atm_common.hpp), through
its host build (lib/libatmos_column_host.so).  The reference forms
p**kappa with the libm pow (mod_sw_dyn.f90:99-101, mod_remap.f90:48-70);
pow_kappa is pinned against the exact power (decimal arithmetic at 40
digits, correctly rounded to double) and against the libm pow the oracle
calls: within 0.7 ulp of the exact value, never more than one ulp from
libm, and the grouped form of the column scans gives the same doubles."""
import ctypes
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ATM_COLUMN_HOST_LIB names another build (the sanitizer build)
LIB = os.environ.get("ATM_COLUMN_HOST_LIB") or os.path.join(ROOT, "atmosphere-model_amd", "lib", "libatmos_column_host.so")
KAPPA = float(np.float32(287.052874)) / 1005.0  # R_dry (binary32 literal) / cp_dry


@pytest.fixture(scope="module")
def powk():
    if not os.path.exists(LIB):
        pytest.skip("host build of the column numerics not built")
    lib = ctypes.CDLL(LIB)
    f = lib.atmc_pow_kappa
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int]

    def run(x, width=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        f(x.ctypes.data, y.ctypes.data, len(x), width)
        return y
    return run


def _exact(xs):
    getcontext().prec = 40
    k = Decimal(KAPPA)
    return [(k * Decimal(float(v)).ln()).exp() for v in xs]


def _libm_pow():
    libm = ctypes.CDLL("libm.so.6")
    libm.pow.restype = ctypes.c_double
    libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    return lambda v: libm.pow(float(v), KAPPA)


def test_accuracy_against_exact_power(powk):
    rng = np.random.default_rng(7)
    x = np.exp(rng.uniform(math.log(0.3), math.log(3e5), 3000))
    edges = [0.69405305, 1.0, 1.0 + 2 ** -52, 1.0 - 2 ** -53, 2 ** -0.5, 2 ** 0.5, 2.0, 1000.0, 1e5, 1.1e5,
             110000.0 + 2 ** -36, 4.0, 65536.0]
    x = np.concatenate([x, edges])
    y = powk(x)
    ex = _exact(x)
    err = [abs((Decimal(float(a)) - e) / Decimal(math.ulp(float(e)))) for a, e in zip(y, ex)]
    assert max(err) <= Decimal("0.7"), max(err)
    # correctly rounded in nearly every case
    cr = np.array([float(e) for e in ex])
    assert np.mean(y != cr) < 0.05


def test_within_one_ulp_of_libm(powk):
    rng = np.random.default_rng(8)
    x = np.exp(rng.uniform(math.log(0.5), math.log(2e5), 20000))
    y = powk(x)
    p = _libm_pow()
    ref = np.array([p(v) for v in x])
    d = np.abs(y - ref) / np.spacing(ref)
    assert d.max() <= 1.0
    assert np.mean(d > 0) < 0.06


def test_grouped_form_is_identical(powk):
    rng = np.random.default_rng(9)
    x = np.exp(rng.uniform(math.log(0.3), math.log(3e5), 4096))
    assert np.array_equal(powk(x, 1), powk(x, 4))


def test_special_values(powk):
    y = powk(np.array([0.0, np.inf, -1.0, np.nan, 5e-324, 2.2250738585072014e-308, 1e300]))
    assert y[0] == 0.0 and y[1] == np.inf and np.isnan(y[2]) and np.isnan(y[3])
    assert np.all(np.isfinite(y[4:])) and np.all(y[4:] > 0)
    # the subnormal and huge arguments against libm
    p = _libm_pow()
    for v, r in zip([5e-324, 2.2250738585072014e-308, 1e300], y[4:]):
        assert abs(r - p(v)) <= math.ulp(p(v))
