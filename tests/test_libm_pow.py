"""The remelting velocity's pow (amca3d_amd/csrc/libm_pow.h) against glibc.

The reference's remelting growth velocity (ProblemPhysics.cpp:142-158) calls
glibc's pow(dT, b).  The device code restates glibc 2.35's pow; here the same
header, compiled for the host, is counted against the host's glibc pow on
random arguments over every branch of the algorithm and on arguments whose
exact power lies next to a rounding midpoint, where any difference in the
tables, coefficients or fused operations would show."""
import os
import subprocess
import sys

import pytest

import pow_check

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "amca3d_amd", "csrc")


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    return pow_check.build(tmp_path_factory.mktemp("powcheck"))[0]


def test_tables_follow_their_construction():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_libm_pow_tables.py")],
                         check=True, capture_output=True, text=True).stdout
    with open(os.path.join(CSRC, "libm_pow_tables.h")) as f:
        assert f.read() == out


@pytest.mark.parametrize("y", [1.0, 2.0, 3.0])
def test_undercooling_range(check, y):
    """dT from 2^-45 K to 8192 K: the reference's exponents 1, 2, 3."""
    bad, _ = check(2_000_000, y, -45, 12, seed=int(y))
    assert bad == 0


@pytest.mark.parametrize("y", [2.0, 3.0])
def test_next_to_rounding_midpoints(check, y):
    """Arguments whose exact square or cube lies within 1e-3 ulp of a midpoint:
    the product forms miss glibc on a large share of them, the restatement on
    none."""
    bad, naive = check(200_000, y, -30, 11, seed=10 + int(y), hard=1, band=1e-3)
    assert bad == 0
    assert naive > 20_000, naive


@pytest.mark.parametrize("y,lo_e,hi_e", [
    (2.0, -1074, -1023),    # subnormal bases
    (3.0, -1074, -1023),
    (3.0, -358, -341),      # subnormal results (underflow branch, k < 0)
    (2.0, -537, -511),
    (3.0, -400, -100),      # |y log x| in [512, 1024): rescaled results
    (3.0, 150, 345),        # large results and overflow (k > 0)
    (2.5, -1074, 1023),     # any exponent, the whole double range
    (-2.75, -1074, 1023),
    (0.5, -1074, 1023),
])
def test_every_branch(check, y, lo_e, hi_e):
    bad, _ = check(300_000, y, lo_e, hi_e, seed=lo_e & 0xffff)
    assert bad == 0


@pytest.mark.parametrize("y", [1.0, 2.0, 3.0])
def test_exact_product_shortcut(check, y):
    """glibc_pow_fast, the form the device velocity calls, answers most
    undercoolings from exact products and equals glibc everywhere."""
    bad, sc = check.fast(2_000_000, y, -45, 12, seed=20 + int(y))
    assert bad == 0
    assert sc > 1_700_000, sc


@pytest.mark.parametrize("y", [2.0, 3.0])
def test_exact_product_shortcut_at_its_threshold(check, y):
    """Arguments whose exact square or cube lies within 0.09 ulp of a rounding
    midpoint, on both sides of the shortcut's 1/16-ulp threshold."""
    bad, sc = check.fast(300_000, y, -30, 11, seed=30 + int(y), hard=1, band=0.09)
    assert bad == 0
    assert 50_000 < sc < 250_000, sc


@pytest.mark.parametrize("y,lo_e,hi_e", [(1.0, -1074, 1023), (2.0, -1074, 1023),
                                         (3.0, -1074, 1023), (2.5, -60, 60)])
def test_exact_product_shortcut_whole_range(check, y, lo_e, hi_e):
    bad, _ = check.fast(300_000, y, lo_e, hi_e, seed=40 + int(4 * y))
    assert bad == 0
