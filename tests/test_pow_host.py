"""pow_pos (the flux kernels' power, pyminiweather_amd/csrc/pmw_pow.h) built
for the host and checked against 113-bit quadmath and glibc."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO

CSRC = os.path.join(REPO, "pyminiweather_amd", "csrc")
GAMMA = "1.40027894002789400278940027894"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("pow") / "pow_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                        os.path.join(REPO, "tests", "pow_check.cpp"), "-o", exe,
                        "-lquadmath"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("quadmath not available: " + r.stderr[-200:])
    return exe


def _run(exe, n, lo, hi, y):
    out = subprocess.run([exe, str(n), str(lo), str(hi), y], capture_output=True, text=True,
                         check=True).stdout.split()
    return [int(v) for v in out[:5]] + [float(out[5])]


@pytest.mark.parametrize("lo,hi", [(50.0, 2000.0), (1e-3, 1e5), (1e-200, 1e200)])
def test_pow_gamma_correctly_rounded_almost_always(checker, lo, hi):
    """x^gamma over the range rho*theta takes (and far beyond): the result
    is the correctly rounded one in all but <= 2e-5 of the arguments (glibc
    misses ~8e-4 in the physical range); +-0, +-inf, -1, NaN behave as pow."""
    n, bad_cr, _, glibc_cr, special, _ = _run(checker, 200000, lo, hi, GAMMA)
    assert special == 0
    assert bad_cr <= 2e-5 * n, (bad_cr, glibc_cr)


def test_pow_gamma_minus_one(checker):
    """y = gamma - 1 (dp/dq of the adjoint): pow_pos as above, and the
    adjoint's few-ulp pow_pos_fast within 5e-15 relative."""
    n, bad_cr, _, _, special, fast_err = _run(checker, 100000, 1e-3, 1e5,
                                              repr(float(GAMMA) - 1.0))
    assert special == 0 and bad_cr <= 2e-5 * n
    assert fast_err <= 5e-15, fast_err


@pytest.mark.parametrize("lo,hi", [(50.0, 2000.0), (1e-3, 1e5)])
def test_pow_near_errors_stay_inside_its_guard(checker, lo, hi):
    """pow_pos_near decides that its rounding equals pow_pos's when its value
    before rounding is more than 2^-68 relative from a midpoint between
    doubles; that is sound while the two values before rounding together
    lie within that of the exact power. Measured maxima (200k arguments and
    continuations up to 2^-10 away): each ~5e-22, their sum a third of 2^-68."""
    out = subprocess.run([checker, "200000", str(lo), str(hi), GAMMA, "err", "9.7e-4"],
                         capture_output=True, text=True, check=True).stdout.split()
    pp_err, near_err = float(out[1]), float(out[2])
    assert pp_err + near_err < 0.5 * 2.0 ** -68, (pp_err, near_err)


@pytest.mark.parametrize("lo,hi", [(50.0, 2000.0), (1e-3, 1e5)])
def test_pow_near_equals_pow_pos(checker, lo, hi):
    """The fused kernels' stages 2 and 3 continue each face's stage-1 power
    (pow_pos_near): bit for bit pow_pos on 4e6 continuations (and on 2e9 in
    a longer run of the same check over [50, 2000])."""
    out = subprocess.run([checker, "4000000", str(lo), str(hi), GAMMA, "q", "9.7e-4"],
                         capture_output=True, text=True).stdout.split()
    assert int(out[1]) == 0, out


@pytest.mark.parametrize("rxerr", ["1e-15", "1e-14"])
def test_pow_near_insensitive_to_reference_reciprocal(checker, rxerr):
    """On the device the reference's 1/x0 is the hardware reciprocal with
    two Newton steps (an ulp or two), not an IEEE division. With 1/x0 off
    by 1e-15 or 1e-14 relative (~9 and ~90 ulps) pow_pos_near still returns
    pow_pos bit for bit (at 1e-13 it does not: the e^2 terms are evaluated
    from eh, which carries the reciprocal's error)."""
    out = subprocess.run([checker, "2000000", "50", "2000", GAMMA, "q", "9.7e-4", rxerr],
                         capture_output=True, text=True).stdout.split()
    assert int(out[1]) == 0, out


def test_tables_regenerate_identically():
    """pmw_pow_tables.h is what gen_pow_tables.py writes."""
    r = subprocess.run(["python3", os.path.join(CSRC, "gen_pow_tables.py")],
                       capture_output=True, text=True, check=True)
    with open(os.path.join(CSRC, "pmw_pow_tables.h")) as fh:
        assert fh.read() == r.stdout
