"""The reverse-communication mu searches (lresdft_amd/csrc/lres_musearch.h) on the
CPU: the code that both the host loops (lres_mu_solve, lres_mu_solve_warm) and the
device controller of the queued warm step (k_mu_begin / k_mu_step) run, driven here
by Python count functions through a host-only shim compiled with hipcc.

Pins: Brent's roots against scipy.optimize.brentq (the reference's root finder,
lcao3d.py:394-398) on the Butterworth count of a random spectrum (the pole route's
count function, r_M of lrft/core.py:120-127); the keep / lag / re-solve outcomes of
the keep-mu rule (lcao3d.py:412-428); the log-space Brent branch on an exponential
count; the no-bracket exit; and that a search started with the keep test takes the
same trial points as one started from the known count (the device step and the
host search take the same steps)."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle.lres_oracle import r_M

HERE = os.path.dirname(os.path.abspath(__file__))
FN = C.CFUNCTYPE(C.c_double, C.c_double)
KDONE, KNOBRACKET = 3, 4


@pytest.fixture(scope="module")
def shim():
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    d = tempfile.mkdtemp(prefix="lres_musearch_")
    try:
        so = os.path.join(d, "musearch_shim.so")
        subprocess.run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        os.path.join(HERE, "native", "musearch_shim.cpp"), "-o", so], check=True,
                       capture_output=True)
        lib = C.CDLL(so)
        lib.sim_warm.argtypes = [C.c_double] * 8 + [C.c_int, C.c_int, FN, C.POINTER(C.c_double)]
        lib.sim_brent.argtypes = [C.c_double] * 4 + [FN, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        yield lib
    finally:
        shutil.rmtree(d, ignore_errors=True)


def warm(lib, f, N, e_lo, mu0, slope, keep_tol, ftol, xtol=1e-12, lag=0, keep_first=1, f0=0.0):
    trials = []

    def cb(mu):
        trials.append(mu)
        return f(mu)

    out = (C.c_double * 8)()
    lib.sim_warm(N, e_lo, mu0, f0, slope, keep_tol, ftol, xtol, lag, keep_first, FN(cb), out)
    return list(out), trials


def butterworth_count(seed=0, nstate=400, N=200.0, M=32):
    rng = np.random.default_rng(seed)
    lam = np.sort(np.concatenate([rng.uniform(-0.8, -0.3, nstate // 2), rng.uniform(0.1, 0.9, nstate // 2)]))
    c = rng.uniform(0.5, 1.5, nstate)
    c *= (N / 2) / c[: nstate // 2].sum()
    e_lo = lam[0] - 0.3
    return (lambda mu: 2.0 * float(np.dot(c, r_M(lam, mu, e_lo, M))) - N), e_lo


def test_brent_matches_brentq(shim):
    for seed in range(4):
        f, e_lo = butterworth_count(seed)
        a, b = -0.3, 0.5
        root, ev = C.c_double(), C.c_int()
        assert shim.sim_brent(a, b, 1e-12, 0.0, FN(f), C.byref(root), C.byref(ev)) == 1
        ref = brentq(f, a, b, xtol=1e-12)
        assert abs(root.value - ref) < 4e-12
        assert ev.value < 60


def test_keep_lag_and_resolve(shim):
    f, e_lo = butterworth_count(1)
    N = 200.0
    mu_star = brentq(f, -0.3, 0.5, xtol=1e-14)
    # kept: the count at mu0 within keep_tol
    out, tr = warm(shim, f, N, e_lo, mu_star, 50.0, 1e-6 * N, 1e-9 * N)
    assert out[0] == KDONE and out[1] == 0 and out[2] == 1 and out[3] == mu_star and tr == [mu_star]
    # lagged: |count| < 0.25 with lag
    mu_l = mu_star + 0.02
    assert 1e-9 * N < abs(f(mu_l)) < 0.25
    out, tr = warm(shim, f, N, e_lo, mu_l, 50.0, 1e-9 * N, 1e-9 * N, lag=1)
    assert out[1] == 1 and out[3] == mu_l and len(tr) == 1
    # re-solved: the root to the count tolerance
    out, tr = warm(shim, f, N, e_lo, mu_l, 0.0, 1e-12 * N, 1e-12 * N)
    assert out[0] == KDONE and out[1] == 2
    assert abs(f(out[3])) <= 1e-12 * N or abs(out[3] - mu_star) < 1e-11
    assert out[7] in tr


def test_keep_first_takes_the_host_searchs_steps(shim):
    # a search that opens with the keep test at mu0 (the device step) makes the same trials
    # as one started from the count at mu0 (lres_mu_solve_warm after a host keep test)
    f, e_lo = butterworth_count(2)
    N = 200.0
    for mu0, slope in ((0.3, 0.0), (-0.15, 40.0), (0.05, 500.0)):
        a, ta = warm(shim, f, N, e_lo, mu0, slope, 1e-10 * N, 1e-10 * N)
        b, tb = warm(shim, f, N, e_lo, mu0, slope, 0.0, 1e-10 * N, keep_first=0, f0=f(mu0))
        assert ta[0] == mu0 and ta[1:] == tb
        assert a[3] == b[3] and a[4] == b[4] and a[2] == b[2] + 1


def test_log_space_brent_on_an_exponential_count(shim):
    # a flat deficit, then an exponential rise: f(hi) >> |f(lo)| takes the log branch
    N = 100.0
    f = lambda mu: -0.06 + 1e-3 * (np.exp(60.0 * (mu - 0.1)) - 1.0)   # noqa: E731
    root = brentq(f, 0.0, 0.3, xtol=1e-14)
    out, tr = warm(shim, f, N, -1.0, 0.0, 0.05, 0.0, 1e-9 * N)
    assert out[0] == KDONE and out[5] == 1 and out[6] == 1
    assert abs(f(out[3])) <= 1e-9 * N or abs(out[3] - root) < 1e-11
    assert len(tr) < 40


def test_no_bracket_exit(shim):
    out, tr = warm(shim, lambda mu: -1.0, 10.0, -1.0, 0.0, 1.0, 1e-8, 1e-8)
    assert out[0] == KNOBRACKET and len(tr) == 61
