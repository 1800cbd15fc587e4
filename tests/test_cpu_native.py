"""The C++/OpenMP CPU restatement of the window passes (oracle/lres_cpu.cpp, the CPU
baseline of the benchmark) against the numpy oracle: eigen route with full spectra
and truncated at c + 1.5 R (dsygvd / dsygvx, src/lrft_cpp.cpp:79-114), the float32
Y^2 store, the pole route (zgesv per pole), 1 and 4 threads giving the same bits,
and a whole LRes density build through LResEigen."""
import numpy as np
import pytest

from oracle import cpu_native
from oracle.lres_oracle import (BASIS, LCAOBasis, LResEigen, build_windows, cell_centres, eig_window_pass,
                                guess_v, h2_crystal, pole_window_pass, r_M)


@pytest.fixture(scope="module")
def sys3():
    cr = h2_crystal((3, 3, 3), a=6.0, pts_per_cell=10, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    v = guess_v(cr)
    W, own = build_windows(b, cell_centres((3, 3, 3), 6.0), 8.0)
    return cr, b, v, W, own, cpu_native.Tables(b, W, own)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not __import__("os").path.exists(cpu_native.LIB):
        cpu_native.build()


def test_native_eigen_pass_matches_numpy(sys3):
    cr, b, v, W, own, tab = sys3
    H = b.H(v)
    e_lo, mu = -0.9, -0.1
    for upper in (None, 0.5 * (mu + e_lo) + 1.5 * 0.5 * (mu - e_lo)):
        d0, l0, c0, o0, y0 = eig_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, 32, upper, True, keep_y2=True)
        d1, l1, c1, o1, y1 = cpu_native.eig_pass(tab, H, mu, e_lo, 32, upper, True, 1, keep_y2=True)
        assert np.array_equal(o0, o1)
        assert np.max(np.abs(l0 - l1)) < 1e-12 and np.max(np.abs(d0 - d1)) < 1e-12
        # eigenvectors inside degenerate window subspaces are LAPACK-path dependent, so
        # the per-state weights c_k are compared through the filtered sums N_B, E_B
        occ = r_M(l0, mu, e_lo, 32)
        for B in range(len(W)):
            sl = slice(o0[B], o0[B + 1])
            assert abs(np.dot(c0[sl], occ[sl]) - np.dot(c1[sl], occ[sl])) < 1e-12
            assert abs(np.dot(c0[sl], l0[sl] * occ[sl]) - np.dot(c1[sl], l1[sl] * occ[sl])) < 1e-12
            # the float32 Y^2 store reproduces the density to float32 rounding
            dB = y1[B].astype(float) @ occ[sl]
            assert np.max(np.abs(dB - d1[own[B]])) < 1e-6 * max(1.0, np.max(np.abs(dB)))
    # threads: same bits
    d4, l4, c4, _, _ = cpu_native.eig_pass(tab, H, mu, e_lo, 32, None, True, 4)
    d1, l1, c1, _, _ = cpu_native.eig_pass(tab, H, mu, e_lo, 32, None, True, 1)
    assert np.array_equal(d1, d4) and np.array_equal(l1, l4) and np.array_equal(c1, c4)


def test_native_pole_pass_matches_numpy(sys3):
    cr, b, v, W, own, tab = sys3
    H = b.H(v)
    e_lo, mu = -0.9, -0.1
    dp, NBp, EBp = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, 32, True)
    dn, NBn, EBn = cpu_native.pole_pass(tab, H, mu, e_lo, 32, True, 4)
    assert np.max(np.abs(dn - dp)) < 1e-12
    assert np.max(np.abs(NBn - NBp)) < 1e-12 and np.max(np.abs(EBn - EBp)) < 1e-12
    de, lam, c, _, _ = cpu_native.eig_pass(tab, H, mu, e_lo, 32, None, True, 4)
    assert np.max(np.abs(dn - de)) < 1e-12      # pole route = full-spectrum eigen route (Lemma 1)


def test_native_density_build_matches_numpy(sys3):
    """LResEigen on the native passes (reference orchestration, float32 Y^2 re-density)
    vs the numpy passes: cold then warm at R_w 8.5, M 64 (tests/test_three_d.py:86-100)."""
    cr, b, v, W, own, tab = sys3
    cells = cell_centres((3, 3, 3), 6.0)
    A = LResEigen(b, cells, R_w=8.5, M=64)
    B = LResEigen(b, cells, R_w=8.5, M=64, native_threads=4, fast_windows=True)
    for _ in range(2):
        na, ea, ma = A.density(v)
        nb_, eb, mb = B.density(v)
        assert abs(ea - eb) < 1e-10 and np.max(np.abs(na - nb_)) < 1e-10 and abs(ma - mb) < 1e-10
    A.mu += 0.02
    B.mu += 0.02
    na, ea, _ = A.density(0.999 * v)
    nb_, eb, _ = B.density(0.999 * v)
    assert abs(ea - eb) < 1e-9 and np.max(np.abs(na - nb_)) < 1e-6
