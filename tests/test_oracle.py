"""Pins of the CPU oracle against the reference's published numbers and against
what the mathematics fixes (closed forms, library eigensolvers, invariants).

Golden values are the reference's committed results (file:key noted per test);
they are restated here as literals."""
import numpy as np
import pytest
from scipy import linalg

from oracle.lres_oracle import (BASIS, LCAOBasis, LResEigen, KSDense, amorphous_h2, build_windows,
                                cell_centres, eig_window_pass, guess_v, h2_crystal, ks_lcao, lda_xc,
                                nodes, pole_window_block, pole_window_pass, r_M, scf_lcao, ion_potential)


def crystal(n, a=6.0):
    return h2_crystal((n, n, n), a=a, pts_per_cell=10, rc=0.6)


@pytest.fixture(scope="module")
def sys2():
    cr = crystal(2)
    b = LCAOBasis(cr, **BASIS)
    return cr, b, guess_v(cr)


@pytest.fixture(scope="module")
def sys3():
    cr = crystal(3)
    b = LCAOBasis(cr, **BASIS)
    return cr, b, guess_v(cr)


@pytest.mark.parametrize("M", [24, 32, 64])
def test_lemma1_pole_sum_is_closed_form_filter(M):
    """sum_j w_j (z_j S - H)^-1 = U r_M(Lambda) U^T for a random pencil (tests/test_lrft.py:42-56,
    generalised to S != I as in lcao3d.py:17)."""
    rng = np.random.default_rng(3)
    n = 40
    A = rng.normal(size=(n, n))
    H = 0.5 * (A + A.T)
    B = rng.normal(size=(n, n))
    S = B @ B.T / n + np.eye(n)
    lam, U = linalg.eigh(H, S)
    mu, e_lo = 0.5 * (lam[10] + lam[11]), lam[0] - 0.3
    zs, ws = nodes(M, mu, e_lo)
    P = np.zeros((n, n), complex)
    for z, w in zip(zs, ws):
        P += w * np.linalg.inv(z * S - H)
    P = 2 * P.real
    ref = (U * r_M(lam, mu, e_lo, M)) @ U.T
    assert np.max(np.abs(P - ref)) < 1e-12
    # weights sum to zero over the full circle: the band-energy identity F = U lam r_M U^T
    F = np.zeros((n, n), complex)
    for z, w in zip(zs, ws):
        F += w * z * np.linalg.inv(z * S - H)
    assert np.max(np.abs(2 * F.real - (U * (lam * r_M(lam, mu, e_lo, M))) @ U.T)) < 1e-12


def test_r_M_special_values():
    mu, e_lo = 0.2, -1.0
    c, R = -0.4, 0.6
    assert r_M(c, mu, e_lo, 32) == 1.0
    assert abs(r_M(mu, mu, e_lo, 32) - 0.5) < 1e-13       # (lam - c)/R = +-1 up to rounding
    assert abs(r_M(e_lo, mu, e_lo, 32) - 0.5) < 1e-13
    assert r_M(1e6, mu, e_lo, 64) == 0.0                 # overflow guard (lrft_cpp.cpp:73)
    with pytest.raises(ValueError):
        nodes(31, mu, e_lo)


def test_crystal_generator_geometry():
    cr = crystal(2)
    d = np.linalg.norm(cr.pos[0::2] - cr.pos[1::2], axis=1)
    assert np.allclose(d, 1.4, atol=1e-12)
    assert cr.shape == (20, 20, 20) and abs(cr.dV - 0.216) < 1e-15 and cr.n_elec == 16


def test_basis_normalised_and_symmetric(sys2):
    cr, b, v = sys2
    assert np.max(np.abs(b.S.diagonal() - 1.0)) < 1e-13
    H = b.H(v)
    assert abs(H - H.T).max() < 1e-13
    assert b.nb == 6 * len(cr.pos)


def test_ion_potential_neutral():
    cr = h2_crystal((1, 1, 1), a=6.0, pts_per_cell=10, rc=0.6)      # tests/test_three_d.py:26-32
    v = ion_potential(cr)
    assert abs(v.mean()) < 1e-10
    i = np.unravel_index(np.argmin(v), v.shape)
    assert min(np.linalg.norm(np.array(i) * cr.dx - p) for p in cr.pos) < 1.0


def test_lda_potential_is_derivative():
    n = np.logspace(-4, 1, 40)                                          # tests/test_three_d.py:18-23
    e, v = lda_xc(n)
    h = 1e-6 * n
    fd = ((n + h) * lda_xc(n + h)[0] - (n - h) * lda_xc(n - h)[0]) / (2 * h)
    assert np.max(np.abs(fd - v) / np.abs(v)) < 1e-6


def test_published_window_sizes():
    """results/three_d.json "window" and window_resolvent.json: n_W at n=6."""
    cr = crystal(6)
    b = LCAOBasis(cr, **BASIS)
    for Rw, nW in ((6.1, 84), (8.0, 84), (8.5, 228), (10.0, 228), (12.1, 396), (13.5, 684)):
        W, own = build_windows(b, cell_centres((6, 6, 6), 6.0), Rw)
        assert all(len(w) == nW for w in W), Rw
    own_all = np.concatenate(own)
    assert own_all.size == cr.ng and np.unique(own_all).size == cr.ng        # partition
    assert all(np.all(np.diff(w) > 0) for w in W)                            # sorted


def test_published_amorphous_window_sizes():
    """results/referee_checks.json R3_amorphous: n_W mean 145.632, max 192 at R_w 8.5."""
    cr, centres = amorphous_h2(125, 30.0)
    b = LCAOBasis(cr, **BASIS)
    W, own = build_windows(b, centres, 8.5)
    nW = np.array([len(w) for w in W])
    assert abs(nW.mean() - 145.632) < 1e-9 and nW.max() == 192


def test_published_gap_and_window_error_n4():
    """results/three_d.json "cost" n=4 (R_w 12.1, M 64, guess potential):
    gap 0.328082031159947, dEband_per_e 1.0988457643801297e-08, dn_per_e 7.982349563092475e-07."""
    cr = crystal(4)
    b = LCAOBasis(cr, **BASIS)
    v = guess_v(cr)
    N = cr.n_elec
    nk, ebk, w = ks_lcao(b, v)
    assert abs((w[N // 2] - w[N // 2 - 1]) - 0.328082031159947) < 1e-12
    nl, ebl, mu = LResEigen(b, cell_centres((4, 4, 4), 6.0), R_w=12.1, M=64, nthreads=4).density(v)
    assert abs((ebl - ebk) / N - 1.0988457643801297e-08) < 1e-14
    assert abs(np.abs(nl - nk).sum() * cr.dV / N - 7.982349563092475e-07) < 1e-13


def test_lres_equals_ks_when_window_is_whole_cell(sys2):
    cr, b, v = sys2                                                      # tests/test_three_d.py:51-60
    nk, ebk, _ = ks_lcao(b, v)
    nl, ebl, _ = LResEigen(b, cell_centres((2, 2, 2), 6.0), R_w=20.0, M=64).density(v)
    assert abs(ebl - ebk) < 1e-10
    assert np.sum(np.abs(nl - nk)) * cr.dV < 1e-9


@pytest.mark.parametrize("M", [32, 64])
def test_pole_route_equals_full_spectrum_eigen_route(sys3, M):
    """Criterion of results/window_resolvent.json (max diff <= 2.6e-14 at M 64), on all
    27 windows of n=3 at R_w 8 (n_W 84), fixed mu / e_lo."""
    cr, b, v = sys3
    H = b.H(v)
    W, own = build_windows(b, cell_centres((3, 3, 3), 6.0), 8.0)
    lam_min = min(linalg.eigh(H[w][:, w].toarray(), b.S[w][:, w].toarray(), eigvals_only=True)[0] for w in W)
    e_lo, mu = lam_min - 0.3, lam_min + 0.35
    de, lam, c, lamoff, _ = eig_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, M, None, True)
    dp, NB, EB = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, M, True)
    assert np.max(np.abs(dp - de)) < 1e-12
    occ = r_M(lam, mu, e_lo, M)
    for B in range(len(W)):
        sl = slice(lamoff[B], lamoff[B + 1])
        assert abs(NB[B] - np.dot(c[sl], occ[sl])) < 1e-12
        assert abs(EB[B] - np.dot(c[sl], lam[sl] * occ[sl])) < 1e-12


def test_pole_block_matches_dense_inverse():
    rng = np.random.default_rng(5)
    n = 30
    A = rng.normal(size=(n, n))
    H = 0.5 * (A + A.T)
    S = np.eye(n) + 0.01 * (A @ A.T) / n
    zs, ws = nodes(16, 0.3, -2.0)
    T = np.array([3, 4, 10, 29])
    P, F = pole_window_block(H, S, T, zs, ws)
    Pr = sum(2 * np.real(w * np.linalg.inv(z * S - H)) for z, w in zip(zs, ws))
    assert np.max(np.abs(P - Pr[np.ix_(T, T)])) < 1e-13


def test_published_scf_energies_n2():
    """results/three_d.json "scf" n=2 (tol 1e-7): E_KS -4.634687711976325 after 13 iterations.

    The LRes row of that file (17 iterations) predates the reference's current
    warm-path logic: with R_w 12.1 the window is the whole cell, so the current
    LRFTLCAOCpp follows the KS iteration to ~1e-10; that is what is checked."""
    cr = crystal(2)
    b = LCAOBasis(cr, **BASIS)
    ks = scf_lcao(cr, KSDense(b), tol=1e-7)
    lr = scf_lcao(cr, LResEigen(b, cell_centres((2, 2, 2), 6.0), R_w=12.1, M=64), tol=1e-7)
    assert ks["converged"] and lr["converged"]
    assert abs(ks["energy"] - (-4.634687711976325)) < 1e-10 and ks["iterations"] == 13
    assert abs(lr["energy"] - ks["energy"]) < 1e-9 and lr["iterations"] == 13


def test_published_scf_energy_n3_scaling_rigorous():
    """results/scaling_rigorous.json n=3 (scf defaults tol 1e-6, R_w 12.1, M 64):
    KS -15.746532895761453 and LRes -15.746532895084757, both after 11 iterations.

    KS is deterministic to 1e-13.  The LRes run re-solves mu once from float32
    Y^2 (lrft_cpp.cpp:185); Y inside degenerate window subspaces is LAPACK-build
    dependent, so its float32 rounding leaves ~1e-10 of path noise in E."""
    cr = crystal(3)
    b = LCAOBasis(cr, **BASIS)
    ks = scf_lcao(cr, KSDense(b))
    assert ks["converged"] and ks["iterations"] == 11
    assert abs(ks["energy"] - (-15.746532895761453)) < 1e-12
    lr = scf_lcao(cr, LResEigen(b, cell_centres((3, 3, 3), 6.0), R_w=12.1, M=64, nthreads=4))
    assert lr["converged"] and lr["iterations"] == 11
    assert abs(lr["energy"] - (-15.746532895084757)) < 1e-9


def test_stored_vectors_reproduce_second_pass(sys3):
    """tests/test_three_d.py:142-156 on the oracle: forced mu re-solve."""
    cr, b, v = sys3
    out = []
    for reuse in (True, False):
        C = LResEigen(b, cell_centres((3, 3, 3), 6.0), R_w=8.5, reuse_vectors=reuse)
        C.density(v)
        C.mu += 0.02
        out.append(C.density(0.999 * v))
    assert abs(out[0][1] - out[1][1]) < 1e-9
    assert np.max(np.abs(out[0][0] - out[1][0])) < 1e-6


@pytest.mark.parametrize("n,gap", [(2, 0.40642405101199747), (3, 0.4048221304541437)])
def test_published_ks_gaps_n2_n3(n, gap):
    """results/three_d.json "cost" n=2, 3: KS gap at the guess potential (uniform start density)."""
    cr = crystal(n)
    b = LCAOBasis(cr, **BASIS)
    _, _, w = ks_lcao(b, guess_v(cr))
    N = cr.n_elec
    assert abs((w[N // 2] - w[N // 2 - 1]) - gap) < 1e-12


@pytest.mark.parametrize("case", ["lattice", "compressed", "random"])
def test_fast_window_builder_equals_restatement(case):
    """build_windows_fast (k-d tree candidates, the restatement's distance formula) gives
    the same W(B) and own(B) as build_windows, Voronoi ties included: the a = 6 lattice
    puts ~27% of grid points on cell boundaries, the a = 5.1 lattice (8 points per cell)
    and a random-centre box test the general case."""
    from oracle.lres_oracle import build_windows_fast
    if case == "lattice":
        cr = crystal(3)
        cells = cell_centres((3, 3, 3), 6.0)
    elif case == "compressed":
        cr = h2_crystal((3, 3, 3), a=5.1, pts_per_cell=8, rc=0.6)
        cells = cell_centres((3, 3, 3), 5.1)
    else:
        cr, cells = amorphous_h2(27, 18.0, dmin=3.0)
    b = LCAOBasis(cr, **BASIS)
    W0, own0 = build_windows(b, cells, 8.0)
    W1, own1 = build_windows_fast(b, cells, 8.0)
    assert len(W0) == len(W1)
    assert all(np.array_equal(x, y) for x, y in zip(W0, W1))
    assert all(np.array_equal(x, y) for x, y in zip(own0, own1))


def test_spectra_pass_equals_eigen_pass(sys3):
    """window_spectra + spectra_pass (one decomposition, several filters) equal
    eig_window_pass with full spectra at M 32 and 64."""
    from oracle.lres_oracle import spectra_pass, window_spectra
    cr, b, v = sys3
    H = b.H(v)
    W, own = build_windows(b, cell_centres((3, 3, 3), 6.0), 8.0)
    sp = window_spectra(H, b.S, b.Xr, W, own)
    lmin = min(s[0].min() for s in sp)
    for M in (32, 64):
        e_lo, mu = lmin - 0.3, lmin + 0.35
        de, lam, c, lamoff, _ = eig_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, M, None, True)
        ds, NB, EB = spectra_pass(sp, own, cr.ng, cr.dV, mu, e_lo, M)
        assert np.max(np.abs(ds - de)) < 1e-14
        occ = r_M(lam, mu, e_lo, M)
        assert abs(NB.sum() - np.dot(c, occ)) < 1e-12 and abs(EB.sum() - np.dot(c, lam * occ)) < 1e-12
