"""Adjoint of the window step (gradients with respect to the window blocks).

Bars (tolerances written per assertion):
  directional derivatives of 2 sum_B E_B and 2 sum_B N_B along random symmetric
  perturbations of every H_W and S_W at a fixed circle vs Richardson-extrapolated
  central differences (steps 2e-5 and 1e-5 times a direction of spectral norm
  ~20; truncation O(h^4) well below 1e-8 against the nearest pole's distance
  0.06 from the real axis): 1e-6 relative;
  whole-cell windows (every window = the whole cell, sum_B Q_B = S): the
  gradient mapped back to the global H equals the dense-KS closed form
  sum_k g'(lam_k) c_k c_k^T with g(lam) = lam r_M(lam) (first-order eigenvalue
  perturbation, c_k S-normalised, library eigh), and the S gradient (Q fixed)
  equals -sum_k (lam_k g'(lam_k) + g(lam_k)) c_k c_k^T: 1e-9 absolute;
  the torch.autograd.Function's backward equals the adjoint, bitwise;
  the fixed-N derivative (mu re-solved, e_lo fixed) vs the same extrapolated
  differences of the energy at fixed N: 1e-6 relative.
"""
import numpy as np
import pytest
from scipy import linalg

from oracle.lres_oracle import BASIS, LCAOBasis, cell_centres, guess_v, h2_crystal, r_M

pytestmark = pytest.mark.gpu


def _solver(n, Rw, M=32):
    from lresdft_amd.lcao3d import LRFTLCAOCpp, csr64
    cr = h2_crystal((n, n, n), a=6.0, pts_per_cell=10, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    v = guess_v(cr)
    C = LRFTLCAOCpp(b, cell_centres((n, n, n), 6.0), R_w=Rw, M=M)
    H = b.H(v)
    C.plan.set_h(csr64(H))
    lam_min = C.plan.lambda_min()
    return cr, b, H, C, lam_min - 0.3, lam_min + 0.35


def _sym_dirs(plan, seed):
    """Random symmetric direction per window block (flat, layout order)."""
    import torch
    rng = np.random.default_rng(seed)
    mat_off, _ = plan.layout()
    d = np.zeros(mat_off[-1])
    for w, n in enumerate(plan.n):
        a = rng.standard_normal((n, n))
        d[mat_off[w]:mat_off[w + 1]] = (0.5 * (a + a.T)).ravel()
    return torch.from_numpy(d).cuda()


def _energy(plan, HW, SW, mu, e_lo, alpha=1.0, beta=0.0):
    plan.set_window_blocks(HW.contiguous(), SW.contiguous())
    n, e = plan.run(mu, e_lo, want_density=False)
    return 2.0 * (alpha * e + beta * n)


def _richardson(f, h=2e-5):
    """(4 D(h/2) - D(h)) / 3 with D the central difference of f at 0."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("alpha,beta,Rw", [(1.0, 0.0, 8.0), (0.0, 1.0, 8.0), (1.0, 0.0, 10.0)])
def test_adjoint_matches_finite_differences(alpha, beta, Rw):
    """R_w 8: every window function touches the owned points (k = n); R_w 10: n_W 228
    with k < n (functions outside T_B, the blocked out-of-register kernels)."""
    print(f"[gpu] adjoint vs FD n=3 R_w={Rw} alpha={alpha} beta={beta}", flush=True)
    cr, b, H, C, e_lo, mu = _solver(3, Rw)
    plan = C.plan
    if Rw > 8.0:
        assert (plan.k < plan.n).all() and plan.n.min() > 128
    HW, SW = plan.window_blocks()
    gH, gS = plan.adjoint(mu, e_lo, alpha, beta)
    for which, g in (("H", gH), ("S", gS)):
        d = _sym_dirs(plan, 1 if which == "H" else 2) * (1.0 if which == "H" else 0.1)
        if which == "H":
            fd = _richardson(lambda h: _energy(plan, HW + h * d, SW, mu, e_lo, alpha, beta))
        else:
            fd = _richardson(lambda h: _energy(plan, HW, SW + h * d, mu, e_lo, alpha, beta))
        an = 2.0 * float((g * d).sum())
        print(f"    d{which}: adjoint {an:.12e} fd {fd:.12e}", flush=True)
        assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-3)
    plan.set_window_blocks(HW, SW)


def test_adjoint_whole_cell_matches_dense_ks():
    print("[gpu] adjoint whole-cell vs dense KS n=2 R_w=20", flush=True)
    M = 32
    cr, b, H, C, e_lo, mu = _solver(2, 20.0, M)
    plan = C.plan
    assert all(n == b.nb for n in plan.n)        # every window is the whole cell
    gH, gS = plan.adjoint(mu, e_lo, 1.0, 0.0)
    mat_off, fn = plan.layout()
    gHg = np.zeros((b.nb, b.nb))
    gSg = np.zeros((b.nb, b.nb))
    gHc, gSc = gH.cpu().numpy(), gS.cpu().numpy()
    woff = np.concatenate([[0], np.cumsum(plan.n)])
    for w, n in enumerate(plan.n):
        f = fn[woff[w]:woff[w + 1]]
        gHg[np.ix_(f, f)] += gHc[mat_off[w]:mat_off[w + 1]].reshape(n, n)
        gSg[np.ix_(f, f)] += gSc[mat_off[w]:mat_off[w + 1]].reshape(n, n)
    lam, c = linalg.eigh(H.toarray(), b.S.toarray())
    u = (lam - 0.5 * (mu + e_lo)) / (0.5 * (mu - e_lo))
    r = r_M(lam, mu, e_lo, M)
    dr = -M * u ** (M - 1) / (0.5 * (mu - e_lo)) * r * r      # d r_M / d lam
    g, dg = lam * r, r + lam * dr
    refH = (c * dg) @ c.T
    refS = -(c * (lam * dg + g)) @ c.T
    assert np.max(np.abs(gHg - refH)) < 1e-9
    assert np.max(np.abs(gSg - refS)) < 1e-9


def test_autograd_function_backward_is_the_adjoint():
    import torch
    from lresdft_amd.lcao3d import WindowBandEnergy
    print("[gpu] autograd WindowBandEnergy n=3", flush=True)
    cr, b, H, C, e_lo, mu = _solver(3, 8.0)
    plan = C.plan
    HW, SW = plan.window_blocks()
    n0, e0 = plan.run(mu, e_lo, want_density=False)
    HWg = HW.clone().requires_grad_(True)
    SWg = SW.clone().requires_grad_(True)
    E = WindowBandEnergy.apply(plan, HWg, SWg, mu, e_lo)
    assert float(E.detach()) == 2.0 * e0
    E.backward()
    gH, gS = plan.adjoint(mu, e_lo)
    assert torch.equal(HWg.grad, 2.0 * gH) and torch.equal(SWg.grad, 2.0 * gS)


def test_fixed_n_derivative_matches_finite_differences():
    print("[gpu] fixed-N derivative vs FD n=3", flush=True)
    cr, b, H, C, e_lo, mu = _solver(3, 8.0)
    plan = C.plan
    N = cr.n_elec
    HW, SW = plan.window_blocks()
    # mu at fixed N for the unperturbed blocks
    lo, hi = e_lo + 1e-6, mu + 1.0
    mu0, _ = plan.mu_solve(N, e_lo, lo, hi, xtol=1e-14, ftol=1e-13 * N)
    # dE/dmu and dN/dmu at fixed blocks (central differences in mu)
    hm = 1e-5
    np_, ep = plan.run(mu0 + hm, e_lo, want_density=False)
    nm_, em = plan.run(mu0 - hm, e_lo, want_density=False)
    dEdmu, dNdmu = (ep - em) / (2 * hm), (np_ - nm_) / (2 * hm)
    gH, _ = plan.adjoint(mu0, e_lo, 1.0, -dEdmu / dNdmu)
    d = _sym_dirs(plan, 3)

    def e_fixed_n(h):
        plan.set_window_blocks((HW + h * d).contiguous(), SW)
        m, _ = plan.mu_solve(N, e_lo, lo, hi, xtol=1e-14, ftol=1e-13 * N)
        return 2.0 * plan.run(m, e_lo, want_density=False)[1]

    fd = _richardson(e_fixed_n)
    an = 2.0 * float((gH * d).sum())
    print(f"    fixed-N: adjoint {an:.12e} fd {fd:.12e}", flush=True)
    assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-3)
    plan.set_window_blocks(HW, SW)
