"""CPU oracle for the LRes 3D LCAO window step (numpy/scipy restatement).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product path:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product package never imports or calls it.

It restates, in plain numpy/scipy, the reference's algorithm for the path:

* geometry and grid          lrft/three_d.py:41-87 (Crystal, h2_crystal),
                             benchmarks/referee_checks.py:128-141 (amorphous_h2)
* potentials                 lrft/three_d.py:90-164 (FFT ion/Hartree, PZ81 LDA,
                             7-point FD kinetic operator)
* LCAO basis and matrices    lrft/lcao3d.py:49-100 (LCAOBasis: X, S, T, H(v))
* windows and ownership      lrft/lcao3d.py:132-161 (LRFTLCAO.__init__)
* filter and pole rule       lrft/three_d.py:221-225 (r_M), lrft/core.py:130-139 (nodes)
* eigen-route window pass    src/lrft_cpp.cpp:129-233 (lrft_windows),
                             lrft/lcao3d.py:163-173 (LRFTLCAO._window)
* pole-route window pass     benchmarks/window_resolvent_vs_eig.py:58-65
* mu solve and orchestration lrft/lcao3d.py:359-456 (LRFTLCAOCpp)
* SCF driver                 lrft/lcao3d.py:237-284 (scf_lcao)
* dense Kohn-Sham            lrft/lcao3d.py:110-117 (ks_lcao)

Pinning (each function's checks live in tests/test_oracle.py):
  h2_crystal/LCAOBasis/potentials: published KS gaps and energies
      (results/three_d.json "cost"/"scf", results/scaling_rigorous.json) and the
      exact invariant diag(S) = 1.
  windows/ownership: published n_W (results/three_d.json, referee_checks.json R3)
      and the partition invariant of the owned sets.
  r_M/nodes: Lemma 1 (pole sum = closed form) on a random symmetric pencil.
  eigen pass: published dE_band/e and dn/e vs KS (results/three_d.json "cost").
  pole pass: equality with the full-spectrum eigen pass (the criterion of
      results/window_resolvent.json, <= 2.6e-14) and with dense inverses.
  mu solve / LResEigen / scf_lcao: published SCF energies (three_d.json "scf",
      scaling_rigorous.json) and KS = LRes when the window is the whole cell.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np
import scipy.fft
from scipy import linalg, sparse
from scipy.optimize import brentq

BASIS = dict(rcut=4.5, s_exps=(1.2, 0.4, 0.15), p_exps=(0.5,))   # tests/test_three_d.py:10


# ---------------------------------------------------------------------------
# geometry (lrft/three_d.py:41-87)
# ---------------------------------------------------------------------------


@dataclass
class Crystal:
    pos: np.ndarray
    Z: np.ndarray
    cell: np.ndarray
    shape: tuple
    rc: float = 1.0

    @property
    def dx(self):
        return self.cell[0] / self.shape[0]

    @property
    def ng(self):
        return int(np.prod(self.shape))

    @property
    def dV(self):
        return float(np.prod(self.cell / np.array(self.shape)))

    @property
    def n_elec(self):
        return int(round(self.Z.sum()))


def h2_crystal(n=(2, 2, 2), a=4.5, bond=1.4, pts_per_cell=8, jitter=0.15, seed=0, rc=1.0):
    """One H2 per cell, centre jittered, axis a random unit vector (three_d.py:74-87).

    The random stream is consumed in the reference's order: per cell (i, j, k)
    in C order, first the 3 jitter normals, then the 3 axis normals."""
    rng = np.random.default_rng(seed)
    atoms = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                centre = (np.array([i, j, k]) + 0.5) * a + rng.normal(scale=jitter, size=3)
                axis = rng.normal(size=3)
                axis /= np.linalg.norm(axis)
                atoms.append(centre - 0.5 * bond * axis)
                atoms.append(centre + 0.5 * bond * axis)
    pos = np.array(atoms)
    return Crystal(pos, np.ones(len(pos)), a * np.array(n, float),
                   tuple(int(pts_per_cell * m) for m in n), rc)


def min_image(d, L):
    """Minimum-image displacement in an orthorhombic box (lcao3d.py:45-46)."""
    return d - L * np.round(d / L)


def amorphous_h2(nmol, box, dmin=4.2, bond=1.4, seed=7, pts=10, rc=0.6):
    """Random-centre H2 box by rejection sampling (referee_checks.py:128-141)."""
    rng = np.random.default_rng(seed)
    L = np.full(3, box)
    centres = []
    while len(centres) < nmol:
        c = rng.uniform(0, box, 3)
        if all(np.linalg.norm(min_image(c - q, L)) >= dmin for q in centres):
            centres.append(c)
    atoms = []
    for c in centres:
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        atoms.append(c - 0.5 * bond * axis)
        atoms.append(c + 0.5 * bond * axis)
    m = int(round(box / 0.6))
    return Crystal(np.array(atoms), np.ones(2 * nmol), L.copy(), (m, m, m), rc), np.array(centres)


def cell_centres(n, a):
    """Centres of the n0 x n1 x n2 lattice cells in C order (lcao3d.py:287-289)."""
    g = np.stack(np.meshgrid(*[np.arange(m) for m in n], indexing="ij"), -1).reshape(-1, 3)
    return (g.astype(float) + 0.5) * a


# ---------------------------------------------------------------------------
# potentials (lrft/three_d.py:90-164)
# ---------------------------------------------------------------------------


def _g2(cr):
    ks = [2 * np.pi * np.fft.fftfreq(m, d=L / m) for m, L in zip(cr.shape, cr.cell)]
    kx, ky, kz = np.meshgrid(*ks, indexing="ij")
    return kx, ky, kz, kx**2 + ky**2 + kz**2


def ion_potential(cr):
    """Periodic -Z erf(r/rc)/r summed in reciprocal space, G=0 removed (three_d.py:102-119)."""
    kx, ky, kz, G2 = _g2(cr)
    Sg = np.zeros(cr.shape, complex)
    for p, Z in zip(cr.pos, cr.Z):
        Sg += Z * np.exp(-1j * (kx * p[0] + ky * p[1] + kz * p[2]) - G2 * cr.rc**2 / 4)
    with np.errstate(divide="ignore", invalid="ignore"):
        vG = -4 * np.pi * Sg / (G2 * float(np.prod(cr.cell)))
    vG[0, 0, 0] = 0.0
    return np.real(scipy.fft.ifftn(vG)) * cr.ng


def hartree(cr, n):
    """Periodic Poisson solve by FFT, neutralising background (three_d.py:122-129)."""
    _, _, _, G2 = _g2(cr)
    nG = scipy.fft.fftn(n)
    with np.errstate(divide="ignore", invalid="ignore"):
        vG = 4 * np.pi * nG / G2
    vG[0, 0, 0] = 0.0
    return np.real(scipy.fft.ifftn(vG))


def lda_xc(n):
    """Slater exchange + Perdew-Zunger 1981 correlation, unpolarised (three_d.py:132-148)."""
    n = np.maximum(n, 1e-14)
    rs = (3.0 / (4.0 * np.pi * n)) ** (1.0 / 3.0)
    ex = -0.75 * (3.0 / np.pi) ** (1.0 / 3.0) * n ** (1.0 / 3.0)
    vx = 4.0 / 3.0 * ex
    gam, b1, b2 = -0.1423, 1.0529, 0.3334
    A, B, C, D = 0.0311, -0.048, 0.0020, -0.0116
    srs = np.sqrt(rs)
    den = 1.0 + b1 * srs + b2 * rs
    ec_h = gam / den
    vc_h = ec_h * (1.0 + 7.0 / 6.0 * b1 * srs + 4.0 / 3.0 * b2 * rs) / den
    lnr = np.log(rs)
    ec_l = A * lnr + B + C * rs * lnr + D * rs
    vc_l = A * lnr + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lnr + (2.0 * D - C) / 3.0 * rs
    high = rs >= 1
    return ex + np.where(high, ec_h, ec_l), vx + np.where(high, vc_h, vc_l)


def fd_hamiltonian(cr, v):
    """7-point periodic FD Hamiltonian -1/2 Lap + v as CSR (three_d.py:151-164)."""
    idx = np.arange(cr.ng).reshape(cr.shape)
    h = -0.5 / cr.dx**2
    r, c = [], []
    for ax in range(3):
        for s in (1, -1):
            r.append(idx.ravel())
            c.append(np.roll(idx, s, axis=ax).ravel())
    r, c = np.concatenate(r), np.concatenate(c)
    off = sparse.csr_matrix((np.full(r.size, h), (r, c)), shape=(cr.ng, cr.ng))
    return (off + sparse.diags(v.ravel() + 3.0 / cr.dx**2)).tocsr()


def guess_v(cr):
    """Potential of the uniform start density (tests/test_three_d.py:13-15)."""
    d0 = np.full(cr.shape, cr.n_elec / (cr.ng * cr.dV))
    return ion_potential(cr) + hartree(cr, d0) + lda_xc(d0)[1]


# ---------------------------------------------------------------------------
# LCAO basis (lrft/lcao3d.py:49-100)
# ---------------------------------------------------------------------------


class LCAOBasis:
    """Cut-off Gaussians tabulated on the grid; X (N_g x N_b), S, T, H(v).

    Function order: per atom, the s exponents then (x, y, z) per p exponent
    (lcao3d.py:61-84).  Each column is normalised by the grid quadrature, so
    diag(S) = 1 up to rounding."""

    def __init__(self, cr, rcut=5.5, s_exps=(1.0, 0.3, 0.12), p_exps=(0.5,)):
        self.cr, self.rcut = cr, rcut
        rad = int(np.ceil(rcut / cr.dx))
        o = np.arange(-rad, rad + 1)
        box = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)
        rows, cols, vals, owner = [], [], [], []
        fn = 0
        for a, R in enumerate(cr.pos):
            pts = np.round(R / cr.dx).astype(int) + box
            d = min_image(pts * cr.dx - R, cr.cell)
            r = np.sqrt(np.sum(d * d, axis=1))
            keep = r < rcut
            pts, d, r = pts[keep], d[keep], r[keep]
            gidx = np.ravel_multi_index(tuple((pts % np.array(cr.shape)).T), cr.shape)
            env = (1.0 - (r / rcut) ** 2) ** 2
            funcs = [np.exp(-al * r * r) * env for al in s_exps]
            for al in p_exps:
                g = np.exp(-al * r * r) * env
                funcs.extend([d[:, 0] * g, d[:, 1] * g, d[:, 2] * g])
            for f in funcs:
                rows.append(gidx)
                cols.append(np.full(gidx.size, fn))
                vals.append(f / np.sqrt(np.sum(f * f) * cr.dV))
                owner.append(a)
                fn += 1
        self.nb = fn
        self.atom_of_fn = np.array(owner)
        self.X = sparse.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                   shape=(cr.ng, fn))
        self.Xr = self.X.tocsr()
        dV = cr.dV
        self.S = (self.X.T @ self.X).tocsr() * dV
        T_fd = fd_hamiltonian(cr, np.zeros(cr.shape))
        self.T = (self.X.T @ (T_fd @ self.X)).tocsr() * dV

    def H(self, v):
        """H(v) = T + X^T diag(v) X dV (lcao3d.py:98-100)."""
        Xv = self.Xr.multiply(v.reshape(-1, 1)).tocsc()
        return (self.T + (self.X.T @ Xv) * self.cr.dV).tocsr()

    def density_from_orbitals(self, C):
        Y = self.Xr @ C
        return 2.0 * np.sum(Y * Y, axis=1).reshape(self.cr.shape)


def ks_lcao(basis, v):
    """Dense generalised KS: lowest nocc+1 pairs, closed-shell density (lcao3d.py:110-117)."""
    nocc = basis.cr.n_elec // 2
    w, C = linalg.eigh(basis.H(v).toarray(), basis.S.toarray(), subset_by_index=(0, nocc), driver="gvx")
    return basis.density_from_orbitals(C[:, :nocc]), 2.0 * w[:nocc].sum(), w


# ---------------------------------------------------------------------------
# windows and ownership (lrft/lcao3d.py:132-161)
# ---------------------------------------------------------------------------


def build_windows(basis, cells, R_w):
    """Window function lists W(B) and Voronoi-owned grid points own(B).

    Atoms belong to their nearest cell centre (min-image, argmin -> lowest index
    on ties).  Grid points belong to the nearest centre by squared distance with
    a strict '<' scanned in ascending B, so ties go to the lowest B.  W(B) holds,
    in ascending order, the functions whose cell centre lies within R_w of B."""
    cr = basis.cr
    cells = np.asarray(cells, float)
    da = min_image(cr.pos[:, None, :] - cells[None, :, :], cr.cell)
    cell_of_atom = np.argmin(np.linalg.norm(da, axis=2), axis=1)
    cell_of_fn = cell_of_atom[basis.atom_of_fn]
    dc = np.linalg.norm(min_image(cells[:, None, :] - cells[None, :, :], cr.cell), axis=2)
    grid = np.stack(np.meshgrid(*[np.arange(m) for m in cr.shape], indexing="ij"), -1).reshape(-1, 3) * cr.dx
    owner = np.zeros(cr.ng, np.int64)
    best = np.full(cr.ng, np.inf)
    for B, cB in enumerate(cells):
        dd = min_image(grid - cB, cr.cell)
        d2 = np.einsum("ij,ij->i", dd, dd)
        closer = d2 < best
        best[closer] = d2[closer]
        owner[closer] = B
    windows, own = [], []
    for B in range(len(cells)):
        windows.append(np.flatnonzero(np.isin(cell_of_fn, np.flatnonzero(dc[B] <= R_w))).astype(np.int64))
        own.append(np.flatnonzero(owner == B).astype(np.int64))
    return windows, own


def _d2_ref(g, c, L):
    """Squared minimum-image distance in the operation order of build_windows'
    einsum ((x^2 + y^2) + z^2), elementwise over broadcast g, c."""
    dd = min_image(g - c, L)
    return (dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1]) + dd[..., 2] * dd[..., 2]


def voronoi_owner_fast(grid, cells, L, k=16, chunk=65536):
    """The owner array of build_windows (nearest centre by the same squared
    distance, lowest index on ties) without the N_g x N_cells scan: a periodic
    k-d tree proposes the k nearest centres of every point, their distances are
    recomputed with build_windows' formula, and the smallest value with the
    lowest index wins.  A point is resolved this way only if its k-th proposal is
    farther than its best by a margin far above rounding (so no centre at the
    minimal distance can be missing); the rest take the full scan.  Same result
    as build_windows (checked against it on tie-heavy lattices and random boxes)."""
    from scipy.spatial import cKDTree
    cells = np.asarray(cells, float)
    ncell = len(cells)
    k = min(k, ncell)
    tree = cKDTree(np.mod(cells, L), boxsize=L)
    owner = np.empty(len(grid), np.int64)
    for s in range(0, len(grid), chunk):
        g = grid[s:s + chunk]
        dk, ik = tree.query(g, k=k)
        if k == 1:
            dk, ik = dk[:, None], ik[:, None]
        d2 = _d2_ref(g[:, None, :], cells[ik], L)
        best = d2.min(axis=1)
        # lowest cell index among the candidates at the minimal value
        cand = np.where(d2 == best[:, None], ik, ncell)
        own_c = cand.min(axis=1)
        safe = (dk[:, -1] ** 2 > best * (1 + 1e-9) + 1e-12) if k < ncell else np.ones(len(g), bool)
        bad = np.flatnonzero(~safe)
        for q in bad:
            full = _d2_ref(g[q][None, :], cells, L)
            own_c[q] = int(np.flatnonzero(full == full.min())[0])
        owner[s:s + chunk] = own_c
    return owner


def build_windows_fast(basis, cells, R_w):
    """build_windows with voronoi_owner_fast for the grid ownership (the same
    tables; for the full-size configurations, where the N_g x N_cells scan of the
    restatement takes about a minute)."""
    cr = basis.cr
    cells = np.asarray(cells, float)
    da = min_image(cr.pos[:, None, :] - cells[None, :, :], cr.cell)
    cell_of_atom = np.argmin(np.linalg.norm(da, axis=2), axis=1)
    cell_of_fn = cell_of_atom[basis.atom_of_fn]
    dc = np.linalg.norm(min_image(cells[:, None, :] - cells[None, :, :], cr.cell), axis=2)
    grid = np.stack(np.meshgrid(*[np.arange(m) for m in cr.shape], indexing="ij"), -1).reshape(-1, 3) * cr.dx
    owner = voronoi_owner_fast(grid, cells, np.asarray(cr.cell, float))
    order = np.argsort(owner, kind="stable")
    bounds = np.searchsorted(owner[order], np.arange(len(cells) + 1))
    windows, own = [], []
    for B in range(len(cells)):
        windows.append(np.flatnonzero(np.isin(cell_of_fn, np.flatnonzero(dc[B] <= R_w))).astype(np.int64))
        own.append(order[bounds[B]:bounds[B + 1]].astype(np.int64))
    return windows, own


# ---------------------------------------------------------------------------
# filter and pole rule
# ---------------------------------------------------------------------------


def r_M(lam, mu, e_lo, M):
    """Butterworth occupation 1 / (1 + ((lam - c)/R)^M) (three_d.py:221-225, lrft_cpp.cpp:69-74)."""
    c, R = 0.5 * (mu + e_lo), 0.5 * (mu - e_lo)
    u = np.abs((np.asarray(lam, float) - c) / R)
    with np.errstate(over="ignore"):
        p = u ** M
    return np.where(np.isfinite(p), 1.0 / (1.0 + p), 0.0)


def nodes(M, mu, e_lo):
    """Upper-half-plane poles z_j and weights w_j, j < M/2 (core.py:130-139).

    sum over all M poles of w_j (z_j - lam)^-1 = r_M(lam); the lower half are the
    conjugates, so the full sum is 2 Re of the upper-half sum."""
    if M % 2:
        raise ValueError("M must be even")
    c, R = 0.5 * (mu + e_lo), 0.5 * (mu - e_lo)
    zeta = np.exp(1j * np.pi * (2 * np.arange(M // 2) + 1) / M)
    return c + R * zeta, R * zeta / M


# ---------------------------------------------------------------------------
# window passes
# ---------------------------------------------------------------------------


def dense_block(A, W):
    return A[W][:, W].toarray()


def _eig_one(H, S, Xr, W, own, dV, upper):
    HW, SW = dense_block(H, W), dense_block(S, W)
    if upper is None or not np.isfinite(upper):
        lam, U = linalg.eigh(HW, SW)
    else:
        lam, U = linalg.eigh(HW, SW, subset_by_value=(-np.inf, upper), driver="gvx")
    Y = Xr[own][:, W] @ U
    return lam, Y


def eig_window_pass(H, S, Xr, windows, own, dV, mu, e_lo, M, upper=None, want_density=True, nthreads=1,
                    keep_y2=False):
    """Eigen route of lrft_windows (src/lrft_cpp.cpp:129-233).

    Per window: dense (H_W, S_W) generalised eigenpairs (all of them, or those
    <= upper), Y = X[own, W] U, c_k = sum_own Y^2 dV, density on owned points
    sum_k r_M(lam_k) Y^2.  Returns (dens (N_g,) or None, lam, c, lamoff, y2),
    y2 being the per-window float32 Y^2 (lrft_cpp.cpp:183-185) when keep_y2 is
    set and None otherwise."""
    ng = Xr.shape[0]
    dens = np.zeros(ng) if want_density else None

    def one(B):
        return _eig_one(H, S, Xr, windows[B], own[B], dV, upper)

    if nthreads > 1:
        with ThreadPoolExecutor(nthreads) as ex:
            res = list(ex.map(one, range(len(windows))))
    else:
        res = [one(B) for B in range(len(windows))]
    lams, cs, y2 = [], [], []
    for B, (lam, Y) in enumerate(res):
        lams.append(lam)
        cs.append(np.sum(Y * Y, axis=0) * dV)
        if want_density:
            dens[own[B]] = (Y * Y) @ r_M(lam, mu, e_lo, M)
        if keep_y2:
            y2.append((Y * Y).astype(np.float32))
    lamoff = np.concatenate([[0], np.cumsum([len(l) for l in lams])]).astype(np.int64)
    out = (dens, np.concatenate(lams), np.concatenate(cs), lamoff)
    return out + (y2,) if keep_y2 else out + (None,)


def window_spectra(H, S, Xr, windows, own):
    """Full generalised spectra of every window and the squared owned-point
    amplitudes: [(lam, Y^2)] with Y = X[own, W] U (src/lrft_cpp.cpp:159-185 with
    upper = inf, Y^2 kept in float64).  One decomposition serves any number of
    (mu, e_lo, M) evaluations by spectra_pass."""
    out = []
    for W, o in zip(windows, own):
        lam, Y = _eig_one(H, S, Xr, W, o, None, None)
        out.append((lam, Y * Y))
    return out


def spectra_pass(spectra, own, ng, dV, mu, e_lo, M):
    """The eigen-route pass from window_spectra: (dens (N_g,), N_B, E_B) with
    N_B = sum_k c_k r_M(lam_k), E_B = sum_k c_k lam_k r_M(lam_k), c_k = sum_own Y^2 dV,
    dens[own] = Y^2 r_M (no spin factor; lrft_cpp.cpp:179-190)."""
    dens = np.zeros(ng)
    NB, EB = np.zeros(len(spectra)), np.zeros(len(spectra))
    for B, (lam, Y2) in enumerate(spectra):
        occ = r_M(lam, mu, e_lo, M)
        c = np.sum(Y2, axis=0) * dV
        NB[B], EB[B] = np.dot(c, occ), np.dot(c, lam * occ)
        dens[own[B]] = Y2 @ occ
    return dens, NB, EB


def pole_window_block(HW, SW, T, zs, ws):
    """P[T, T] and F[T, T] of one window by the pole sum (window_resolvent_vs_eig.py:58-65).

    P = 2 Re sum_j w_j (z_j S - H)^-1,  F = 2 Re sum_j w_j z_j (z_j S - H)^-1,
    both restricted to rows/columns T (local indices)."""
    n = HW.shape[0]
    E = np.zeros((n, len(T)))
    E[T, np.arange(len(T))] = 1.0
    A = zs[:, None, None] * SW[None] - HW[None]
    G = np.linalg.solve(A, np.broadcast_to(E, (len(zs), n, len(T))).astype(complex))[:, T, :]
    P = 2.0 * np.real(np.einsum("j,jab->ab", ws, G))
    F = 2.0 * np.real(np.einsum("j,jab->ab", ws * zs, G))
    return P, F


def window_touch_sets(Xr, windows, own):
    """T(B): local window indices of functions with support on own(B), and the
    owned rows of X restricted to them (the data the pole-route density needs)."""
    Ts, Xo = [], []
    for W, o in zip(windows, own):
        Xw = Xr[o][:, W].tocsc()
        T = np.flatnonzero(np.diff(Xw.indptr) > 0)
        Ts.append(T)
        Xo.append(Xw[:, T].tocsr())
    return Ts, Xo


def pole_window_pass(H, S, Xr, windows, own, dV, mu, e_lo, M, want_density=True):
    """Pole route of the window step: per window N_B, E_B and owned density.

    n(r) = x_r^T P_TT x_r,  N_B = sum_own n dV,  E_B = sum_own x_r^T F_TT x_r dV
    (spin factor 2 excluded, as in lrft_windows)."""
    zs, ws = nodes(M, mu, e_lo)
    Ts, Xo = window_touch_sets(Xr, windows, own)
    ng = Xr.shape[0]
    dens = np.zeros(ng) if want_density else None
    NB, EB = np.zeros(len(windows)), np.zeros(len(windows))
    for B, W in enumerate(windows):
        P, F = pole_window_block(dense_block(H, W), dense_block(S, W), Ts[B], zs, ws)
        X = Xo[B]
        nr = np.asarray(np.sum(X.multiply(X @ P), axis=1)).ravel()
        er = np.asarray(np.sum(X.multiply(X @ F), axis=1)).ravel()
        NB[B], EB[B] = nr.sum() * dV, er.sum() * dV
        if want_density:
            dens[own[B]] = nr
    return dens, NB, EB


def lowest_window_eigenvalue(H, S, windows):
    return min(linalg.eigh(dense_block(H, W), dense_block(S, W), eigvals_only=True,
                           subset_by_index=(0, 0))[0] for W in windows)


def window_inertia(H, S, W, sigma):
    """Number of window eigenvalues below sigma (library eigensolver)."""
    lam = linalg.eigh(dense_block(H, W), dense_block(S, W), eigvals_only=True)
    return int(np.sum(lam < sigma))


# ---------------------------------------------------------------------------
# LRes back end with the reference's cold/warm orchestration (lcao3d.py:340-456)
# ---------------------------------------------------------------------------


class StaleMu(RuntimeError):
    """Truncated window spectra hold fewer than N electrons (lcao3d.py:41-42)."""


def solve_mu(lam, c, e_lo, N, M):
    """Brent on the closed-form count, bracket as in lcao3d.py:387-398."""
    def count(m):
        return 2.0 * np.dot(c, r_M(lam, m, e_lo, M)) - N
    lo = max(lam.min(), e_lo + 1e-6)
    if count(lo) > 0:
        raise RuntimeError("electron count positive at the bottom of the spectrum")
    hi = lo + 1.0
    while count(hi) < 0:
        hi += 0.5
        if hi > lam.max() + 10.0:
            raise StaleMu("truncated window spectra hold fewer than N electrons")
    return brentq(count, lo, hi, xtol=1e-12)


class LResEigen:
    """Eigen-route LRes density builder with the reference's cold/warm logic.

    Mirrors LRFTLCAOCpp (lcao3d.py:340-456) with reuse_vectors=True.  The
    float32 storage of Y^2 (lrft_cpp.cpp:185) is reproduced by keeping Y^2 as
    float32 for the mu re-solve path.  ``full_spectrum=True`` replaces the
    dsygvx truncation at c + 1.5 R by full spectra (needed to compare with the
    pole route at M = 32, where the truncation drops r_M up to 1.5^-32)."""

    def __init__(self, basis, cells, R_w=12.1, M=64, margin=0.3, nthreads=1, full_spectrum=False,
                 reuse_vectors=True, keep_tol=1e-8, native_threads=None, fast_windows=False):
        self.b, self.M, self.margin, self.nthreads = basis, M, margin, nthreads
        self.windows, self.own = (build_windows_fast if fast_windows else build_windows)(basis, cells, R_w)
        # native_threads: run the window passes in the C++/OpenMP restatement
        # (oracle/lres_cpu.cpp) on that many threads instead of scipy per window
        self.native_threads = native_threads
        if native_threads:
            from . import cpu_native
            self._tab = cpu_native.Tables(basis, self.windows, self.own)
        self.nW = np.array([len(w) for w in self.windows])
        self.full_spectrum, self.reuse_vectors = full_spectrum, reuse_vectors
        self.keep_tol = keep_tol   # keep mu when |count| <= keep_tol N (the reference: 1e-8, lcao3d.py:412)
        self.mu = self.e_lo = None
        self.passes = 0
        self.history = []

    def _upper(self, mu, e_lo):
        if self.full_spectrum:
            return None
        c, R = 0.5 * (mu + e_lo), 0.5 * (mu - e_lo)
        return c + 1.5 * R

    def _pass(self, H, mu, e_lo, upper, want_density):
        self.passes += 1
        b = self.b
        if self.native_threads:
            from . import cpu_native
            dens, lam, c, lamoff, y2 = cpu_native.eig_pass(self._tab, H, mu, e_lo, self.M, upper, want_density,
                                                           self.native_threads, keep_y2=want_density)
        else:
            dens, lam, c, lamoff, y2 = eig_window_pass(H, b.S, b.Xr, self.windows, self.own, b.cr.dV, mu, e_lo,
                                                       self.M, upper, want_density, self.nthreads,
                                                       keep_y2=want_density)
        self._lamoff, self._y2 = lamoff, y2
        return dens, lam, c

    def _redensity(self, lam, mu, e_lo):
        """Density for a new mu from the stored float32 Y^2 (lcao3d.py:368-381)."""
        dens = np.zeros(self.b.cr.ng)
        occ = r_M(lam, mu, e_lo, self.M)
        for B in range(len(self.windows)):
            l0, l1 = self._lamoff[B], self._lamoff[B + 1]
            dens[self.own[B]] = self._y2[B].astype(np.float64) @ occ[l0:l1] if l1 > l0 else 0.0
        return dens

    def density(self, v, N=None):
        try:
            return self._density(v, N)
        except StaleMu:
            self.mu = None
            return self._density(v, N)

    def _density(self, v, N=None):
        b = self.b
        cr = b.cr
        N = cr.n_elec if N is None else N
        H = b.H(v)
        tol = self.keep_tol * N
        M = self.M
        if self.mu is not None:
            mu, e_lo = self.mu, self.e_lo
            dens, lam, c = self._pass(H, mu, e_lo, self._upper(mu, e_lo), True)
            if lam.size and e_lo + self.margin / 3 < lam.min() < e_lo + 2 * self.margin:
                cnt = 2.0 * np.dot(c, r_M(lam, mu, e_lo, M)) - N
                if abs(cnt) > tol and self.reuse_vectors:
                    mu = solve_mu(lam, c, e_lo, N, M)
                    dens = self._redensity(lam, mu, e_lo)
                    self.history.append("warm-resolve")
                elif abs(cnt) > tol:
                    mu = solve_mu(lam, c, e_lo, N, M)
                    dens, lam, c = self._pass(H, mu, e_lo, self._upper(mu, e_lo), True)
                    self.history.append("warm-second-pass")
                else:
                    self.history.append("warm-keep")
                self.mu = mu
                e_band = 2.0 * np.dot(c, lam * r_M(lam, mu, e_lo, M))
                return 2.0 * dens.reshape(cr.shape), e_band, mu
        _, lam, c = self._pass(H, 0.0, -1.0, None, False)
        e_lo = lam.min() - self.margin
        mu = solve_mu(lam, c, e_lo, N, M)
        dens, lam, c = self._pass(H, mu, e_lo, self._upper(mu, e_lo), True)
        self.mu, self.e_lo = mu, e_lo
        self.history.append("cold")
        e_band = 2.0 * np.dot(c, lam * r_M(lam, mu, e_lo, M))
        return 2.0 * dens.reshape(cr.shape), e_band, mu


class KSDense:
    def __init__(self, basis):
        self.basis = basis

    def density(self, v):
        return ks_lcao(self.basis, v)


# ---------------------------------------------------------------------------
# SCF driver (lrft/lcao3d.py:237-284)
# ---------------------------------------------------------------------------


def scf_lcao(cr, solver, tol=1e-6, maxiter=60, mix=0.4, ndiis=6, dens0=None, vion=None):
    """Pulay-mixed SCF with Harris-Foulkes energy at the last input density.
    vion: the ion potential if already formed (otherwise ion_potential(cr), as
    lcao3d.py:239 does)."""
    vion = ion_potential(cr) if vion is None else vion
    dens = np.full(cr.shape, cr.n_elec / (cr.ng * cr.dV)) if dens0 is None else dens0.copy()
    hist_n, hist_r = [], []
    converged = False
    for it in range(1, maxiter + 1):
        v = vion + hartree(cr, dens) + lda_xc(dens)[1]
        out = solver.density(v)
        nout, eband = out[0], out[1]
        resid = nout - dens
        err = np.sum(np.abs(resid)) * cr.dV / cr.n_elec
        if err < tol:
            converged = True
            break
        hist_n.append(dens.copy())
        hist_r.append(resid.copy())
        if len(hist_n) > ndiis:
            hist_n.pop(0)
            hist_r.pop(0)
        k = len(hist_n)
        if k > 1:
            Bm = np.zeros((k + 1, k + 1))
            for i in range(k):
                for j in range(k):
                    Bm[i, j] = np.vdot(hist_r[i], hist_r[j])
            Bm[k, :k] = Bm[:k, k] = 1.0
            rhs = np.zeros(k + 1)
            rhs[k] = 1.0
            coef = np.linalg.lstsq(Bm, rhs, rcond=None)[0][:k]
            dens = sum(cf * h for cf, h in zip(coef, hist_n)) + mix * sum(cf * r for cf, r in zip(coef, hist_r))
        else:
            dens = dens + mix * resid
        dens = np.maximum(dens, 1e-12)
        dens *= cr.n_elec / (dens.sum() * cr.dV)
    vh = hartree(cr, dens)
    exc, vxc = lda_xc(dens)
    E_H = 0.5 * np.sum(dens * vh) * cr.dV
    E_xc = np.sum(dens * exc) * cr.dV
    E = eband - np.sum((vh + vxc) * dens) * cr.dV + E_H + E_xc
    return dict(energy=E, density=nout, iterations=it, converged=converged, dens_in=dens, eband=eband,
                last=out, E_H=E_H, E_xc=E_xc)
