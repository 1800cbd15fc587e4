"""Geometry, grid and LCAO basis: the inputs of the window step.

Same construction and numbers as the reference's lrft/three_d.py:41-87
(Crystal, h2_crystal), benchmarks/referee_checks.py:128-141 (random-centre
boxes) and lrft/lcao3d.py:49-100 (LCAOBasis), built with whole-array numpy
instead of per-atom loops.  The basis additionally keeps its functions per atom
(``atom_of_fn``), which the device H(v) assembly groups by."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import sparse

DEFAULT_BASIS = dict(rcut=4.5, s_exps=(1.2, 0.4, 0.15), p_exps=(0.5,))


@dataclass
class Crystal:
    pos: np.ndarray        # (natoms, 3) bohr
    Z: np.ndarray          # ionic charges
    cell: np.ndarray       # (3,) orthorhombic box lengths
    shape: tuple           # grid points per axis
    rc: float = 1.0        # softening radius of -Z erf(r/rc)/r

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
    """n0 x n1 x n2 cells with one randomly oriented, jittered H2 each (three_d.py:74-87).

    Random draws per cell in C order: 3 jitter normals, then 3 axis normals."""
    rng = np.random.default_rng(seed)
    ncell = int(np.prod(n))
    centres = (np.stack(np.meshgrid(*[np.arange(m) for m in n], indexing="ij"), -1).reshape(-1, 3) + 0.5) * a
    pos = np.empty((2 * ncell, 3))
    for c in range(ncell):
        centre = centres[c] + rng.normal(scale=jitter, size=3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos[2 * c] = centre - 0.5 * bond * u
        pos[2 * c + 1] = centre + 0.5 * bond * u
    return Crystal(pos, np.ones(2 * ncell), a * np.array(n, float), tuple(int(pts_per_cell * m) for m in n), rc)


def cell_centres(n, a):
    return (np.stack(np.meshgrid(*[np.arange(m) for m in n], indexing="ij"), -1).reshape(-1, 3) + 0.5) * a


def random_h2_box(nmol, box, dmin=3.0, bond=1.4, seed=7, dx=0.6, rc=0.6):
    """H2 molecules at uniform random centres with minimum separation dmin
    (periodic), random axes; rejection sampling as referee_checks.py:128-141."""
    rng = np.random.default_rng(seed)
    L = np.full(3, float(box))
    centres = np.empty((0, 3))
    while len(centres) < nmol:
        c = rng.uniform(0, box, 3)
        if len(centres):
            d = c - centres
            d -= L * np.round(d / L)
            if np.min(np.sqrt(np.sum(d * d, axis=1))) < dmin:
                continue
        centres = np.vstack([centres, c])
    pos = np.empty((2 * nmol, 3))
    for i, c in enumerate(centres):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos[2 * i] = c - 0.5 * bond * u
        pos[2 * i + 1] = c + 0.5 * bond * u
    m = int(round(box / dx))
    return Crystal(pos, np.ones(2 * nmol), L, (m, m, m), rc), centres


def fd_kinetic(cr):
    """-1/2 Laplacian, 7-point periodic stencil, as CSR (three_d.py:151-164 with v = 0)."""
    idx = np.arange(cr.ng).reshape(cr.shape)
    rows = np.tile(idx.ravel(), 6)
    cols = np.concatenate([np.roll(idx, s, axis=ax).ravel() for ax in range(3) for s in (1, -1)])
    off = sparse.csr_matrix((np.full(rows.size, -0.5 / cr.dx**2), (rows, cols)), shape=(cr.ng, cr.ng))
    return (off + sparse.diags(np.full(cr.ng, 3.0 / cr.dx**2))).tocsr()


class LCAOBasis:
    """Cut-off Gaussians on the grid (lcao3d.py:49-100): per atom the s exponents,
    then (x, y, z) per p exponent; envelope (1 - (r/rcut)^2)^2; each function
    normalised by the grid quadrature.  X (N_g x N_b, CSC and CSR), S = X^T X dV,
    T = X^T T_FD X dV, H(v) = T + X^T diag(v) X dV."""

    def __init__(self, cr, rcut=5.5, s_exps=(1.0, 0.3, 0.12), p_exps=(0.5,)):
        self.cr, self.rcut = cr, rcut
        self.s_exps, self.p_exps = tuple(s_exps), tuple(p_exps)
        rad = int(np.ceil(rcut / cr.dx))
        o = np.arange(-rad, rad + 1)
        box = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)
        nat = len(cr.pos)
        nf = len(s_exps) + 3 * len(p_exps)
        pts = np.round(cr.pos / cr.dx).astype(int)[:, None, :] + box[None]          # (nat, nbox, 3)
        d = pts * cr.dx - cr.pos[:, None, :]
        d = d - cr.cell * np.round(d / cr.cell)
        r = np.sqrt(np.sum(d * d, axis=2))
        keep = r < rcut
        shape = np.array(cr.shape)
        gidx = np.ravel_multi_index(tuple(np.moveaxis(pts % shape, -1, 0)), cr.shape)
        env = (1.0 - (r / rcut) ** 2) ** 2
        vals = []
        for al in s_exps:
            vals.append(np.exp(-al * r * r) * env)
        for al in p_exps:
            g = np.exp(-al * r * r) * env
            vals.extend([d[..., 0] * g, d[..., 1] * g, d[..., 2] * g])
        F = np.stack(vals, 1)                                  # (nat, nf, nbox)
        F = np.where(keep[:, None, :], F, 0.0)
        norm = np.sqrt(np.sum(F * F, axis=2) * cr.dV)          # (nat, nf)
        F = F / norm[..., None]
        counts = keep.sum(1)
        # CSC columns: atom-major, function-minor, box order within a column
        col_len = np.repeat(counts, nf)
        indptr = np.concatenate([[0], np.cumsum(col_len)])
        rows = np.concatenate([np.tile(gidx[a][keep[a]], nf) for a in range(nat)])
        data = np.concatenate([F[a][:, keep[a]].ravel() for a in range(nat)])
        self.nb = nat * nf
        self.atom_of_fn = np.repeat(np.arange(nat), nf)
        self.X = sparse.csc_matrix((data, rows, indptr), shape=(cr.ng, self.nb))
        self.Xr = self.X.tocsr()
        dV = cr.dV
        self.S = (self.X.T @ self.X).tocsr() * dV
        self.T = (self.X.T @ (fd_kinetic(cr) @ self.X)).tocsr() * dV

    def H(self, v):
        Xv = self.Xr.multiply(np.asarray(v).reshape(-1, 1)).tocsc()
        return (self.T + (self.X.T @ Xv) * self.cr.dV).tocsr()

    def derivatives(self):
        """d x_c(g) / d R_A for every function c (on atom A) at its grid points, as three
        CSC matrices with X's sparsity and entry order (x, y, z components of R_A).

        x_c = phi_c(g - R_A) / |phi_c|, |phi_c|^2 = sum_g phi_c^2 dV, so
        dx/dR = -grad phi / |phi| + phi (sum_g phi . grad phi dV) / |phi|^3.  With
        u = r^2 / rcut^2, e = exp(-al r^2), env = (1 - u)^2 and
        q = e (-2 al env - 4 (1 - u) / rcut^2): grad phi_s = q d, and for the p
        function along axis i, grad (d_i e env) = e_i e env + d_i q d.  The
        envelope's derivative vanishes at rcut, so the changing set of grid points
        inside the cut-off contributes nothing."""
        cr, rcut = self.cr, self.rcut
        rad = int(np.ceil(rcut / cr.dx))
        o = np.arange(-rad, rad + 1)
        box = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)
        nat = len(cr.pos)
        pts = np.round(cr.pos / cr.dx).astype(int)[:, None, :] + box[None]
        d = pts * cr.dx - cr.pos[:, None, :]
        d = d - cr.cell * np.round(d / cr.cell)
        r2 = np.sum(d * d, axis=2)
        keep = r2 < rcut * rcut
        u = r2 / (rcut * rcut)
        env = (1.0 - u) ** 2
        vals, grads = [], []
        for al in self.s_exps:
            e = np.exp(-al * r2)
            q = e * (-2.0 * al * env - 4.0 * (1.0 - u) / rcut ** 2)
            vals.append(e * env)
            grads.append(q[..., None] * d)
        for al in self.p_exps:
            e = np.exp(-al * r2)
            q = e * (-2.0 * al * env - 4.0 * (1.0 - u) / rcut ** 2)
            for i in range(3):
                vals.append(d[..., i] * e * env)
                g = d[..., i][..., None] * q[..., None] * d
                g[..., i] += e * env
                grads.append(g)
        F = np.where(keep[:, None, :], np.stack(vals, 1), 0.0)                         # (nat, nf, nbox)
        dF = np.where(keep[:, None, :, None], -np.stack(grads, 1), 0.0)                # d/dR = -d/dd
        norm = np.sqrt(np.sum(F * F, axis=2) * cr.dV)                                   # (nat, nf)
        dnorm = np.sum(F[..., None] * dF, axis=2) * cr.dV / norm[..., None]             # (nat, nf, 3)
        dX = dF / norm[..., None, None] - F[..., None] * (dnorm / norm[..., None] ** 2)[:, :, None, :]
        shape = np.array(cr.shape)
        gidx = np.ravel_multi_index(tuple(np.moveaxis(pts % shape, -1, 0)), cr.shape)
        nf = F.shape[1]
        rows = np.concatenate([np.tile(gidx[a][keep[a]], nf) for a in range(nat)])
        indptr = np.concatenate([[0], np.cumsum(np.repeat(keep.sum(1), nf))])
        out = []
        for i in range(3):
            data = np.concatenate([dX[a][:, keep[a], i].ravel() for a in range(nat)])
            out.append(sparse.csc_matrix((data, rows, indptr), shape=self.X.shape))
        return out
