"""Atomic forces of the band energy from the adjoint of the window step.

The reference's 3D path has no forces (README.md:35); this is the chain rule
from the window-block gradients of lres_adjoint to the atom positions.  With
the grid potential v frozen and N fixed (mu moves with the geometry, e_lo
fixed), E_band = 2 sum_B E_B depends on the positions through the basis X(R):
H = X^T (T_FD + diag v) X dV, S = X^T X dV (lrft/lcao3d.py:92-100) and
Q_B = sum_{g in own(B)} x_g x_g^T dV.  With beta = -(dE/dmu)/(dN/dmu) (fixed-N
correction of the fixed-circle derivative) and the spin factor 2:

  G^H = 2 sum_W scatter(dE_B/dH_W + beta dN_B/dH_W)   (global, symmetric)
  G^S = 2 sum_W scatter(dE_B/dS_W + beta dN_B/dS_W)
  G^Q_B = 2 (F_B + beta D_B)                           (dE_B/dQ_B = F_B, dN_B/dQ_B = D_B)
  dE/dx_c(g) = 2 dV [ sum_b G^H_cb ((T_FD + v) x_b)(g) + sum_b G^S_cb x_b(g)
                      + sum_u G^Q_{B(g),cu} x_u(g) ]   (B(g): the window owning g)
  F_A = - sum_{c on A} sum_g dE/dx_c(g) dx_c(g)/dR_A.

The window gradients, D_B and F_B come from the GPU; the contraction with the
basis derivatives (one small sparse product per atom) runs on the host, like
the basis construction itself."""
from __future__ import annotations

import numpy as np
from scipy import sparse

from .system import fd_kinetic


def _scatter(blocks, mat_off, fn, n, nb):
    """Sum of per-window n_W x n_W blocks (flat, layout order) into an nb x nb CSR."""
    woff = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    rows, cols = [], []
    for w, nw in enumerate(n):
        f = fn[woff[w]:woff[w + 1]]
        rows.append(np.repeat(f, nw))
        cols.append(np.tile(f, nw))
    return sparse.coo_matrix((blocks, (np.concatenate(rows), np.concatenate(cols))), shape=(nb, nb)).tocsr()


def band_forces(solver, v, mu=None, e_lo=None, N=None, hmu=1e-5):
    """Fixed-N band-energy forces -dE_band/dR (Ha/bohr, shape (natoms, 3)) of
    ``solver`` (an LRFTLCAOCpp) at (mu, e_lo) (default: its last ones), with the
    grid potential v (host array over the grid) frozen.  Also returns the parts
    {"beta", "e_band", "count"} for diagnostics."""
    b = solver.b
    cr = b.cr
    plan = solver.plan
    mu = solver.mu if mu is None else float(mu)
    e_lo = solver.e_lo if e_lo is None else float(e_lo)
    N = cr.n_elec if N is None else N
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    if solver.device_h:
        plan.set_potential(plan.torch.from_numpy(v).cuda())
    else:
        from .lcao3d import csr64
        plan.set_h(csr64(b.H(v)))
    # fixed-N correction: dE/dmu and dN/dmu at fixed H by central differences in mu
    n_p, e_p = plan.run(mu + hmu, e_lo, want_density=False)
    n_m, e_m = plan.run(mu - hmu, e_lo, want_density=False)
    beta = -(e_p - e_m) / (n_p - n_m)
    nsum, esum = plan.run(mu, e_lo, want_density=False)
    D, F = (t.cpu().numpy() for t in plan.pass_blocks())
    gH, gS = (t.cpu().numpy() for t in plan.adjoint(mu, e_lo, 1.0, beta))
    mat_off, fn = plan.layout()
    n = plan.n.astype(np.int64)
    k = plan.k.astype(np.int64)
    GH = 2.0 * _scatter(gH, mat_off, fn, n, b.nb)
    GS = 2.0 * _scatter(gS, mat_off, fn, n, b.nb)
    GHc, GSc = GH.tocsc(), GS.tocsc()

    dV = cr.dV
    X = b.X.tocsc()
    Xr = b.Xr.tocsr()
    A = (fd_kinetic(cr) @ X + sparse.diags(v) @ X).tocsr()     # (T_FD + v) x_b on the grid
    # dE/dQ part: rows = owned grid points, columns = T_B functions, one dense block per window
    woff = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
    kk = np.concatenate([[0], np.cumsum(k * k)]).astype(np.int64)
    q_rows, q_cols, q_vals = [], [], []
    for w in range(len(n)):
        fT = fn[woff[w] + n[w] - k[w]:woff[w + 1]]
        own = solver.own[w]
        if len(own) == 0 or k[w] == 0:
            continue
        GQ = 2.0 * (F[kk[w]:kk[w + 1]] + beta * D[kk[w]:kk[w + 1]]).reshape(k[w], k[w])
        Xo = Xr[own][:, fT].toarray()
        q_rows.append(np.repeat(own, k[w]))
        q_cols.append(np.tile(fT, len(own)))
        q_vals.append((2.0 * dV * (Xo @ GQ)).ravel())
    EQ = sparse.coo_matrix((np.concatenate(q_vals), (np.concatenate(q_rows), np.concatenate(q_cols))),
                           shape=X.shape).tocsc()

    dX = [m.tocsc() for m in b.derivatives()]
    nat = len(cr.pos)
    forces = np.zeros((nat, 3))
    aof = np.asarray(b.atom_of_fn)
    first = np.searchsorted(aof, np.arange(nat))
    last = np.searchsorted(aof, np.arange(nat), side="right")
    for a in range(nat):
        cols = np.arange(first[a], last[a])
        rows = X.indices[X.indptr[cols[0]]:X.indptr[cols[0] + 1]]   # the atom's grid support (shared by its functions)
        Ex = 2.0 * dV * (A[rows] @ GHc[:, cols].toarray() + Xr[rows] @ GSc[:, cols].toarray())
        Ex += EQ[:, cols][rows].toarray()
        for i in range(3):
            forces[a, i] = -float(np.sum(Ex * dX[i][:, cols][rows].toarray()))
    return forces, {"beta": beta, "e_band": 2.0 * esum, "count": 2.0 * nsum - N}


def total_forces(solver, grid, vion, dens_in, mu=None, e_lo=None, N=None):
    """Forces -dE/dR_A (Ha/bohr, natoms x 3) of the Harris-Foulkes SCF energy
    E = e_band - int (v_H + v_xc) n_in + E_H + E_xc (lcao3d.py:275-281, no ion-ion
    term, as in the reference) at a converged SCF with last input density
    ``dens_in`` (device tensor): the fixed-N band-energy forces with the converged
    potential v = v_ion + v_H(n_in) + v_xc(n_in) frozen (band_forces: basis
    derivatives through H, S and the Q_B), plus the ion-potential response
    -int n_in dv_ion/dR_A (GridOps.ion_forces).  The derivative of E through n_in
    vanishes to the extent that the band energy's density response equals n_in
    (Harris-Foulkes stationarity); for the windowed LRes energy the two differ at
    the level of the window error, which the FD checks measure.  Also returns the
    parts {"band", "ion", "beta", "e_band", "count"}."""
    v = grid.potential(dens_in, vion)
    F_band, parts = band_forces(solver, v.cpu().numpy(), mu, e_lo, N)
    F_ion = grid.ion_forces(dens_in).cpu().numpy()
    parts = dict(parts, band=F_band, ion=F_ion)
    return F_band + F_ion, parts
