"""Atomic forces from the adjoint of the window step.

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

Everything runs on the GPU: the adjoint kernels give the window gradients, the
pass gives D_B and F_B, and lres_band_forces contracts them with the basis
derivatives tabulated on per-atom boxes (lres_force_setup).  total_forces adds
the ion-potential response for forces of the SCF energy."""
from __future__ import annotations

import numpy as np

from . import _native as nat


def _force_setup(solver):
    """Tabulate the basis of the solver's geometry on the device (once per plan)."""
    plan = solver.plan
    if getattr(plan, "_forces_ready", False):
        return
    b = solver.b
    cr = b.cr
    exps = nat.f64(list(b.s_exps) + list(b.p_exps))
    nat.check(nat.lib().lres_force_setup(plan.h, nat.pf(nat.f64(cr.pos)), len(cr.pos), nat.p64(nat.i64(b.atom_of_fn)),
                                         int(b.nb), float(b.rcut), nat.pf(exps), len(b.s_exps), len(b.p_exps),
                                         nat.pf(nat.f64(cr.cell)), nat.p64(nat.i64(cr.shape)), float(cr.dx)))
    plan._forces_ready = True


def band_forces(solver, v, mu=None, e_lo=None, N=None, hmu=1e-5):
    """Fixed-N band-energy forces -dE_band/dR (Ha/bohr, shape (natoms, 3)) of
    ``solver`` (an LRFTLCAOCpp with a basis that keeps rcut, s_exps and p_exps) at
    (mu, e_lo) (default: its last ones), with the grid potential v (host array or
    device tensor over the grid) frozen.  Also returns the parts
    {"beta", "e_band", "count"} for diagnostics."""
    b = solver.b
    cr = b.cr
    plan = solver.plan
    torch = plan.torch
    mu = solver.mu if mu is None else float(mu)
    e_lo = solver.e_lo if e_lo is None else float(e_lo)
    N = cr.n_elec if N is None else N
    if isinstance(v, np.ndarray):
        v_dev = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).reshape(-1)).cuda()
    else:
        v_dev = v.reshape(-1).contiguous()
    if solver.device_h:
        plan.set_potential(v_dev)
    else:
        from .lcao3d import csr64
        plan.set_h(csr64(b.H(v_dev.cpu().numpy())))
    # fixed-N correction: dE/dmu and dN/dmu at fixed H by central differences in mu
    n_p, e_p = plan.run(mu + hmu, e_lo, want_density=False)
    n_m, e_m = plan.run(mu - hmu, e_lo, want_density=False)
    beta = -(e_p - e_m) / (n_p - n_m)
    nsum, esum = plan.run(mu, e_lo, want_density=False)
    D, F = plan.pass_blocks()
    gH, gS = plan.adjoint(mu, e_lo, 1.0, beta)
    _force_setup(solver)
    out = torch.empty(3 * len(cr.pos), dtype=torch.float64, device="cuda")
    nat.check(nat.lib().lres_band_forces(plan.h, gH.data_ptr(), gS.data_ptr(), D.data_ptr(), F.data_ptr(),
                                         float(beta), v_dev.data_ptr(), out.data_ptr()))
    forces = out.cpu().numpy().reshape(-1, 3)
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
    F_band, parts = band_forces(solver, v, mu, e_lo, N)
    F_ion = grid.ion_forces(dens_in).cpu().numpy()
    parts = dict(parts, band=F_band, ion=F_ion)
    return F_band + F_ion, parts
