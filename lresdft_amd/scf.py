"""SCF driver on the GPU: the caller of the window step.

Same iteration as the reference's scf_lcao (lrft/lcao3d.py:237-284): LDA + FFT
Hartree potential of the input density, one density build, Pulay/DIIS over the
last ``ndiis`` (density, residual) pairs with mixing ``mix``, clamp at 1e-12
and renormalise to N, Harris-Foulkes energy at the last input density.  The grid
arrays stay on the device (torch); only DIIS's (k+1) x (k+1) system and scalars
go to the host.  Potentials follow lrft/three_d.py:102-148."""
from __future__ import annotations

import math
import time

import numpy as np


class GridOps:
    """FFT Hartree, ion potential and PZ81 LDA on the device for one crystal."""

    def __init__(self, cr, device="cuda"):
        import torch
        self.torch, self.cr, self.device = torch, cr, device
        ks = [2 * np.pi * np.fft.fftfreq(m, d=L / m) for m, L in zip(cr.shape, cr.cell)]
        self.k1 = [torch.tensor(k, dtype=torch.float64, device=device) for k in ks]
        kx, ky, kz = torch.meshgrid(*self.k1, indexing="ij")
        G2 = kx * kx + ky * ky + kz * kz
        G2[0, 0, 0] = 1.0
        self.inv_G2 = 1.0 / G2
        self.inv_G2[0, 0, 0] = 0.0
        self.G2 = G2
        self.G2[0, 0, 0] = 0.0

    def ion_potential(self, chunk=256):
        """Periodic sum of -Z erf(r/rc)/r, G = 0 dropped (three_d.py:102-119).
        The structure factor sum_a Z_a e^{-i G.R_a} is separable per axis and is
        formed as one batched product per chunk of atoms."""
        torch, cr = self.torch, self.cr
        pos = torch.tensor(cr.pos, dtype=torch.float64, device=self.device)
        Z = torch.tensor(cr.Z, dtype=torch.complex128, device=self.device)
        n1, n2, n3 = cr.shape
        Sg = torch.zeros((n1 * n2, n3), dtype=torch.complex128, device=self.device)
        for a0 in range(0, len(cr.pos), chunk):
            p = pos[a0:a0 + chunk]
            ex = torch.exp(-1j * p[:, 0:1] * self.k1[0][None, :]) * Z[a0:a0 + chunk, None]
            ey = torch.exp(-1j * p[:, 1:2] * self.k1[1][None, :])
            ez = torch.exp(-1j * p[:, 2:3] * self.k1[2][None, :])
            P = (ex[:, :, None] * ey[:, None, :]).reshape(len(p), n1 * n2)
            Sg += P.transpose(0, 1) @ ez
        vol = float(np.prod(cr.cell))
        vG = -4 * math.pi * Sg.reshape(cr.shape) * torch.exp(-self.G2 * cr.rc**2 / 4) * self.inv_G2 / vol
        return torch.fft.ifftn(vG).real * cr.ng

    def hartree(self, n):
        torch = self.torch
        nG = torch.fft.fftn(n)
        return torch.fft.ifftn(4 * math.pi * nG * self.inv_G2).real

    def lda_xc(self, n):
        """Slater exchange + Perdew-Zunger 1981 correlation (three_d.py:132-148)."""
        torch = self.torch
        n = torch.clamp(n, min=1e-14)
        rs = (3.0 / (4.0 * math.pi * n)) ** (1.0 / 3.0)
        ex = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * n ** (1.0 / 3.0)
        vx = 4.0 / 3.0 * ex
        gam, b1, b2 = -0.1423, 1.0529, 0.3334
        A, B, C, D = 0.0311, -0.048, 0.0020, -0.0116
        srs = torch.sqrt(rs)
        den = 1.0 + b1 * srs + b2 * rs
        ec_h = gam / den
        vc_h = ec_h * (1.0 + 7.0 / 6.0 * b1 * srs + 4.0 / 3.0 * b2 * rs) / den
        lnr = torch.log(rs)
        ec_l = A * lnr + B + C * rs * lnr + D * rs
        vc_l = A * lnr + (B - A / 3.0) + 2.0 / 3.0 * C * rs * lnr + (2.0 * D - C) / 3.0 * rs
        high = rs >= 1
        return ex + torch.where(high, ec_h, ec_l), vx + torch.where(high, vc_h, vc_l)


def scf_lcao(cr, solver, tol=1e-6, maxiter=60, mix=0.4, ndiis=6, grid=None, vion=None, dens0=None, fp32_floor=1e-4):
    """Device SCF; ``solver.density_device(v)`` returns (density tensor [ng], e_band, mu).

    A solver with ``count_tol``/``refine`` (``LRFTLCAOCpp``) solves mu each step
    to a count accuracy 1e-2 x the current density residual (clamped to
    [1e-11, 1e-4] N), and the converged step is finished in fp64 to the strict
    tolerance (``solver.refine()``).  With mixed precision the iterations run on
    complex64 passes until the residual falls below ``fp32_floor`` (then fp64
    passes): complex64 densities differ from fp64 by ~2e-6 in this residual's
    norm, and DIIS steps built from residuals within a few times that noise
    converge slowly (measured at 1024 electrons: 15 iterations with the
    hand-over at 1e-5, 13 at 1e-4, 12 in pure fp64), so the default floor is
    1e-4.  The energy and the returned density are always fp64 results at the
    converged input density.  One device-to-host transfer per iteration (the
    residual norm and the new row of the DIIS Gram matrix together)."""
    grid = GridOps(cr) if grid is None else grid
    torch = grid.torch
    t0 = time.perf_counter()
    vion = grid.ion_potential() if vion is None else vion
    dV, N = cr.dV, cr.n_elec
    dens = (torch.full(cr.shape, N / (cr.ng * dV), dtype=torch.float64, device=grid.device) if dens0 is None
            else dens0.clone())
    # DIIS history in device rings (slot order oldest first in `slots`), Gram matrix
    # of the stored residuals on the host; one device-to-host transfer per iteration
    Rh = torch.zeros((ndiis, cr.ng), dtype=torch.float64, device=grid.device)
    Nh = torch.zeros((ndiis, cr.ng), dtype=torch.float64, device=grid.device)
    slots = []
    gram = np.zeros((ndiis, ndiis))
    converged = False
    handed_over = False
    errs = []
    t_solver = 0.0
    err = None
    adaptive = hasattr(solver, "count_tol") and hasattr(solver, "refine")
    for it in range(1, maxiter + 1):
        v = vion + grid.hartree(dens) + grid.lda_xc(dens)[1]
        if adaptive:
            # mu to a count accuracy matched to the current density residual: a count
            # error dN moves the residual by about |dN| / N, so |dN| <= 1e-2 err N
            # costs the iteration nothing; the converged step is polished in fp64 to
            # the strict tolerance below
            solver.count_tol = 1e-4 if err is None else min(1e-4, max(1e-2 * err, 1e-11))
        ts = time.perf_counter()
        nout, eband, mu = solver.density_device(v.reshape(-1).contiguous())
        nout = nout.reshape(cr.shape)
        t_solver += time.perf_counter() - ts
        resid = nout - dens
        # the new history slot (the oldest one once the history is full)
        slot = len(slots) if len(slots) < ndiis else slots[0]
        Rh[slot] = resid.reshape(-1)
        Nh[slot] = dens.reshape(-1)
        live = slots[1:] + [slot] if len(slots) == ndiis else slots + [slot]
        # |resid| and the dot products with every ring slot (unused slots are zero)
        stats = torch.cat([torch.sum(torch.abs(resid)).reshape(1), Rh @ resid.reshape(-1)]).cpu().numpy()
        err = float(stats[0]) * dV / N
        errs.append(err)
        if err < tol:
            converged = True
            break
        if err < fp32_floor and getattr(solver, "precision", "fp64") == "mixed":
            # below this residual the complex64 noise would dominate: the remaining
            # iterations use fp64 passes (the converged step is fp64 anyway)
            solver.set_precision("fp64")
            handed_over = True
        slots = live
        for sq in slots:
            gram[slot, sq] = gram[sq, slot] = stats[1 + sq]
        k = len(slots)
        if k > 1:
            Bm = np.zeros((k + 1, k + 1))
            Bm[:k, :k] = gram[np.ix_(slots, slots)]
            Bm[k, :k] = Bm[:k, k] = 1.0
            rhs = np.zeros(k + 1)
            rhs[k] = 1.0
            coef = np.linalg.lstsq(Bm, rhs, rcond=None)[0][:k]
            cfull = np.zeros(ndiis)
            cfull[slots] = coef
            cvec = torch.from_numpy(cfull).to(grid.device)
            dens = (cvec @ Nh + mix * (cvec @ Rh)).reshape(cr.shape)
        else:
            dens = dens + mix * resid
        dens = torch.clamp(dens, min=1e-12)
        dens = dens * (N / (dens.sum() * dV))
    if adaptive:
        # fp64 finish: mu to the strict count tolerance, density and band energy of
        # the converged input potential in fp64
        solver.count_tol = None
        ts = time.perf_counter()
        nout, eband, mu = solver.refine()
        nout = nout.reshape(cr.shape)
        t_solver += time.perf_counter() - ts
    if handed_over:
        solver.set_precision("mixed")
    vh = grid.hartree(dens)
    exc, vxc = grid.lda_xc(dens)
    E_H = 0.5 * float(torch.sum(dens * vh)) * dV
    E_xc = float(torch.sum(dens * exc)) * dV
    E = eband - float(torch.sum((vh + vxc) * dens)) * dV + E_H + E_xc
    torch.cuda.synchronize()
    return dict(energy=E, density=nout, iterations=it, converged=converged, dens_in=dens, eband=eband, mu=mu,
                time=time.perf_counter() - t0, t_solver=t_solver, E_H=E_H, E_xc=E_xc, residuals=errs)
