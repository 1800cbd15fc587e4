"""SCF driver on the GPU: the caller of the window step.

Same iteration as the reference's scf_lcao (lrft/lcao3d.py:237-284): LDA + FFT
Hartree potential of the input density, one density build, Pulay/DIIS over the
last ``ndiis`` (density, residual) pairs with mixing ``mix``, clamp at 1e-12
and renormalise to N, Harris-Foulkes energy at the last input density.  The grid
arrays stay on the device (torch); only DIIS's (k+1) x (k+1) system and scalars
go to the host.  Potentials follow lrft/three_d.py:102-148."""
from __future__ import annotations

import gc
import math
import time
import weakref

import numpy as np

from . import _native as nat


_GELSD = {}


def diis_coefficients(gram, slots):
    """Pulay coefficients of the stored residuals `slots` (oldest first) from their Gram
    matrix: the least-squares solution of [[B, 1], [1^T, 0]] [c; lambda] = [0; 1]
    (lcao3d.py:268-273).  LAPACK's dgelsd with numpy.linalg.lstsq's rcond (eps times
    the order), called directly: the same solution without lstsq's wrapper, which
    costs more than the solve at this size (the host sits between two iterations of
    GPU work while it runs)."""
    from scipy.linalg import lapack
    k = len(slots)
    idx = np.asarray(slots)
    Bm = np.ones((k + 1, k + 1))
    Bm[k, k] = 0.0
    Bm[:k, :k] = gram[idx[:, None], idx]
    rhs = np.zeros((k + 1, 1))
    rhs[k, 0] = 1.0
    lw = _GELSD.get(k)
    if lw is None:
        w = lapack.dgelsd_lwork(k + 1, k + 1, 1)
        lw = _GELSD[k] = (int(w[0]), int(w[1]))
    x, _, _, info = lapack.dgelsd(Bm, rhs, lw[0], lw[1], np.finfo(np.float64).eps * (k + 1))
    if info != 0:
        return np.linalg.lstsq(Bm, rhs[:, 0], rcond=None)[0][:k]
    return x[:k, 0]


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
        # 4 pi / G^2 on the half spectrum of a real-to-complex transform
        self.w_half = (4 * math.pi * self.inv_G2[:, :, : cr.shape[2] // 2 + 1]).contiguous()
        self.part = None
        # the FFT library builds (and on the GPU compiles) its transform plans on first
        # use: do that here, at set-up, rather than inside the first SCF step
        self.hartree(torch.zeros(cr.shape, dtype=torch.float64, device=device))

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

    def ion_forces(self, n, chunk=256):
        """The ion-potential part of the force, -dV sum_r n(r) dv_ion(r)/dR_A, for every
        atom (natoms x 3 tensor).  With v_ion = Re IFFT(vG) as ion_potential forms it
        (three_d.py:102-119), vG = -4 pi sum_A Z_A e^{-iG.R_A} e^{-G^2 rc^2/4} / (G^2 vol),
        the sum is Re sum_G (-i G) vG_A(G) conj(FFT(n)(G)); the per-atom phase e^{-iG.R_A}
        factorises per axis, so it is two batched products per chunk of atoms (the
        second for the G_z weight)."""
        torch, cr = self.torch, self.cr
        n1, n2, n3 = cr.shape
        vol = float(np.prod(cr.cell))
        nG = torch.fft.fftn(n.reshape(cr.shape))
        P = (-4 * math.pi / vol) * torch.exp(-self.G2 * cr.rc**2 / 4) * self.inv_G2 * torch.conj(nG)
        P2 = P.reshape(n1 * n2, n3)
        k1, k2, k3 = self.k1
        K1 = k1.repeat_interleave(n2)          # G_x of the flattened (i, j) index
        K2 = k2.repeat(n1)                     # G_y
        pos = torch.tensor(cr.pos, dtype=torch.float64, device=self.device)
        Z = torch.tensor(cr.Z, dtype=torch.float64, device=self.device)
        out = torch.empty((len(cr.pos), 3), dtype=torch.float64, device=self.device)
        for a0 in range(0, len(cr.pos), chunk):
            p = pos[a0:a0 + chunk]
            ex = torch.exp(-1j * p[:, 0:1] * k1[None, :])
            ey = torch.exp(-1j * p[:, 1:2] * k2[None, :])
            ez = torch.exp(-1j * p[:, 2:3] * k3[None, :])
            Q = (P2 @ ez.transpose(0, 1)).transpose(0, 1)                        # (atoms, n1 n2)
            Qz = (P2 @ (k3[:, None] * ez.transpose(0, 1))).transpose(0, 1)
            E = (ex[:, :, None] * ey[:, None, :]).reshape(len(p), n1 * n2)
            t = torch.stack([(E * Q * K1[None, :]).sum(1), (E * Q * K2[None, :]).sum(1), (E * Qz).sum(1)], 1)
            # sum_r n dv/dR = Re(-i t); the force is -dV times it
            out[a0:a0 + chunk] = -cr.dV * Z[a0:a0 + chunk, None] * torch.imag(t)
        return out

    def hartree(self, n):
        """v_H = IFFT(4 pi FFT(n) / G^2), G = 0 dropped (three_d.py:122-129); the
        density is real, so real-to-complex transforms over the half spectrum."""
        torch = self.torch
        n = n.reshape(self.cr.shape)
        return torch.fft.irfftn(torch.fft.rfftn(n) * self.w_half, s=tuple(self.cr.shape))

    def _workspace(self):
        if self.part is None:
            rows = max(3, nat.LRES_MAX_HIST + 1)
            self.part = self.torch.empty(rows * nat.LRES_MIX_PARTS, dtype=self.torch.float64, device=self.device)
        return self.part

    def potential(self, n, vion):
        """v = v_ion + v_H(n) + v_xc(n) as a flat device tensor: the FFT Hartree
        potential, then one fused kernel for the LDA potential and the sum
        (lres_lda_potential)."""
        torch = self.torch
        n = n.reshape(-1).contiguous()
        vh = self.hartree(n).reshape(-1)
        v = torch.empty_like(n)
        nat.check(nat.lib().lres_lda_potential(n.data_ptr(), vion.reshape(-1).contiguous().data_ptr(), vh.data_ptr(),
                                               n.numel(), v.data_ptr(),
                                               nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return v

    def harris(self, n, eband):
        """Harris-Foulkes energy at input density n (lcao3d.py:275-280):
        (E, E_H, E_xc) with E = e_band - int (v_H + v_xc) n + E_H + E_xc.  One
        pass over the grid (lres_harris_terms) and one transfer of three sums."""
        torch = self.torch
        n = n.reshape(-1).contiguous()
        vh = self.hartree(n).reshape(-1)
        out = torch.empty(3, dtype=torch.float64, device=self.device)
        nat.check(nat.lib().lres_harris_terms(n.data_ptr(), vh.data_ptr(), n.numel(), self._workspace().data_ptr(),
                                              out.data_ptr(), nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        s_vh, s_exc, s_v = out.cpu().tolist()
        dV = self.cr.dV
        E_H = 0.5 * s_vh * dV
        E_xc = s_exc * dV
        return eband - s_v * dV + E_H + E_xc, E_H, E_xc

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


class PotentialStep:
    """The device work of one SCF iteration between two host decisions: the DIIS
    step `dens = max(sum_h c_h (N_h + mix R_h), 1e-12)` renormalised to N (one
    fixed-order library call, lres_diis_mix), then `v = v_ion + v_H(dens) +
    v_xc(dens)` (FFT Hartree, PZ81) and the window blocks H_W(v) (lres_set_potential).
    With graph=True the sequence is captured once into a hipGraph (torch.cuda.CUDAGraph
    on a side stream, the plan switched to that stream while capturing) and replayed
    every iteration; the coefficients travel through a pinned host buffer and a fixed
    device buffer, the outputs (dens, v) live at fixed addresses.  Eager and replayed
    steps run the same kernels in the same order, so they give the same bits."""

    def __init__(self, cr, grid, solver, vion, Nh, Rh, mix, graph=True):
        torch = grid.torch
        # the solver keeps this step (solver._scf_step): a weak reference back, so that a
        # dropped solver and its captured graph are freed at once, not by a later cyclic
        # collection (which may run while another graph is being captured, where
        # destroying a graph is not permitted)
        self.torch, self.cr, self.grid, self.vion = torch, cr, grid, vion
        self._solver = weakref.ref(solver)
        self.Nh, self.Rh, self.mix = Nh, Rh, float(mix)
        self.use_graph = bool(graph)
        self.key = (id(grid), id(vion), vion.data_ptr(), Nh.shape[0], self.mix, self.use_graph)
        nh = Nh.shape[0]
        self.coef_host = torch.zeros(nh, dtype=torch.float64).pin_memory()
        self.coef = torch.zeros(nh, dtype=torch.float64, device=grid.device)
        self.dens = torch.empty(cr.ng, dtype=torch.float64, device=grid.device)
        self.part = torch.empty(nat.LRES_MIX_PARTS, dtype=torch.float64, device=grid.device)
        # the residual statistics (sum |r| and the Gram row): pinned, written by the
        # residual kernel (also from a replayed iteration graph), and on the device
        self.stats_pin = torch.empty(nh + 1, dtype=torch.float64).pin_memory()
        self.stats_dev = torch.empty(nh + 1, dtype=torch.float64, device=grid.device)
        self.v = None
        self.graph = None
        # whole-iteration graphs per (precision, history slot): (graph, potential)
        self.iter_graphs = {}
        self.replays = 0
        # the first iteration's potential and H(v) from a given density: (graph, potential)
        self.pot_graph = None
        self.pot_replays = 0

    @property
    def solver(self):
        return self._solver()

    def _body(self):
        torch = self.torch
        nat.check(nat.lib().lres_diis_mix(self.coef.data_ptr(), self.Nh.shape[0], self.Nh.data_ptr(),
                                          self.Rh.data_ptr(), self.cr.ng, self.mix, 1e-12, float(self.cr.n_elec),
                                          float(self.cr.dV), self.dens.data_ptr(), self.part.data_ptr(),
                                          nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        v = self.grid.potential(self.dens, self.vion)
        self.solver.plan.set_potential(v)
        return v

    def _potential_body(self):
        v = self.grid.potential(self.dens, self.vion)
        self.solver.plan.set_potential(v)
        return v

    def _capture(self, tail=None, body=None):
        """Capture body() (default _body(); and tail(), queued behind it) into a new graph on
        a side stream, the plan switched to that stream meanwhile; returns (graph, potential)."""
        torch = self.torch
        plan = self.solver.plan
        home = plan.stream
        side = torch.cuda.Stream()
        torch.cuda.synchronize()
        # garbage holding other graphs is collected before the capture, not during it
        gc.collect()
        g = torch.cuda.CUDAGraph()
        plan.set_stream(side)
        try:
            with torch.cuda.graph(g, stream=side):
                v = (self._body if body is None else body)()
                if tail is not None:
                    tail()
        finally:
            plan.set_stream(home)
        return g, v

    def run(self, coef):
        """Mix with the coefficient vector `coef` (numpy, one per history row) and
        assemble H_W of the new potential; returns (density [ng], potential [ng])."""
        self.coef_host.numpy()[:] = coef
        self.coef.copy_(self.coef_host, non_blocking=True)
        if self.use_graph:
            if self.graph is None:
                self.graph, self.v = self._capture()
            self.graph.replay()
            self.replays += 1
        else:
            self.v = self._body()
        return self.dens, self.v

    def run_potential(self, dens):
        """The potential of a given density and the window blocks H_W(v) (the SCF's first
        iteration, before any history): dens copied into the step's density buffer, then
        the potential and H(v) of run(), replayed from a graph of their own with graphs on
        (captured at the first use); returns the potential [ng]."""
        self.dens.copy_(dens.reshape(-1))
        if not self.use_graph:
            return self._potential_body()
        if self.pot_graph is None:
            # once eagerly first: library state the potential needs (the FFT plans of this
            # grid) is then set up outside the capture
            self._potential_body()
            self.pot_graph = self._capture(body=self._potential_body)
        self.pot_graph[0].replay()
        self.pot_replays += 1
        return self.pot_graph[1]

    def run_iteration(self, coef, key, tail):
        """The whole iteration as one graph: the step of run() and the work tail()
        queues behind it (the solver's warm density step from its parameter block and
        the residual of its density into history slot key[1]), captured at the first
        use of `key` and replayed after; returns (density [ng], potential [ng]).  The
        step's inputs are read from pinned memory when the graph runs, so the caller
        writes them before this call and arms the solver after it."""
        self.coef_host.numpy()[:] = coef
        self.coef.copy_(self.coef_host, non_blocking=True)
        entry = self.iter_graphs.get(key)
        if entry is None:
            entry = self.iter_graphs[key] = self._capture(tail)
        entry[0].replay()
        self.replays += 1
        return self.dens, entry[1]


def scf_lcao(cr, solver, tol=1e-6, maxiter=60, mix=0.4, ndiis=6, grid=None, vion=None, dens0=None, fp32_floor=3e-5,
             graph=True, count_rel=1e-2, count_cap=1e-4):
    """Device SCF; ``solver.density_device(v)`` returns (density tensor [ng], e_band, mu).

    A solver with ``count_tol``/``refine`` (``LRFTLCAOCpp``) solves mu each step
    to a count accuracy ``count_rel`` x the current density residual (clamped to
    [1e-11, 1e-4] N), and the converged step is finished in fp64 to the strict
    tolerance (``solver.refine()``).  The default 1e-2 is the measured best at
    432 and 1024 electrons: 3e-2 and 1e-1 add one or two iterations (the count
    error enters the residual) and more passes.  With mixed precision the iterations run on
    complex64 passes until the residual falls below ``fp32_floor`` (then fp64
    passes): complex64 densities differ from fp64 by ~2e-6 in this residual's
    norm, and DIIS steps built from residuals within a few times that noise
    converge slowly.  Measured at 1024 electrons (13 iterations in each case,
    14 at 1e-5): 45.7 ms with the hand-over at 1e-4, 41.2 ms at 3e-5, so the
    default floor is 3e-5.  The energy and the returned density are always fp64 results at the
    converged input density.  A warm step (a previous mu) is queued with its mu
    search on the device (solver.density_queue) and the residual behind it, so
    the iteration waits once: for the step's outcome and, in the same transfer,
    the residual norm and the new row of the DIIS Gram matrix (a new filter edge
    or a stale mu falls back to the host-driven paths).  The
    mixing, potential and H(v) of every iteration after the first form one
    PotentialStep, replayed from a captured graph when ``graph`` is true (same
    results as eager launches, bit for bit)."""
    grid = GridOps(cr) if grid is None else grid
    torch = grid.torch
    t0 = time.perf_counter()
    vion = grid.ion_potential() if vion is None else vion
    dV, N = cr.dV, cr.n_elec
    dens = (torch.full(cr.shape, N / (cr.ng * dV), dtype=torch.float64, device=grid.device) if dens0 is None
            else dens0.clone())
    # DIIS history in device rings (slot order oldest first in `slots`), Gram matrix
    # of the stored residuals on the host; one device-to-host transfer per iteration
    # the step (its graph and the history rings it reads by address) is kept on the
    # solver and reused by later SCF runs with the same grid, ion potential and mixing
    step = None
    if hasattr(solver, "plan"):
        step = getattr(solver, "_scf_step", None)
        key = (id(grid), id(vion), vion.data_ptr(), ndiis, float(mix), bool(graph))
        if step is None or step.key != key or step.cr is not cr:
            ring = torch.zeros((2, ndiis, cr.ng), dtype=torch.float64, device=grid.device)
            step = PotentialStep(cr, grid, solver, vion, ring[0], ring[1], mix, graph)
            solver._scf_step = step
        Nh, Rh = step.Nh, step.Rh
        Nh.zero_()
        Rh.zero_()
    else:
        Rh = torch.zeros((ndiis, cr.ng), dtype=torch.float64, device=grid.device)
        Nh = torch.zeros((ndiis, cr.ng), dtype=torch.float64, device=grid.device)
    replays0 = step.replays if step is not None else 0
    slots = []
    gram = np.zeros((ndiis, ndiis))
    converged = False
    handed_over = False
    errs = []
    t_solver = 0.0
    err = None
    adaptive = hasattr(solver, "count_tol") and hasattr(solver, "refine")
    cfull = None
    if step is not None:
        # residual statistics buffers held by the step (a captured iteration graph writes
        # them by address)
        stats_dev, stats_pin = step.stats_dev, step.stats_pin
        if ndiis > nat.LRES_MAX_HIST:
            raise ValueError(f"ndiis must be at most {nat.LRES_MAX_HIST}")
        # warm steps queued on the device with the residual behind them: one wait per
        # iteration (needs the solver's plan on this stream)
        queue_ok = hasattr(solver, "density_queue") and solver.plan.stream == torch.cuda.current_stream()
    for it in range(1, maxiter + 1):
        # a warm step of a solver that can hold it in a graph: the whole iteration (mixing,
        # potential, H(v), the step and its residual) is one replay, below
        whole = False
        if cfull is None:
            if step is not None:
                # the first potential and its H(v) (one replay with graphs on)
                v = step.run_potential(dens)
                preset = True
            else:
                v = grid.potential(dens, vion)
                preset = False
        elif step is not None:
            whole = step.use_graph and queue_ok and solver.step_capturable()
            if not whole:
                dflat, v = step.run(cfull)
                dens = dflat.reshape(cr.shape)
            else:
                dens = step.dens.reshape(cr.shape)
            preset = True
        else:
            cvec = torch.from_numpy(cfull).to(grid.device)
            dens = torch.clamp((cvec @ Nh + mix * (cvec @ Rh)).reshape(cr.shape), min=1e-12)
            dens = dens * (N / (dens.sum() * dV))
            v = grid.potential(dens, vion)
            preset = False
        if adaptive:
            # mu to a count accuracy matched to the current density residual: a count
            # error dN moves the residual by about |dN| / N, so |dN| <= 1e-2 err N
            # costs the iteration nothing (the first step, before any residual is
            # known, takes count_rel x 1e-2); the converged step is polished in fp64
            # to the strict tolerance below
            solver.count_tol = min(count_cap, max(count_rel * (1e-2 if err is None else err), 1e-11))
        ts = time.perf_counter()
        # the new history slot (the oldest one once the history is full)
        free = [q for q in range(ndiis) if q not in slots]
        slot = free[0] if free else slots[0]
        live = slots[1:] + [slot] if len(slots) == ndiis else slots + [slot]
        if step is not None:
            # density without the spin factor; one fused pass forms the residual
            # 2 n_out - n_in, stores the history row and the Gram row
            din = dens.reshape(-1).contiguous()

            def residual(nraw, stats_out):
                nat.check(nat.lib().lres_scf_residual(
                    nraw.data_ptr(), 2.0, din.data_ptr(), cr.ng, ndiis, slot, Nh.data_ptr(), Rh.data_ptr(),
                    grid._workspace().data_ptr(), stats_out.data_ptr(),
                    nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)))

            out = None
            if whole:
                # the step's parameters to pinned memory, then one replay (captured at the
                # first use of this precision and history slot)
                solver.density_params()

                def tail():
                    solver.step_launch()
                    residual(solver.plan.dens, stats_pin)

                dflat, v = step.run_iteration(cfull, (solver.precision, slot), tail)
                solver.density_armed(v)
                queued = True
            else:
                queued = queue_ok and solver.density_queue(v, potential_set=preset)
                if queued:
                    residual(solver.plan.dens, stats_pin)
            if queued:
                # the warm step and the residual of its density queued together, one wait;
                # the residual's totals go straight to pinned host memory
                out = solver.density_collect()
                if out is not None:
                    nraw, eband, mu = out
                    if solver.step_batches > 1:
                        # the search took a further batch: the residual above ran
                        # before this density was formed
                        residual(nraw, stats_dev)
                        stats = stats_dev.cpu().numpy()
                    else:
                        stats = stats_pin.numpy().copy()
            if out is None:
                # first step, new filter edge or stale mu: the host-driven paths
                nraw, eband, mu = solver.density_device(v, potential_set=preset or queued, spin_scaled=False)
                residual(nraw, stats_dev)
                stats = stats_dev.cpu().numpy()
            t_solver += time.perf_counter() - ts
            nout = None
        else:
            nout, eband, mu = solver.density_device(v)
            nout = nout.reshape(cr.shape)
            t_solver += time.perf_counter() - ts
            resid = nout - dens
            Rh[slot] = resid.reshape(-1)
            Nh[slot] = dens.reshape(-1)
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
        ls = np.asarray(slots)
        gram[slot, ls] = gram[ls, slot] = stats[1 + ls]
        k = len(slots)
        cfull = np.zeros(ndiis)
        if k > 1:
            cfull[ls] = diis_coefficients(gram, slots)
        else:
            cfull[slot] = 1.0      # plain mixing: dens + mix * resid
    if adaptive:
        # fp64 finish: mu to the strict count tolerance, density and band energy of
        # the converged input potential in fp64
        solver.count_tol = None
        ts = time.perf_counter()
        nout, eband, mu = solver.refine()
        nout = nout.reshape(cr.shape)
        t_solver += time.perf_counter() - ts
    elif nout is None:
        nout = 2.0 * nraw.reshape(cr.shape)
    if handed_over:
        solver.set_precision("mixed")
    dens = dens.clone()
    E, E_H, E_xc = grid.harris(dens, eband)
    torch.cuda.synchronize()
    return dict(energy=E, density=nout, iterations=it, converged=converged, dens_in=dens, eband=eband, mu=mu,
                time=time.perf_counter() - t0, t_solver=t_solver, E_H=E_H, E_xc=E_xc, residuals=errs,
                graph_replays=step.replays - replays0 if step is not None else 0)
