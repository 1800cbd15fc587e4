"""The SCF iteration's device step (DIIS mix, potential, H(v)) replayed from a hipGraph.

Bars: the mixing kernel vs numpy to 1e-13 relative and bitwise reproducible; every
graph replay equals the eager launch of the same step bit for bit (density, potential
and every window block H_W), over several replays with different coefficients and
eager work and synchronisations in between; the whole device SCF with the graph gives
the same bits as without it (energy, iteration count, residual history).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n=2):
    from lresdft_amd import system
    from lresdft_amd.lcao3d import LRFTLCAOCpp
    from lresdft_amd.scf import GridOps
    cr = system.h2_crystal((n, n, n), a=6.0, pts_per_cell=10, rc=0.6)
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    C = LRFTLCAOCpp(b, system.cell_centres((n, n, n), 6.0), R_w=8.0, M=32)
    g = GridOps(cr)
    return cr, C, g, g.ion_potential()


def _history(torch, cr, nh, seed):
    rng = np.random.default_rng(seed)
    n0 = cr.n_elec / (cr.ng * cr.dV)
    Nh = n0 * (1.0 + 0.2 * rng.random((nh, cr.ng)))
    Rh = n0 * 0.05 * rng.standard_normal((nh, cr.ng))
    return (torch.from_numpy(Nh).cuda(), torch.from_numpy(Rh).cuda(), Nh, Rh)


def test_diis_mix_kernel_matches_numpy():
    import torch
    from lresdft_amd import _native as nat
    cr, _, _, _ = _system(2)
    print("[gpu] lres_diis_mix vs numpy", flush=True)
    Nh, Rh, Nn, Rn = _history(torch, cr, 6, 1)
    coef = np.array([0.1, -0.3, 0.5, 0.2, 0.4, 0.1])
    mix = 0.4
    c = torch.from_numpy(coef).cuda()
    out = [torch.empty(cr.ng, dtype=torch.float64, device="cuda") for _ in range(2)]
    part = torch.empty(nat.LRES_MIX_PARTS, dtype=torch.float64, device="cuda")
    for o in out:
        nat.check(nat.lib().lres_diis_mix(c.data_ptr(), 6, Nh.data_ptr(), Rh.data_ptr(), cr.ng, mix, 1e-12,
                                          float(cr.n_elec), float(cr.dV), o.data_ptr(), part.data_ptr(),
                                          nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    ref = np.maximum(coef @ (Nn + mix * Rn), 1e-12)
    ref *= cr.n_elec / (ref.sum() * cr.dV)
    got = out[0].cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert torch.equal(out[0], out[1])
    assert abs(got.sum() * cr.dV - cr.n_elec) < 1e-12 * cr.n_elec


def test_graph_replays_match_eager_step():
    import torch
    from lresdft_amd.scf import PotentialStep
    cr, C, g, vion = _system(2)
    print("[gpu] potential step: graph replays vs eager, n=2", flush=True)
    Nh, Rh, _, _ = _history(torch, cr, 6, 2)
    eager = PotentialStep(cr, g, C, vion, Nh, Rh, 0.4, graph=False)
    graph = PotentialStep(cr, g, C, vion, Nh, Rh, 0.4, graph=True)
    rng = np.random.default_rng(3)
    for trial in range(4):
        coef = rng.standard_normal(6)
        coef /= coef.sum()
        d_e, v_e = eager.run(coef)
        d_e, v_e = d_e.clone(), v_e.clone()
        H_e = C.plan.window_blocks()[0].clone()
        # eager work and a synchronisation between the two
        C.plan.run(0.1, -1.0, want_density=False)
        torch.cuda.synchronize()
        d_g, v_g = graph.run(coef)
        H_g = C.plan.window_blocks()[0].clone()
        torch.cuda.synchronize()
        assert torch.equal(d_e, d_g), trial
        assert torch.equal(v_e, v_g), trial
        assert torch.equal(H_e, H_g), trial
    assert graph.replays == 4 and graph.graph is not None


def test_scf_with_graph_equals_eager_scf():
    cr, C, g, vion = _system(2)
    from lresdft_amd.scf import scf_lcao
    print("[gpu] device SCF n=2: graph vs eager", flush=True)
    r0 = scf_lcao(cr, C, tol=1e-8, grid=g, vion=vion, graph=False)
    C.mu = None
    r1 = scf_lcao(cr, C, tol=1e-8, grid=g, vion=vion, graph=True)
    assert r0["converged"] and r1["converged"]
    assert r1["graph_replays"] == r1["iterations"] - 1 and r0["graph_replays"] == 0
    assert r0["iterations"] == r1["iterations"] and r0["residuals"] == r1["residuals"]
    assert r0["energy"] == r1["energy"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_whole_iteration_graph_scf_equals_eager_scf(precision):
    """Warm iterations as one graph replay each (mixing, potential, H(v), the warm
    density step with its keep-edge test and mu search, the residual; one graph per
    precision and history slot, the step's inputs read from pinned memory): the SCF
    equals the eager one bit for bit, the graphs are replayed in later SCFs of the
    same solver, and those SCFs equal the eager one too (eager work and waits between
    every replay)."""
    import torch
    cr, C, g, vion = _system(3)
    from lresdft_amd.scf import scf_lcao
    print(f"[gpu] device SCF n=3 {precision}: whole-iteration graphs vs eager", flush=True)
    C.set_precision(precision)
    r0 = scf_lcao(cr, C, tol=1e-8, grid=g, vion=vion, graph=False)
    h0 = list(C.history)
    runs = []
    for _ in range(2):
        C.mu = None
        C.history = []
        runs.append(scf_lcao(cr, C, tol=1e-8, grid=g, vion=vion, graph=True))
        assert C.history == h0
    step = C._scf_step
    print(f"    iterations {r0['iterations']}, iteration graphs {sorted(step.iter_graphs)}", flush=True)
    assert len(step.iter_graphs) > 0
    for r1 in runs:
        assert r1["converged"] and r1["iterations"] == r0["iterations"]
        assert r1["residuals"] == r0["residuals"] and r1["energy"] == r0["energy"] and r1["mu"] == r0["mu"]
        assert torch.equal(r0["density"], r1["density"])
