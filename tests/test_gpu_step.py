"""The warm density step queued on the device (lres_step_enqueue / lres_step_finish:
keep-edge test, keep pass and mu search decided by the device controller between
the passes) against the host-driven step it replaces (the keep test and
lres_mu_solve_warm, one wait per pass), lcao3d.py:412-448.

Both run the same search code (lres_musearch.h) on the same pass sums, so every
decision, trial mu and result is the same: the SCFs are compared bit for bit
(energy, mu, per-step histories and pass counts) in both precisions, and the
number of trial slots queued per batch changes nothing."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys3():
    from lresdft_amd import system
    from lresdft_amd.scf import GridOps
    cr = system.h2_crystal((3, 3, 3), a=6.0, pts_per_cell=10, rc=0.6)
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    cells = system.cell_centres((3, 3, 3), 6.0)
    g = GridOps(cr)
    return cr, b, cells, g, g.ion_potential()


def _scf(sys3, precision, **kw):
    from lresdft_amd.lcao3d import LRFTLCAOCpp
    from lresdft_amd.scf import scf_lcao
    cr, b, cells, g, vion = sys3
    C = LRFTLCAOCpp(b, cells, R_w=8.0, M=32, precision=precision, **kw)
    r = scf_lcao(cr, C, tol=1e-7, grid=g, vion=vion)
    assert r["converged"]
    return r, C


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_step_scf_equals_host_step_scf(sys3, precision):
    print(f"[gpu] device vs host warm step, SCF n=3 R_w 8 M 32 {precision}", flush=True)
    rd, Cd = _scf(sys3, precision, device_step=True)
    rh, Ch = _scf(sys3, precision, device_step=False)
    print(f"    histories {Cd.history}\n    passes {Cd.pass_log} vs {Ch.pass_log}", flush=True)
    assert "warm-keep" in Cd.history and "warm-resolve" in Cd.history
    assert Cd.history == Ch.history
    assert Cd.pass_log == Ch.pass_log
    assert rd["iterations"] == rh["iterations"]
    assert rd["mu"] == rh["mu"] and rd["energy"] == rh["energy"]
    assert np.array_equal(rd["density"].cpu().numpy(), rh["density"].cpu().numpy())


def test_trial_slots_per_batch_change_nothing(sys3):
    print("[gpu] device step with 1 and 6 trial slots per batch, SCF n=3 fp64", flush=True)
    r1, C1 = _scf(sys3, "fp64", step_depth=1)
    r6, C6 = _scf(sys3, "fp64", step_depth=6)
    assert C1.history == C6.history and C1.pass_log == C6.pass_log
    assert r1["mu"] == r6["mu"] and r1["energy"] == r6["energy"]
    assert np.array_equal(r1["density"].cpu().numpy(), r6["density"].cpu().numpy())


def test_single_warm_step_equals_host_step(sys3):
    """One warm density() call after a converged SCF, on a perturbed potential that
    needs a mu re-solve: same mu, density and band energy bit for bit."""
    print("[gpu] single device warm step vs host step, n=3 fp64", flush=True)
    import torch
    from lresdft_amd.lcao3d import LRFTLCAOCpp
    cr, b, cells, g, vion = sys3
    out = {}
    for dev in (True, False):
        C = LRFTLCAOCpp(b, cells, R_w=8.0, M=32, device_step=dev)
        r = _scf(sys3, "fp64", device_step=dev)[0]
        v = g.potential(r["dens_in"], vion)
        C.density_device(v)                      # cold step at this potential
        C.count_tol = 1e-10
        n, e, mu = C.density_device(1.002 * v)   # warm: the count moves, mu is re-solved
        out[dev] = (n.cpu().numpy(), e, mu, C.history[-1], C.pass_log[-1])
        torch.cuda.synchronize()
    assert out[True][3] == out[False][3] == "warm-resolve"
    assert out[True][2] == out[False][2] and out[True][1] == out[False][1] and out[True][4] == out[False][4]
    a, h = out[True][0], out[False][0]
    bad = np.flatnonzero(~(a == h))
    assert bad.size == 0, (f"{bad.size} points differ (first {bad[:5]}), NaN {np.isnan(a).sum()} / {np.isnan(h).sum()}, "
                           f"max |diff| {np.nanmax(np.abs(a - h))}, values {a[bad[:3]]} vs {h[bad[:3]]}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_keep_edge_bound_changes_no_decision(sys3, precision, monkeypatch):
    """The keep-edge test decided from the potential-change bound (every window
    eigenvalue moves by at most max |v - v_ref| since the last lambda_min) where it
    can be: the same decisions as the two inertia counts, so the SCF is the same bit
    for bit, and most warm steps need no counts."""
    print(f"[gpu] keep-edge bound on and off, SCF n=3 R_w 8 M 32 {precision}", flush=True)
    monkeypatch.setenv("LRES_EDGE_BOUND", "0")
    r0, C0 = _scf(sys3, precision)
    assert C0.plan.edge_stats()[1] == 0
    monkeypatch.setenv("LRES_EDGE_BOUND", "1")
    r1, C1 = _scf(sys3, precision)
    checked, skipped = C1.plan.edge_stats()
    print(f"    warm keep-edge tests {checked}, decided by the bound {skipped}; histories {C1.history}", flush=True)
    assert checked > 0 and skipped > 0
    assert C0.history == C1.history and C0.pass_log == C1.pass_log
    assert r0["iterations"] == r1["iterations"]
    assert r0["mu"] == r1["mu"] and r0["energy"] == r1["energy"]
    assert np.array_equal(r0["density"].cpu().numpy(), r1["density"].cpu().numpy())


@pytest.mark.parametrize("n", [3, 8])
def test_device_lambda_min_equals_host_search(n, monkeypatch):
    """lambda_min searched on the device (probe, bracket and multisection rounds decided
    by controller kernels, one wait) against the host-driven search: the same bracket
    arithmetic on the same counts, so the same value bit for bit, at the guess
    potential and at a shifted one; both bracket the smallest window eigenvalue
    (inertia counts on either side)."""
    print(f"[gpu] device vs host lambda_min, {n}^3 H2 lattice, R_w 8", flush=True)
    import torch
    from lresdft_amd import system
    from lresdft_amd.lcao3d import LRFTLCAOCpp
    from lresdft_amd.scf import GridOps
    cr = system.h2_crystal((n, n, n), a=6.0, pts_per_cell=10, rc=0.6)
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    g = GridOps(cr)
    vion = g.ion_potential()
    C = LRFTLCAOCpp(b, system.cell_centres((n, n, n), 6.0), R_w=8.0, M=32)
    assert len(C.plan.buckets()) == 1
    dens = torch.full(cr.shape, cr.n_elec / (cr.ng * cr.dV), dtype=torch.float64, device="cuda")
    v = g.potential(dens, vion).reshape(-1).contiguous()
    for shift in (0.0, 0.3):
        C.plan.set_potential(v + shift)
        monkeypatch.setenv("LRES_LMIN", "host")
        lh = C.plan.lambda_min()
        monkeypatch.setenv("LRES_LMIN", "device")
        ld = C.plan.lambda_min()
        print(f"    shift {shift}: lambda_min {ld!r} (host {lh!r})", flush=True)
        assert ld == lh
        lo, hi = C.plan.inertia([ld - 1e-9, ld + 1e-9])
        assert lo == 0 and hi > 0
