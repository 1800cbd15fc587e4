"""Matrix-core (MFMA) pole kernel: parity with the vector-ALU kernel and the oracle.

Bars:
  fp64 MFMA pass vs the oracle pole pass and vs the vector-ALU kernel: density,
  N_B and E_B max-abs 1e-12 (same bar as the vector-ALU kernel's own parity);
  fp32 MFMA pass vs the oracle: density max-abs 2e-5 and sum N_B within 1e-5 N
  (complex64 arithmetic over a window of 84 functions);
  every waves-per-matrix variant gives bitwise the same pass (the autotuner's
  choice never changes results);
  M = 20 (10 poles: partial load batches in the reduction) vs the oracle in both
  precisions;
  uneven windows (random centres, n_W 36..96 in one plan, six buckets) vs the
  oracle at 1e-12 in fp64 (at most one bucket may fall back to the vector-ALU
  kernel when every MFMA variant of its size spills) and at the fp32 bar in
  complex64 (all buckets on the matrix cores).
Windows of up to 128 functions run on the matrix cores in both precisions by
default; the tests still force the family they check (LRES_POLE_KERNEL=mfma /
=valu).  Inertia counts: the matrix-core block elimination equals the
vector-ALU kernel and the eigenvalue counts of the oracle's dense solves
exactly, and lambda_min from both agrees with the dense lowest eigenvalue
to 1e-10.
"""
import os

import numpy as np
import pytest
from scipy import linalg

from oracle.lres_oracle import BASIS, LCAOBasis, build_windows, cell_centres, guess_v, h2_crystal, pole_window_pass

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys3():
    cr = h2_crystal((3, 3, 3), a=6.0, pts_per_cell=10, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    return cr, b, guess_v(cr)


def _mu_window(H, S, W, margin=0.3, above=0.35):
    lam_min = min(linalg.eigh(H[w][:, w].toarray(), S[w][:, w].toarray(), eigvals_only=True)[0] for w in W)
    return lam_min - margin, lam_min + above


def _pass(b, cells, H, mu, e_lo, M=32, precision="fp64", Rw=8.0, env=None):
    from lresdft_amd.lcao3d import LRFTLCAOCpp, csr64
    old = {k: os.environ.get(k) for k in (env or {})}
    os.environ.update(env or {})
    try:
        C = LRFTLCAOCpp(b, cells, R_w=Rw, M=M, precision=precision)
        C.plan.set_h(csr64(H))
        nsum, esum = C.plan.run(mu, e_lo, want_density=True)
        out = (C.plan.dens.cpu().numpy().copy(), C.plan.NB.cpu().numpy().copy(), C.plan.EB.cpu().numpy().copy(),
               nsum, esum, C.plan.buckets())
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out


def test_mfma_fp64_matches_valu_and_oracle(sys3):
    cr, b, v = sys3
    print("[gpu] mfma fp64 vs valu vs oracle, n=3 R_w=8 M=32", flush=True)
    cells = cell_centres((3, 3, 3), 6.0)
    H = b.H(v)
    W, own = build_windows(b, cells, 8.0)
    e_lo, mu = _mu_window(H, b.S, W)
    dm, NBm, EBm, _, _, bk = _pass(b, cells, H, mu, e_lo, env={"LRES_POLE_KERNEL": "mfma"})
    assert all(row[4] == 2 for row in bk), bk          # the matrix-core kernel ran
    dv, NBv, EBv, _, _, bkv = _pass(b, cells, H, mu, e_lo, env={"LRES_POLE_KERNEL": "valu"})
    assert all(row[4] != 2 for row in bkv), bkv
    dp, NBo, EBo = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, 32, True)
    assert np.max(np.abs(dm - dp)) < 1e-12
    assert np.max(np.abs(NBm - NBo)) < 1e-12 and np.max(np.abs(EBm - EBo)) < 1e-12
    assert np.max(np.abs(dm - dv)) < 1e-12
    assert np.max(np.abs(NBm - NBv)) < 1e-12 and np.max(np.abs(EBm - EBv)) < 1e-12


def test_mfma_fp32_error_is_fp32_level(sys3):
    cr, b, v = sys3
    print("[gpu] mfma fp32 vs oracle, n=3 R_w=8 M=32", flush=True)
    cells = cell_centres((3, 3, 3), 6.0)
    H = b.H(v)
    W, own = build_windows(b, cells, 8.0)
    e_lo, mu = _mu_window(H, b.S, W)
    d32, NB32, EB32, n32, e32, bk = _pass(b, cells, H, mu, e_lo, precision="mixed")
    assert all(row[7] == 2 for row in bk), bk
    dp, NBo, EBo = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, 32, True)
    assert np.max(np.abs(d32 - dp)) < 2e-5
    assert abs(n32 - NBo.sum()) < 1e-5 * cr.n_elec
    assert np.max(np.abs(d32 - dp)) > 0.0      # really the complex64 path


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pole_count_not_a_multiple_of_the_load_batch(sys3, precision):
    """M = 20: 10 poles, so the reduction's batches of 8 per-pole slots (complex64)
    and of 4 pole-group partials (fp64, 10 groups at 27 windows) end in a partial
    batch; fp64 vs the oracle at 1e-12, complex64 at the fp32 bar."""
    cr, b, v = sys3
    print(f"[gpu] M=20 pass {precision} vs oracle, n=3 R_w=8", flush=True)
    cells = cell_centres((3, 3, 3), 6.0)
    H = b.H(v)
    W, own = build_windows(b, cells, 8.0)
    e_lo, mu = _mu_window(H, b.S, W)
    dp, NBo, EBo = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, 20, True)
    d, NB, EB, ns, es, _ = _pass(b, cells, H, mu, e_lo, M=20, precision=precision)
    tol = 1e-12 if precision == "fp64" else 2e-5
    assert np.max(np.abs(d - dp)) < tol
    assert np.max(np.abs(NB - NBo)) < tol and np.max(np.abs(EB - EBo)) < (tol if precision == "fp64" else 1e-4)
    assert abs(ns - NBo.sum()) < (1e-11 if precision == "fp64" else 1e-5 * cr.n_elec)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_wave_split_gives_the_same_bits(sys3, precision):
    cr, b, v = sys3
    print(f"[gpu] mfma wave splits {precision}", flush=True)
    cells = cell_centres((3, 3, 3), 6.0)
    H = b.H(v)
    W, _ = build_windows(b, cells, 8.0)
    e_lo, mu = _mu_window(H, b.S, W)
    ref = None
    for nwc in (1, 2, 3, 4, 5, 6):
        d, NB, EB, ns, es, bk = _pass(b, cells, H, mu, e_lo, precision=precision,
                                      env={"LRES_MFMA_WAVES": str(nwc), "LRES_POLE_KERNEL": "mfma"})
        col = 3 if precision == "fp64" else 6
        assert all(row[col] == nwc for row in bk), (nwc, bk)
        if ref is None:
            ref = (d, NB, EB, ns, es)
        else:
            assert np.array_equal(d, ref[0]) and np.array_equal(NB, ref[1]) and np.array_equal(EB, ref[2]), nwc
            assert ns == ref[3] and es == ref[4]


def test_fp64_group_sums_in_lds_equal_global_read_modify_write():
    """216 H2 (n_W 84, |T_B| <= 84), M 64: 10 pole groups per window, 3-4 poles each, so the
    fp64 matrix-core kernel accumulates each group's D / F partial sums in LDS and stores them
    once.  The pass equals the one with the read-modify-write of global memory per pole
    (LRES_POLE_ACC=global) bit for bit, at three wave splits."""
    print("[gpu] fp64 group sums in LDS vs global, n=6 M=64", flush=True)
    cr = h2_crystal((6, 6, 6), a=6.0, pts_per_cell=10, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    v = guess_v(cr)
    cells = cell_centres((6, 6, 6), 6.0)
    H = b.H(v)
    W, _ = build_windows(b, cells, 8.0)
    e_lo, mu = _mu_window(H, b.S, W[:8])
    ref = None
    for nwc in (2, 3, 4):
        outs = [_pass(b, cells, H, mu, e_lo, M=64, env={"LRES_MFMA_WAVES": str(nwc), "LRES_POLE_KERNEL": "mfma",
                                                        **extra})
                for extra in ({}, {"LRES_POLE_ACC": "global"})]
        for a, g in zip(outs[0][:3], outs[1][:3]):
            assert np.array_equal(a, g), nwc
        assert outs[0][3] == outs[1][3] and outs[0][4] == outs[1][4]
        if ref is None:
            ref = outs[0]
        else:
            assert all(np.array_equal(x, y) for x, y in zip(outs[0][:3], ref[:3])), nwc


def test_mfma_uneven_windows_match_oracle():
    """Random centres: windows of 36 to 96 functions, several buckets and tile counts."""
    from lresdft_amd import system
    cr, centres = system.random_h2_box(27, 6.0 * 3, dmin=3.0, seed=7, dx=0.6, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    v = guess_v(cr)
    H = b.H(v)
    W, own = build_windows(b, centres, 6.0)
    sizes = np.array([len(w) for w in W])
    print(f"[gpu] mfma uneven windows n_W {sizes.min()}..{sizes.max()}", flush=True)
    assert sizes.max() <= 128 and sizes.min() < 64
    e_lo, mu = _mu_window(H, b.S, W)
    dp, NBo, EBo = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, 32, True)
    d, NB, EB, _, _, bk = _pass(b, centres, H, mu, e_lo, Rw=6.0, env={"LRES_POLE_KERNEL": "mfma"})
    assert len(bk) > 1 and sum(row[4] == 2 for row in bk) >= len(bk) - 1, bk
    assert np.max(np.abs(d - dp)) < 1e-12
    assert np.max(np.abs(NB - NBo)) < 1e-12 and np.max(np.abs(EB - EBo)) < 1e-12
    d32, _, _, n32, _, bk32 = _pass(b, centres, H, mu, e_lo, Rw=6.0, precision="mixed")
    assert all(row[7] == 2 for row in bk32), bk32
    assert np.max(np.abs(d32 - dp)) < 2e-5 and abs(n32 - NBo.sum()) < 1e-5 * cr.n_elec


def test_mfma_inertia_counts_match_valu_and_dense(sys3):
    cr, b, v = sys3
    print("[gpu] inertia counts: matrix cores vs vector ALU vs dense, n=3 R_w=8", flush=True)
    from lresdft_amd.lcao3d import LRFTLCAOCpp, csr64
    cells = cell_centres((3, 3, 3), 6.0)
    H = b.H(v)
    W, _ = build_windows(b, cells, 8.0)
    C = LRFTLCAOCpp(b, cells, R_w=8.0, M=32)
    C.plan.set_h(csr64(H))
    assert all(row[0] == 84 for row in C.plan.buckets())
    lam = [linalg.eigh(H[w][:, w].toarray(), b.S[w][:, w].toarray(), eigvals_only=True) for w in W]
    lo = min(l[0] for l in lam)
    sig = np.linspace(lo - 0.2, lo + 1.5, 24)
    # positive pivots of sigma S_W - H_W = eigenvalues of (H_W, S_W) below sigma (Sylvester)
    ref = np.array([sum(int(np.sum(l < s)) for l in lam) for s in sig])
    got_m = C.plan.inertia(sig)
    lm_m = C.plan.lambda_min()
    os.environ["LRES_INERTIA_KERNEL"] = "valu"
    try:
        got_v = C.plan.inertia(sig)
        lm_v = C.plan.lambda_min()
    finally:
        os.environ.pop("LRES_INERTIA_KERNEL", None)
    assert np.array_equal(got_m, ref), (got_m, ref)
    assert np.array_equal(got_v, ref), (got_v, ref)
    assert abs(lm_m - lo) < 1e-10 and abs(lm_v - lo) < 1e-10


def _pass_blocks(b, cells, H, mu, e_lo, M, Rw, fused):
    """complex64 pass with the window-fused form forced on ("1") or left out ("0"): density,
    N_B, E_B, the sums, D_B and F_B blocks and the bucket shapes."""
    from lresdft_amd.lcao3d import LRFTLCAOCpp, csr64
    old = os.environ.get("LRES_POLE_FUSED")
    os.environ["LRES_POLE_FUSED"] = fused
    try:
        C = LRFTLCAOCpp(b, cells, R_w=Rw, M=M, precision="mixed")
        C.plan.set_h(csr64(H))
        nsum, esum = C.plan.run(mu, e_lo, want_density=True)
        D, F = C.plan.pass_blocks()
        return (C.plan.dens.cpu().numpy().copy(), C.plan.NB.cpu().numpy().copy(), C.plan.EB.cpu().numpy().copy(),
                D.cpu().numpy(), F.cpu().numpy(), nsum, esum, C.plan.buckets())
    finally:
        if old is None:
            os.environ.pop("LRES_POLE_FUSED", None)
        else:
            os.environ["LRES_POLE_FUSED"] = old


def _same_bits(a, g):
    for x, y in zip(a[:5], g[:5]):
        assert np.array_equal(x, y)
    assert a[5] == g[5] and a[6] == g[6]


@pytest.mark.parametrize("M", [32, 20])
def test_window_fused_complex64_equals_slot_form(sys3, M):
    """The complex64 window-fused form (four waves per window, poles summed in LDS, the
    reduction in the kernel) gives the pass of one wave per matrix with per-pole slots
    and k_reduce bit for bit: density, N_B, E_B, D_B, F_B.  M = 20: 10 poles, so the
    last round of four has two waves without a pole."""
    cr, b, v = sys3
    print(f"[gpu] complex64 window-fused vs per-pole slots, n=3 R_w 8 M {M}", flush=True)
    cells = cell_centres((3, 3, 3), 6.0)
    H = b.H(v)
    W, own = build_windows(b, cells, 8.0)
    e_lo, mu = _mu_window(H, b.S, W)
    slot = _pass_blocks(b, cells, H, mu, e_lo, M, 8.0, "0")
    fused = _pass_blocks(b, cells, H, mu, e_lo, M, 8.0, "1")
    assert all(row[7] == 2 and row[6] > 0 for row in slot[7]), slot[7]
    assert all(row[7] == 2 and row[6] == 0 for row in fused[7]), fused[7]
    _same_bits(slot, fused)
    dp, NBo, _ = pole_window_pass(H, b.S, b.Xr, W, own, cr.dV, mu, e_lo, M, True)
    assert np.max(np.abs(fused[0] - dp)) < 2e-5 and abs(fused[5] - NBo.sum()) < 1e-5 * cr.n_elec


def test_window_fused_uneven_windows_equal_slot_form():
    """Random centres (windows of 36 to 96 functions, several buckets and tile counts):
    the window-fused form equals the per-pole-slot form bit for bit."""
    from lresdft_amd import system
    cr, centres = system.random_h2_box(27, 6.0 * 3, dmin=3.0, seed=7, dx=0.6, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    v = guess_v(cr)
    H = b.H(v)
    W, _ = build_windows(b, centres, 6.0)
    print(f"[gpu] complex64 window-fused vs slots, uneven windows", flush=True)
    e_lo, mu = _mu_window(H, b.S, W)
    slot = _pass_blocks(b, centres, H, mu, e_lo, 32, 6.0, "0")
    fused = _pass_blocks(b, centres, H, mu, e_lo, 32, 6.0, "1")
    assert len(fused[7]) > 1 and all(row[6] == 0 for row in fused[7]), fused[7]
    _same_bits(slot, fused)
