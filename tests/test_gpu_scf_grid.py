"""SCF grid kernels of the C-ABI: LDA potential, residual + DIIS Gram row, Harris integrals.

Bars:
  lres_lda_potential vs the oracle's lda_xc (numpy, three_d.py:132-148) and vs the
  torch formula of GridOps.lda_xc: 1e-13 of max|v| (the device cbrt/log against
  numpy's pow/log differ by a few ulp), over densities from 1e-16 to 30 (both PZ81
  branches and the clamp at 1e-14);
  lres_scf_residual: the stored rows bitwise (r = 2 n_out - n_in is one rounding),
  untouched rows unchanged, sum |r| and the Gram row R_q . r within 1e-13 of the
  sum of |terms| of numpy's sums, at a ragged grid size and history sizes 1, 6, 8;
  lres_harris_terms vs numpy with the oracle's lda_xc: 1e-13 of the sum of |terms|;
  every kernel gives the same bits on a repeated call (fixed-order sums).
"""
import numpy as np
import pytest

from oracle.lres_oracle import lda_xc

pytestmark = pytest.mark.gpu


def _stream():
    import torch
    from lresdft_amd import _native as nat
    return nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _densities(rng, m):
    n = 10.0 ** rng.uniform(-16, 1.5, m)
    n[:5] = [0.0, 1e-14, 5e-15, 1e-13, 30.0]
    # keep clear of rs = 1 (n = 3 / 4 pi), where the two PZ81 branches meet
    edge = 3.0 / (4.0 * np.pi)
    n[np.abs(n - edge) < 1e-9] = 0.5
    return n


def test_lda_potential_matches_oracle_and_torch():
    import torch
    from lresdft_amd import _native as nat
    print("[gpu] lres_lda_potential vs oracle lda_xc, 200003 points", flush=True)
    rng = np.random.default_rng(11)
    m = 200003
    n = _densities(rng, m)
    vion = rng.standard_normal(m)
    vh = rng.standard_normal(m)
    dn, dvi, dvh = (torch.from_numpy(a).cuda() for a in (n, vion, vh))
    out = torch.empty_like(dn)
    nat.check(nat.lib().lres_lda_potential(dn.data_ptr(), dvi.data_ptr(), dvh.data_ptr(), m, out.data_ptr(),
                                           _stream()))
    got = out.cpu().numpy()
    ref = vion + vh + lda_xc(n)[1]
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) < 1e-13 * scale
    tref = (dvi + dvh + _torch_lda(dn)).cpu().numpy()
    assert np.max(np.abs(got - tref)) < 1e-13 * scale
    out2 = torch.empty_like(dn)
    nat.check(nat.lib().lres_lda_potential(dn.data_ptr(), dvi.data_ptr(), dvh.data_ptr(), m, out2.data_ptr(),
                                           _stream()))
    assert torch.equal(out, out2)


def _torch_lda(n):
    import torch
    from lresdft_amd.scf import GridOps
    g = GridOps.__new__(GridOps)
    g.torch = torch
    return g.lda_xc(n)[1]


@pytest.mark.parametrize("nh,slot", [(1, 0), (6, 3), (8, 7)])
def test_scf_residual_rows_and_gram(nh, slot):
    import torch
    from lresdft_amd import _native as nat
    print(f"[gpu] lres_scf_residual nh={nh} slot={slot}, ng=100003", flush=True)
    rng = np.random.default_rng(nh)
    ng = 100003
    nout = rng.uniform(0.0, 1.0, ng)
    nin = rng.uniform(0.0, 1.0, ng)
    Rh = rng.standard_normal((nh, ng))
    Nh = rng.standard_normal((nh, ng))
    if nh > 2:
        Rh[nh - 2] = 0.0          # a history row not in use holds zeros
    dR, dN = torch.from_numpy(Rh.copy()).cuda(), torch.from_numpy(Nh.copy()).cuda()
    dout, din = torch.from_numpy(nout).cuda(), torch.from_numpy(nin).cuda()
    part = torch.empty((nat.LRES_MAX_HIST + 1) * nat.LRES_MIX_PARTS, dtype=torch.float64, device="cuda")
    stats = torch.empty(nh + 1, dtype=torch.float64, device="cuda")

    def call():
        nat.check(nat.lib().lres_scf_residual(dout.data_ptr(), 2.0, din.data_ptr(), ng, nh, slot, dN.data_ptr(),
                                              dR.data_ptr(), part.data_ptr(), stats.data_ptr(), _stream()))
        return stats.cpu().numpy().copy()

    s1 = call()
    r = 2.0 * nout - nin
    R2, N2 = dR.cpu().numpy(), dN.cpu().numpy()
    assert np.array_equal(R2[slot], r) and np.array_equal(N2[slot], nin)
    others = [q for q in range(nh) if q != slot]
    assert np.array_equal(R2[others], Rh[others]) and np.array_equal(N2[others], Nh[others])
    Rn = Rh.copy()
    Rn[slot] = r
    ref = np.concatenate([[np.sum(np.abs(r))], Rn @ r])
    mag = np.concatenate([[np.sum(np.abs(r))], np.abs(Rn) @ np.abs(r)])     # sums of |terms|
    assert np.all(np.abs(s1 - ref) <= 1e-13 * mag), (s1, ref)
    if nh > 2:
        assert s1[1 + nh - 2] == 0.0
    s2 = call()
    assert np.array_equal(s1, s2)


def test_harris_terms_match_numpy():
    import torch
    from lresdft_amd import _native as nat
    print("[gpu] lres_harris_terms vs numpy, 150001 points", flush=True)
    rng = np.random.default_rng(5)
    m = 150001
    n = _densities(rng, m)
    vh = rng.standard_normal(m)
    dn, dvh = torch.from_numpy(n).cuda(), torch.from_numpy(vh).cuda()
    part = torch.empty(3 * nat.LRES_MIX_PARTS, dtype=torch.float64, device="cuda")
    out = torch.empty(3, dtype=torch.float64, device="cuda")
    nat.check(nat.lib().lres_harris_terms(dn.data_ptr(), dvh.data_ptr(), m, part.data_ptr(), out.data_ptr(),
                                          _stream()))
    got = out.cpu().numpy().copy()
    exc, vxc = lda_xc(n)
    ref = np.array([np.sum(n * vh), np.sum(n * exc), np.sum(n * (vh + vxc))])
    mag = np.array([np.sum(np.abs(n * vh)), np.sum(np.abs(n * exc)), np.sum(np.abs(n * (vh + vxc)))])
    assert np.all(np.abs(got - ref) <= 1e-13 * mag), (got, ref)
    nat.check(nat.lib().lres_harris_terms(dn.data_ptr(), dvh.data_ptr(), m, part.data_ptr(), out.data_ptr(),
                                          _stream()))
    assert np.array_equal(got, out.cpu().numpy())
