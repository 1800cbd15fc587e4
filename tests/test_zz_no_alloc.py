"""The allocation-free SCF residual entry point (lres_scf_residual_noalloc) against
lres_scf_residual, through the C-ABI: two calls in a row, the second on a smaller
grid and history, each bit for bit equal to the workspace entry point (the same
kernels and summation order) and to numpy to 1e-12 relative; sizes beyond the
static workspace are rejected with an error."""
import numpy as np
import pytest

from lresdft_amd import _native as nat

pytestmark = pytest.mark.gpu


def _inputs(ng, nh, seed):
    import torch
    rng = np.random.default_rng(seed)
    nout = rng.random(ng)
    nin = rng.random(ng)
    Rh = rng.standard_normal((nh, ng))
    Nh = rng.random((nh, ng))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()   # noqa: E731
    return dev(nout), dev(nin), dev(Nh), dev(Rh), (nout, nin, Nh, Rh)


def _call(noalloc, nout, nin, Nh, Rh, nh, slot):
    import torch
    L = nat.lib()
    ng = nout.numel()
    Nh, Rh = Nh.clone(), Rh.clone()
    stats = torch.full((nh + 1,), float("nan"), dtype=torch.float64, device="cuda")
    st = nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        nat.check(L.lres_scf_residual_noalloc(nout.data_ptr(), 2.0, nin.data_ptr(), ng, nh, slot, Nh.data_ptr(),
                                              Rh.data_ptr(), stats.data_ptr(), st))
    else:
        part = torch.full(((nat.LRES_MAX_HIST + 1) * nat.LRES_MIX_PARTS,), float("nan"), dtype=torch.float64,
                          device="cuda")
        nat.check(L.lres_scf_residual(nout.data_ptr(), 2.0, nin.data_ptr(), ng, nh, slot, Nh.data_ptr(),
                                      Rh.data_ptr(), part.data_ptr(), stats.data_ptr(), st))
    torch.cuda.synchronize()
    return stats.cpu().numpy(), Nh.cpu().numpy(), Rh.cpu().numpy()


def test_noalloc_residual_equals_workspace_entry_point():
    for ng, nh, slot, seed in ((80 ** 3, 6, 2, 1), (30 ** 3, 3, 1, 2)):
        print(f"[gpu] residual noalloc vs workspace: ng {ng}, nhist {nh}, slot {slot}", flush=True)
        nout, nin, Nh, Rh, host = _inputs(ng, nh, seed)
        s_ref, N_ref, R_ref = _call(False, nout, nin, Nh, Rh, nh, slot)
        s_new, N_new, R_new = _call(True, nout, nin, Nh, Rh, nh, slot)
        assert np.array_equal(s_new, s_ref)
        assert np.array_equal(N_new, N_ref) and np.array_equal(R_new, R_ref)
        ho, hi, hN, hR = host
        r = 2.0 * ho - hi
        hR = hR.copy()
        hR[slot] = r
        expect = np.concatenate([[np.abs(r).sum()], hR @ r])
        assert np.allclose(s_new, expect, rtol=1e-12, atol=0.0)
        assert np.array_equal(R_new[slot], r) and np.array_equal(N_new[slot], hi)


def test_noalloc_residual_rejects_oversized_history():
    import torch
    L = nat.lib()
    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    h = torch.zeros((nat.LRES_MAX_HIST + 1) * 16, dtype=torch.float64, device="cuda")
    st = nat.C.c_void_p(torch.cuda.current_stream().cuda_stream)
    with pytest.raises(RuntimeError, match="LRES_MAX_HIST"):
        nat.check(L.lres_scf_residual_noalloc(x.data_ptr(), 2.0, x.data_ptr(), 16, nat.LRES_MAX_HIST + 1, 0,
                                              h.data_ptr(), h.data_ptr(), x.data_ptr(), st))
    with pytest.raises(RuntimeError, match="null buffer"):
        nat.check(L.lres_scf_residual_noalloc(x.data_ptr(), 2.0, x.data_ptr(), 16, 2, 0, h.data_ptr(), h.data_ptr(),
                                              None, st))
