"""Fixed-N band-energy forces from the adjoint vs finite differences of the whole path.

E_band(R) = 2 sum_B E_B with the grid potential frozen (the guess potential),
e_lo fixed and mu solved for N at every geometry; the basis, H, S, Q_B and the
windows' T_B sets are rebuilt from the displaced positions.  Ten atoms drawn
from a seeded generator, one random unit direction each; central differences
with steps 1e-4 and 5e-5 bohr, Richardson-extrapolated (truncation O(h^4)).
Bar: |F . u + dE/dh| <= 1e-6 relative to max(|F . u|, 1e-3 Ha/bohr).
"""
import numpy as np
import pytest

from oracle.lres_oracle import guess_v

pytestmark = pytest.mark.gpu


def test_band_forces_match_finite_differences():
    from lresdft_amd import system
    from lresdft_amd.forces import band_forces
    from lresdft_amd.lcao3d import LRFTLCAOCpp, csr64
    n, Rw, M = 3, 8.0, 32
    cr = system.h2_crystal((n, n, n), a=6.0, pts_per_cell=10, rc=0.6)
    cells = system.cell_centres((n, n, n), 6.0)
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    v = guess_v(cr).reshape(-1)
    N = cr.n_elec
    C = LRFTLCAOCpp(b, cells, R_w=Rw, M=M)
    C.density(v)                               # fixed-N mu and e_lo at the reference geometry
    mu, e_lo = C.mu, C.e_lo
    print(f"[gpu] band forces vs FD n=3 mu={mu:.6f} e_lo={e_lo:.6f}", flush=True)
    Fa, parts = band_forces(C, v, mu, e_lo)
    assert abs(parts["count"]) < 1e-6 * N

    def e_band(pos):
        crp = system.Crystal(pos, cr.Z, cr.cell, cr.shape, cr.rc)
        bp = system.LCAOBasis(crp, **system.DEFAULT_BASIS)
        Cp = LRFTLCAOCpp(bp, cells, R_w=Rw, M=M, device_h=False)
        Cp.plan.set_h(csr64(bp.H(v)))
        m, _ = Cp.plan.mu_solve(N, e_lo, e_lo + 1e-6, mu + 1.0, xtol=1e-14, ftol=1e-13 * N)
        return 2.0 * Cp.plan.run(m, e_lo, want_density=False)[1]

    rng = np.random.default_rng(11)
    atoms = rng.choice(len(cr.pos), size=10, replace=False)
    for A in atoms:
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)

        def f(h):
            p = cr.pos.copy()
            p[A] += h * u
            return e_band(p)

        h = 1e-4
        d1 = (f(h) - f(-h)) / (2 * h)
        d2 = (f(h / 2) - f(-h / 2)) / h
        fd = (4 * d2 - d1) / 3
        an = float(Fa[A] @ u)
        print(f"    atom {A}: F.u {an:.10e}  -dE/dh {-fd:.10e}", flush=True)
        assert abs(an + fd) <= 1e-6 * max(abs(an), 1e-3)
