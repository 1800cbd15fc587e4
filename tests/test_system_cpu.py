"""The product's geometry/basis builders (lresdft_amd.system) against the oracle's
restatement of lrft/three_d.py and lrft/lcao3d.py, and the GPU-side SCF
plumbing's pieces that run on CPU tensors."""
import numpy as np
import pytest

from lresdft_amd import system
from oracle import lres_oracle as orc


@pytest.mark.parametrize("n,a,pts", [((2, 2, 2), 6.0, 10), ((3, 2, 2), 5.1, 8)])
def test_h2_crystal_identical(n, a, pts):
    c1 = system.h2_crystal(n, a=a, pts_per_cell=pts, rc=0.6)
    c2 = orc.h2_crystal(n, a=a, pts_per_cell=pts, rc=0.6)
    assert np.array_equal(c1.pos, c2.pos) and c1.shape == c2.shape and np.array_equal(c1.cell, c2.cell)


def test_basis_matrices_match_oracle():
    cr = system.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6)
    b1 = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    b2 = orc.LCAOBasis(orc.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6), **orc.BASIS)
    assert np.array_equal(b1.atom_of_fn, np.repeat(np.arange(len(cr.pos)), 6))
    assert abs(b1.X - b2.X).max() < 1e-14
    assert abs(b1.S - b2.S).max() < 1e-13
    assert abs(b1.T - b2.T).max() < 1e-12
    v = orc.guess_v(orc.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6))
    assert abs(b1.H(v) - b2.H(v)).max() < 1e-12


def test_cell_centres_match():
    assert np.array_equal(system.cell_centres((3, 2, 4), 6.0), orc.cell_centres((3, 2, 4), 6.0))


def test_random_box_min_separation():
    cr, centres = system.random_h2_box(40, 20.0, dmin=3.0, seed=3)
    L = cr.cell
    for i in range(len(centres)):
        d = centres - centres[i]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(1))
        r[i] = np.inf
        assert r.min() >= 3.0
    assert cr.n_elec == 80 and cr.shape == (33, 33, 33)


def test_grid_ops_match_oracle_on_cpu_tensors():
    import torch
    from lresdft_amd.scf import GridOps
    cr = system.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6)
    cro = orc.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6)
    g = GridOps(cr, device="cpu")
    vi = g.ion_potential(chunk=5).numpy()
    assert np.max(np.abs(vi - orc.ion_potential(cro))) < 1e-12
    rng = np.random.default_rng(0)
    n = rng.uniform(0.001, 0.2, cr.shape)
    assert np.max(np.abs(g.hartree(torch.from_numpy(n)).numpy() - orc.hartree(cro, n))) < 1e-12
    e1, v1 = g.lda_xc(torch.from_numpy(n))
    e2, v2 = orc.lda_xc(n)
    assert np.max(np.abs(e1.numpy() - e2)) < 1e-14 and np.max(np.abs(v1.numpy() - v2)) < 1e-14


def test_basis_derivatives_match_finite_differences():
    """d x_c(g) / d R_A (analytic, with the grid-quadrature normalisation) vs central
    differences of the basis itself: 1e-8 absolute (values are O(1))."""
    from lresdft_amd import system
    cr = system.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6)
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    dX = b.derivatives()
    h = 1e-5
    for A in (0, 7):
        cols = np.where(b.atom_of_fn == A)[0]
        for i in range(3):
            pp, pm = cr.pos.copy(), cr.pos.copy()
            pp[A, i] += h
            pm[A, i] -= h
            bp = system.LCAOBasis(system.Crystal(pp, cr.Z, cr.cell, cr.shape, cr.rc), **system.DEFAULT_BASIS)
            bm = system.LCAOBasis(system.Crystal(pm, cr.Z, cr.cell, cr.shape, cr.rc), **system.DEFAULT_BASIS)
            fd = ((bp.X - bm.X) / (2 * h)).toarray()
            an = dX[i].toarray()
            assert np.max(np.abs(fd[:, cols] - an[:, cols])) < 1e-8
            others = b.atom_of_fn != A
            assert np.max(np.abs(fd[:, others])) == 0.0


def test_ion_forces_match_finite_differences():
    """GridOps.ion_forces (the ion-potential part of the force, -dV sum n dv_ion/dR_A)
    against central differences of dV sum n v_ion(R) with the oracle's per-atom
    ion potential (three_d.py:102-119), on CPU tensors: step 1e-4 bohr, every
    component of four atoms, relative 1e-7."""
    import torch

    from lresdft_amd.scf import GridOps
    cr = system.h2_crystal((2, 2, 2), a=6.0, pts_per_cell=10, rc=0.6)
    rng = np.random.default_rng(4)
    n = 0.01 + rng.random(cr.shape)
    g = GridOps(cr, device="cpu")
    F = g.ion_forces(torch.from_numpy(n)).numpy()
    h = 1e-4
    for A in (0, 3, 8, 15):
        for ax in range(3):
            e = []
            for s in (1, -1):
                pos = cr.pos.copy()
                pos[A, ax] += s * h
                c2 = orc.Crystal(pos, cr.Z, cr.cell, cr.shape, cr.rc)
                e.append(np.sum(n * orc.ion_potential(c2)) * cr.dV)
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(F[A, ax] - fd) <= 1e-7 * max(1.0, abs(fd)), (A, ax, F[A, ax], fd)
