"""CPU checks of the C-ABI library: it loads, exports every entry point of
include/lres.h, and its host-side window/ownership builder equals the oracle's
(lrft/lcao3d.py:132-161 semantics, including Voronoi ties) bit for bit."""
import os
import re

import numpy as np
import pytest

from lresdft_amd import _native as nat
from oracle.lres_oracle import BASIS, LCAOBasis, amorphous_h2, build_windows, cell_centres, h2_crystal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "lres.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(lres_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_header_entry_point():
    names = header_functions()
    assert len(names) >= 14
    L = nat.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in nat.SIGNATURES, n


def test_errors_are_reported_not_thrown():
    L = nat.lib()
    rc = L.lres_plan_destroy(None)
    assert rc == 0
    with pytest.raises(RuntimeError, match="empty geometry"):
        z = np.zeros(3)
        i = np.zeros(3, np.int64)
        nat.check(L.lres_build_windows(nat.pf(z), 0, nat.pf(z), 0, nat.p64(i), 0, nat.pf(z), nat.p64(i), 0.1, 1.0,
                                       None, nat.p64(i), nat.p64(i), nat.p64(i)))


def _native_windows(b, cells, Rw):
    from lresdft_amd.lcao3d import build_windows as native
    Wf, Wo, Of, Oo = native(b, cells, Rw)
    return ([Wf[Wo[i]:Wo[i + 1]] for i in range(len(Wo) - 1)], [Of[Oo[i]:Oo[i + 1]] for i in range(len(Oo) - 1)])


@pytest.mark.parametrize("n,Rw", [(2, 20.0), (3, 8.0), (3, 8.5), (3, 12.1), (4, 10.0)])
def test_native_windows_equal_oracle_lattice(n, Rw):
    cr = h2_crystal((n, n, n), a=6.0, pts_per_cell=10, rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    cells = cell_centres((n, n, n), 6.0)
    Wo, Oo = build_windows(b, cells, Rw)
    Wn, On = _native_windows(b, cells, Rw)
    assert len(Wn) == len(Wo)
    for a, c in zip(Wn, Wo):
        assert np.array_equal(a, c)
    for a, c in zip(On, Oo):          # Voronoi ties (27% of lattice points) must go to the same cell
        assert np.array_equal(a, c)


def test_native_windows_equal_oracle_amorphous():
    cr, centres = amorphous_h2(125, 30.0)
    b = LCAOBasis(cr, **BASIS)
    Wo, Oo = build_windows(b, centres, 8.5)
    Wn, On = _native_windows(b, centres, 8.5)
    for a, c in zip(Wn, Wo):
        assert np.array_equal(a, c)
    for a, c in zip(On, Oo):
        assert np.array_equal(a, c)
    assert abs(np.mean([len(w) for w in Wn]) - 145.632) < 1e-9       # referee_checks.json R3


def test_compressed_lattice_window_sizes():
    """a = 5.1, pts_per_cell = round(5.1/0.6) = 8: R_w 8 -> 19 cells (228), R_w 10 -> 27 cells (324)."""
    cr = h2_crystal((4, 4, 4), a=5.1, pts_per_cell=round(5.1 / 0.6), rc=0.6)
    b = LCAOBasis(cr, **BASIS)
    cells = cell_centres((4, 4, 4), 5.1)
    for Rw, nW in ((8.0, 228), (10.0, 324)):
        Wn, On = _native_windows(b, cells, Rw)
        assert all(len(w) == nW for w in Wn)
