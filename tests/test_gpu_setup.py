"""Window/ownership set-up with the Voronoi owners on the GPU (lres_build_windows_device)
against the host builder and the oracle, bit for bit: lattices whose grid points tie
between cells (27% of the points at a = 6.0), the compressed lattice and random
centres."""
import numpy as np
import pytest

from oracle import lres_oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["lattice3", "lattice4_a51", "random27"])
def test_device_windows_equal_host_and_oracle(case):
    from lresdft_amd import system
    from lresdft_amd.lcao3d import build_windows
    print(f"[gpu] device windows {case}", flush=True)
    if case == "lattice3":
        cr = system.h2_crystal((3, 3, 3), a=6.0, pts_per_cell=10, rc=0.6)
        cells = system.cell_centres((3, 3, 3), 6.0)
    elif case == "lattice4_a51":
        cr = system.h2_crystal((4, 4, 4), a=5.1, pts_per_cell=8, rc=0.6)
        cells = system.cell_centres((4, 4, 4), 5.1)
    else:
        cr, cells = system.random_h2_box(27, 18.0, dmin=3.0, seed=7, dx=0.6, rc=0.6)
    b = system.LCAOBasis(cr, **system.DEFAULT_BASIS)
    dev = build_windows(b, cells, 8.0, device=True)
    host = build_windows(b, cells, 8.0, device=False)
    for x, y in zip(dev, host):
        assert np.array_equal(x, y)
    W, own = orc.build_windows(orc.LCAOBasis(cr, **orc.BASIS), cells, 8.0)
    Wflat, Woff, Oflat, Ooff = dev
    for B in range(len(cells)):
        assert np.array_equal(Wflat[Woff[B]:Woff[B + 1]], W[B])
        assert np.array_equal(Oflat[Ooff[B]:Ooff[B + 1]], own[B])
