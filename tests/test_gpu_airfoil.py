"""A NACA 0012 section in a periodic span (crates/cfd/examples/airfoil.rs)
at +-4 degrees, Re 1000, 48 cells per chord: lift follows the angle's sign
with the same magnitude both ways (the tunnel is mirror-symmetric about the
chord line), the drag does not change with the sign, and both lie where a
thin low-Reynolds-number section puts them (lift under thin-airfoil theory's
2 pi alpha)."""
import math

import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def _coeffs(alpha):
    setup, geom, meta = W.naca0012(48.0, alpha, 1000.0)
    lat = sa.Lattice(setup, geom)
    steps = int(10 * meta["steps_per_pass"])
    h = lat.run_still(steps)
    k = steps // 2
    return h[k:, 1].mean() / meta["q"], h[k:, 0].mean() / meta["q"]


def test_naca0012_lift_and_drag():
    print("case NACA 0012, 48 cells per chord, Re 1000, +-4 degrees", flush=True)
    clp, cdp = _coeffs(4.0)
    clm, cdm = _coeffs(-4.0)
    print(f"  +4: Cl {clp:.4f} Cd {cdp:.4f};  -4: Cl {clm:.4f} Cd {cdm:.4f}", flush=True)
    assert clp > 0.0 > clm
    assert abs(clp + clm) < 0.005 * abs(clp)
    assert abs(cdp / cdm - 1.0) < 0.005
    assert 0.1 < clp < 2.0 * math.pi * math.radians(4.0)
    assert 0.05 < cdp < 0.3
