"""Pins the oracle's Spalding wall model (shader.rs:546-597) on the CPU: the
wall normal and distance from a cell's cut links, and the friction velocity
against Spalding's law of the wall solved independently (scipy brentq):
y+ = u+ + exp(-kappa B) (exp(kappa u+) - 1 - kappa u+ - (kappa u+)^2 / 2 -
(kappa u+)^3 / 6), kappa = 0.41, B = 5.2, u+ = u_t / u*, y+ = y u* / nu; the
cell's tau is 0.5 + 3 max(u*^2 y / u_t, nu), at least the LES tau."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import pyoracle as O

C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [-1, -1, 0],
                [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1],
                [0, 1, -1], [0, -1, 1]], float)
KAP, B = 0.41, 5.2


def plane_links(n, d):
    """Mask and q of a cell at distance d above a plane with unit normal n
    (pointing into the fluid): every link that crosses the plane, q the
    crossing's fraction of the link."""
    n = np.asarray(n, float) / np.linalg.norm(n)
    mask, q = 0, np.ones(19, np.float32)
    for i in range(1, 19):
        cn = C19[i] @ n
        if cn < 0 and -cn >= d - 1e-12:
            mask |= 1 << i
            q[i] = d / -cn
    return mask, q


def spalding_tau(ut, y, nu):
    def g(us):
        up, yp = ut / us, y * us / nu
        k = KAP * up
        return up + math.exp(-KAP * B) * (math.exp(k) - 1 - k - k * k / 2 - k ** 3 / 6) - yp
    us = brentq(g, KAP * ut / 600.0, 10.0 * ut, xtol=1e-16, rtol=1e-14)
    return 0.5 + 3.0 * max(us * us * y / ut, nu), us


# y+ from 1 to 77: the range where the model's 12 Newton steps from the
# viscous-sublayer guess converge (shader.rs:583-593; by y+ ~ 150 they stop
# short of the root, as the reference's do)
@pytest.mark.parametrize("tau0,ut,d", [(0.5012, 0.05, 0.5), (0.5001, 0.08, 0.5), (0.5006, 0.03, 0.25),
                                       (0.52, 0.01, 0.5)])
def test_flat_wall_matches_spalding(tau0, ut, d):
    mask, q = plane_links((0, 0, 1), d)
    # the five links toward -z cross at q = d; the estimated normal is -z
    # and the wall distance d
    assert bin(mask).count("1") == 5 and np.allclose(q[[6, 10, 12, 15, 17]], d)
    nu = (tau0 - 0.5) / 3.0
    want, us = spalding_tau(ut, d, nu)
    # a normal velocity component does not count, only the tangential one
    got = O.wall_tau_f64(mask, q, (0, 0, 0), (ut * 0.6, ut * 0.8, 0.02), tau0, (0, 0, 0), tau0)
    assert got - 0.5 == pytest.approx(want - 0.5, rel=2e-4)
    if tau0 == 0.52:
        # viscous sublayer: u+ ~ y+, the molecular viscosity stands
        assert us * d / nu < 5 and abs((got - 0.5) / (tau0 - 0.5) - 1.0) < 1e-3
    else:
        assert got - 0.5 > 1.02 * (tau0 - 0.5)   # the wall raises the viscosity


def test_les_tau_and_moving_walls():
    mask, q = plane_links((0, 0, 1), 0.5)
    tau0 = 0.5012
    w, _ = spalding_tau(0.05, 0.5, (tau0 - 0.5) / 3.0)
    # a larger LES tau wins
    assert O.wall_tau_f64(mask, q, (0, 0, 0), (0.05, 0, 0), w + 0.1, (0, 0, 0), tau0) == pytest.approx(w + 0.1)
    # a wall moving with the fluid: no slip velocity, the LES tau
    om = (0.0, 0.0, 0.002)
    pc = (10.0, 0.0, 3.0)
    wall_pt = np.array(pc) + np.array([0.0, 0.0, -0.5])
    uw = np.cross(om, wall_pt)
    assert O.wall_tau_f64(mask, q, pc, uw, 0.6, om, tau0) == pytest.approx(0.6)
    # the slip is measured against the wall's velocity at the wall point
    got = O.wall_tau_f64(mask, q, pc, uw + np.array([0.05, 0.0, 0.0]), tau0, om, tau0)
    assert got == pytest.approx(w, rel=1e-9)


def test_cells_far_from_the_wall_keep_their_tau():
    # no cut link: the LES tau; links only nearly parallel to the wall
    # (c.n <= 0.3) cannot give the distance
    assert O.wall_tau_f64(0, np.ones(19, np.float32), (0, 0, 0), (0.05, 0, 0), 0.55, (0, 0, 0), 0.5012) == 0.55


@pytest.mark.parametrize("deg", [10.0, 20.0, 30.0])
def test_tilted_wall_normal(deg):
    # the q-weighted link normal follows a tilted wall: the wall viscosity
    # comes out as for the true normal and distance within 5 %
    tilt = math.radians(deg)
    n = np.array([math.sin(tilt), 0.0, math.cos(tilt)])
    d = 0.45
    mask, q = plane_links(n, d)
    tau0 = 0.5012
    ut = 0.05
    t = np.cross(n, [0.0, 1.0, 0.0])
    t /= np.linalg.norm(t)
    got = O.wall_tau_f64(mask, q, (0, 0, 0), ut * t, tau0, (0, 0, 0), tau0)
    want, _ = spalding_tau(ut, d, (tau0 - 0.5) / 3.0)
    assert abs((got - 0.5) / (want - 0.5) - 1.0) < 0.05
