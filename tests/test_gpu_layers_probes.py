"""The closed forms of test_oracle_layers_probes.py (absorbing layer weight,
density, velocity and viscosity relaxation; probe interpolation) through the
HIP lattice, in FP64 and FP32 storage."""
import numpy as np
import pytest

import simlab_amd as sa
from test_oracle_layers_probes import at_rest, mix_x, setup_periodic, sheared, shear_of

pytestmark = pytest.mark.gpu

STORE = [(sa.Storage.FP64, np.float64, 1e-13), (sa.Storage.FP32, np.float32, 2e-6)]


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("storage,dt,tol", STORE)
def test_absorbing_layer_density(storage, dt, tol):
    _say(f"case layer density {storage.name}")
    dims, w, sigma, rho0 = (16, 6, 6), 12, 0.2, 1.03
    su = setup_periodic(dims, sponge=(w, 0, 0, 0, 0, 0), sponge_tau=0.9, sponge_sigma=sigma)
    su.storage = storage
    lat = sa.Lattice(su)
    lat.set_pops(at_rest(np.full(dims[0], rho0), dims[1], dims[2]).astype(dt))
    sp = np.clip(1.0 - np.arange(dims[0]) / w, 0.0, 1.0)
    lat.run_still(1)
    r1 = rho0 + sigma * sp ** 2 * (1.0 - rho0)
    got = lat.macros()[:, 0].reshape(dims[2], dims[1], dims[0]).astype(np.float64)
    assert np.allclose(got, r1[None, None, :], rtol=0, atol=max(tol, 1e-7))
    lat.run_still(1)
    m = mix_x(r1)
    got2 = lat.macros()[:, 0].reshape(dims[2], dims[1], dims[0]).astype(np.float64)
    assert np.allclose(got2, (m + sigma * sp ** 2 * (1.0 - m))[None, None, :], rtol=0, atol=max(tol, 1e-7))


@pytest.mark.parametrize("storage,dt,tol", STORE)
def test_absorbing_layer_velocity_and_tau(storage, dt, tol):
    _say(f"case layer velocity and tau {storage.name}")
    dims, w, sigma = (16, 6, 6), 10, 0.3
    x = np.arange(dims[0], dtype=float)
    sp = np.clip(1.0 - x / w, 0.0, 1.0)
    su = setup_periodic(dims, sponge=(w, 0, 0, 0, 0, 0), sponge_tau=0.9, sponge_sigma=sigma, u_init=(0.02, 0.0, 0.0),
                        u_inf=(0.05, 0.0, 0.0), absorb_velocity=True)
    su.storage = storage
    lat = sa.Lattice(su)
    lat.run_still(1)
    mac = lat.macros().reshape(dims[2], dims[1], dims[0], 4).astype(np.float64)
    assert np.allclose(mac[..., 1], (0.02 + sigma * sp ** 2 * 0.03)[None, None, :], rtol=0, atol=max(tol, 1e-7))
    # the viscosity ramp
    su2 = setup_periodic(dims, sponge=(w, 0, 0, 0, 0, 0), sponge_tau=0.9, sponge_sigma=0.2)
    su2.storage = storage
    lat2 = sa.Lattice(su2)
    n = int(np.prod(dims))
    pxy = 2e-4
    lat2.set_pops(sheared(pxy, n).astype(dt))
    lat2.run_still(1)
    got = shear_of(lat2.get_pops().astype(np.float64), n).reshape(dims[2], dims[1], dims[0])
    tau = 0.65 + sp ** 2 * (0.9 - 0.65)
    assert np.allclose(got, (1.0 - 1.0 / tau)[None, None, :] * pxy, rtol=1e-5, atol=1e-11)


@pytest.mark.parametrize("storage,dt,tol", STORE)
def test_samplers_read_a_linear_field_exactly(storage, dt, tol):
    _say(f"case samplers {storage.name}")
    dims, c, nrm, pb = (20, 10, 10), (10.0, 5.0, 5.0), (0.6, 0.8, 0.0), (-2.3, 0.4, -0.7)
    probes = sa.Probes(verts=[(pb, nrm)], slices=[sa.SliceDef((1, 5, 5), (1, 0, 0), (0, 1, 0), 18, 1)],
                       vertex_offset=1.0)
    su = setup_periodic(dims, center=c)
    su.storage = storage
    lat = sa.Lattice(su, None, probes)
    a, b = 1.0, 0.01
    lat.set_pops(at_rest(a + b * np.arange(dims[0], dtype=float), dims[1], dims[2]).astype(dt))
    lat.run([sa.StepIn.still(False), sa.StepIn.still(True)])
    assert lat.vertex_means()[0] == pytest.approx(a + b * (c[0] + pb[0] + nrm[0]), abs=2e-6)
    sm = lat.slice_means()
    rho = a + b * np.arange(1, 19, dtype=float)
    assert np.allclose(sm[:, 0], rho, atol=2e-6)
    assert np.allclose(sm[:, 1], -b / 3.0 / rho, rtol=1e-4, atol=1e-8)
