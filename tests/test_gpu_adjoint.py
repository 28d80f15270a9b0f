"""Partial bounce-back and the discrete drag adjoint on the GPU, through the
C ABI. The reference has no porous cells, so nothing of it pins these: the
forward is pinned by its reduction to the plain lattice at phi = 0 and by
momentum balance (force on the porous solid = momentum the fluid loses), the
adjoint by central finite differences of the forward."""
import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def small(storage, seed=7):
    return W.porous_box(dims=(32, 16, 16), box=8, x0=12.0, u=0.04, re=20.0, seed=seed, storage=storage, debug_dump=True)


def test_zero_fraction_is_the_plain_lattice():
    setup, phi, _ = small(sa.Storage.FP64)
    say("case pbb_zero", setup.dims)
    plain = sa.Lattice(setup)
    plain.run_still(24)
    lat = sa.Lattice(setup)
    lat.set_solid_fraction(np.zeros_like(phi))
    j, ft = lat.adjoint_forward(24, 8, 0, 0)
    assert j == 0.0 and np.all(ft == 0.0)
    assert np.array_equal(lat.get_pops(), plain.get_pops())


def test_porous_force_equals_fluid_momentum_loss():
    # periodic box, uniform flow: sum over steps of the force on the porous
    # solid = fluid momentum before - after
    setup, phi, _ = small(sa.Storage.FP64)
    setup.faces = (sa.FaceBc.PERIODIC,) * 6
    say("case pbb_momentum", setup.dims)
    lat = sa.Lattice(setup)
    lat.set_solid_fraction(phi)
    m0 = lat.macros().astype(np.float64)      # initial state (feq of u_init)
    p0 = (m0[:, :1] * m0[:, 1:]).sum(0)
    n = 30
    _, ft = lat.adjoint_forward(n, 10, 0, 0)
    m1 = lat.macros().astype(np.float64)
    p1 = (m1[:, :1] * m1[:, 1:]).sum(0)
    imp = ft[:, :3].sum(0)
    say("  impulse", imp, "momentum loss", p0 - p1)
    assert np.allclose(imp, p0 - p1, rtol=1e-4, atol=1e-6 * np.abs(p0).max())


def fd_cells(phi, k, rng):
    """k cells inside the box with phi in [0.05, 0.95] (room for +-eps)."""
    idx = np.nonzero((phi > 0.05) & (phi < 0.95))[0]
    return [int(c) for c in rng.choice(idx, k, replace=False)]


def drag(lat, p, nsteps, every, avg_from):
    lat.set_solid_fraction(p.astype(np.float32))
    return lat.adjoint_forward(nsteps, every, avg_from, 0)[0]


def central_fd(lat, phi, cells, eps, nsteps, every, avg_from):
    out = []
    for c in cells:
        p = phi.astype(np.float64).copy()
        hi, lo = np.float32(phi[c] + eps), np.float32(phi[c] - eps)
        p[c] = hi
        jp = drag(lat, p, nsteps, every, avg_from)
        p[c] = lo
        jm = drag(lat, p, nsteps, every, avg_from)
        out.append((jp - jm) / (float(hi) - float(lo)))
    return np.array(out)


@pytest.mark.parametrize("storage", [sa.Storage.FP64, sa.Storage.FP32])
def test_adjoint_matches_central_differences(storage):
    setup, phi, _ = small(storage)
    say("case adjoint_fd", storage.name, setup.dims)
    n, every, a0 = 60, 16, 20          # 4 segments, the last one short
    lat = sa.Lattice(setup)
    lat.set_solid_fraction(phi)
    j, _ = lat.adjoint_forward(n, every, a0, 0)
    g = lat.adjoint_backward()
    s64, _, _ = small(sa.Storage.FP64)
    ref = sa.Lattice(s64)
    cells = fd_cells(phi, 8, np.random.default_rng(3))
    fd = central_fd(ref, phi, cells, 2e-4, n, every, a0)
    rel = np.abs(g[cells] - fd) / np.abs(fd).max()
    say("  J", j, "grad", g[cells], "fd", fd, "max rel err", rel.max())
    # FP64: the discrete adjoint is exact up to the difference quotient's
    # O(eps^2) error; FP32: the north star's 1e-3 relative
    assert rel.max() < (5e-7 if storage == sa.Storage.FP64 else 1e-3)
    # a cell outside the box (phi = 0): one-sided difference
    c0 = int(np.nonzero(phi == 0)[0][phi.size // 3])
    p = phi.astype(np.float64).copy()
    p[c0] = 1e-4
    j1 = drag(ref, p, n, every, a0)
    j0 = drag(ref, phi.astype(np.float64), n, every, a0)
    fd0 = (j1 - j0) / float(np.float32(1e-4))
    say("  outside cell", c0, "grad", g[c0], "fd", fd0)
    assert abs(g[c0] - fd0) <= 1e-2 * np.abs(fd).max()


def test_torch_wrapper_gradient():
    torch = pytest.importorskip("torch")
    from simlab_amd.torch_adjoint import porous_drag
    setup, phi, _ = small(sa.Storage.FP64)
    say("case adjoint_torch", setup.dims)
    lat = sa.Lattice(setup)
    t = torch.tensor(phi, dtype=torch.float64, requires_grad=True)
    j = porous_drag(t, lat, 40, 16, 10, 0)
    j.backward()
    lat.set_solid_fraction(phi)
    lat.adjoint_forward(40, 16, 10, 0)
    g = lat.adjoint_backward()
    assert np.allclose(t.grad.numpy(), g, rtol=0, atol=0)
    # directional derivative along a random direction inside the box (cells
    # with room for the step on both sides)
    rng = np.random.default_rng(11)
    d = np.where((phi > 0.01) & (phi < 0.99), rng.uniform(-1, 1, phi.size), 0.0)
    eps = 1e-4
    pp = (phi + eps * d).astype(np.float32)
    pm = (phi - eps * d).astype(np.float32)
    jp = float(porous_drag(torch.tensor(pp, dtype=torch.float64), lat, 40, 16, 10, 0))
    jm = float(porous_drag(torch.tensor(pm, dtype=torch.float64), lat, 40, 16, 10, 0))
    dd = pp.astype(np.float64) - pm.astype(np.float64)
    say("  directional: fd", jp - jm, "adjoint", float(g @ dd))
    assert abs((jp - jm) - float(g @ dd)) <= 1e-5 * abs(jp - jm)


def test_c3_drag_gradient_matches_central_differences():
    # the drag-sensitivity configuration: 128x64x64 channel, seeded solid
    # fraction in a 24^3 box around x = 40, inlet u = 0.04, Re 200 on the box,
    # 1000 steps with checkpoints every 100; FP32 forward and adjoint against
    # central differences (eps = 1e-4) of the FP64 forward at 16 seeded cells
    setup, phi, meta = W.porous_box(storage=sa.Storage.FP32)
    say("case c3_adjoint", setup.dims, meta)
    n, every, a0 = 1000, 100, 0
    lat = sa.Lattice(setup)
    lat.set_solid_fraction(phi)
    j, _ = lat.adjoint_forward(n, every, a0, 0)
    g = lat.adjoint_backward()
    del lat
    s64, _, _ = W.porous_box(storage=sa.Storage.FP64)
    ref = sa.Lattice(s64)
    cells = fd_cells(phi, 16, np.random.default_rng(20261203))
    fd = central_fd(ref, phi, cells, 1e-4, n, every, a0)
    scale = np.abs(fd).max()
    err = np.abs(g[cells] - fd)
    big = np.abs(fd) >= 0.01 * scale
    rel = err[big] / np.abs(fd[big])
    say("  J", j, "max |adjoint - fd| / max |fd|", err.max() / scale, "max rel err (|fd| >= 1 % of max)", rel.max())
    for c, a, b in zip(cells, g[cells], fd):
        say(f"   cell {c}: adjoint {a:+.6e} fd {b:+.6e}")
    # 1e-3 relative: per cell where the sensitivity is not negligible, and
    # against the gradient's scale everywhere (the FP32 trajectory the
    # adjoint linearises about differs from the FP64 one by rounding)
    assert rel.max() < 1e-3 and err.max() < 1e-3 * scale
