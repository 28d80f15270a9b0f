"""Lattice-unit force and moment to SI, and the coefficients a study reports
(tunnel.rs:1081-1082, 1121-1122, 1213-1245), with the sign conventions of
lib.rs:29-39: force and moment are the air's on the part; thrust is the force
along +spin axis; torque the moment about +spin axis (opposite to the
rotation for a driven propeller); power = -torque * omega, the shaft power
the motor supplies (positive for a driven propeller).

`frame` is the study's lattice frame as in tunnel.rs `frame`: three rows,
the lattice x, y, z axes in model coordinates (lattice = frame . model), so
model = frame^T . lattice; drag is along frame[0], lift along frame[2], side
along frame[1].
"""
from __future__ import annotations

import math

import numpy as np


def force_scale(density: float, dx: float, dt: float) -> float:
    """Newtons per lattice force unit: rho dx^4 / dt^2 (tunnel.rs:1081)."""
    return density * dx ** 4 / (dt * dt)


def moment_scale(density: float, dx: float, dt: float) -> float:
    """Newton metres per lattice moment unit (tunnel.rs:1082)."""
    return force_scale(density, dx, dt) * dx


def to_model(frame, v):
    """frame^T . v (tunnel.rs:480-482 mat_t_vec)."""
    r = np.asarray(frame, np.float64)
    return r.T @ np.asarray(v, np.float64)


def lattice_cd(fx: float, u: float, radius: float) -> float:
    """Drag coefficient of a sphere in lattice units: F_x / (u^2 / 2 pi r^2)
    (crates/cfd/examples/compare.rs:97)."""
    return fx / (0.5 * u * u * math.pi * radius * radius)


def summarize(ft_mean, frame, density: float, dx: float, dt: float, speed: float = 0.0, frontal_area: float = 0.0,
              spin_axis=None, rpm: float = 0.0, omega_lat=None, rotor_radius: float = 0.0) -> dict:
    """SI force, moment and coefficients of a window-averaged lattice force
    history row (Fx Fy Fz Tx Ty Tz), as tunnel.rs:1213-1245. For a spinning
    part give spin_axis (model coordinates, unit), rpm (signed), omega_lat
    (rad/step vector in lattice coordinates) and rotor_radius (metres)."""
    ft = np.asarray(ft_mean, np.float64)
    fr = np.asarray(frame, np.float64)
    force = to_model(fr, ft[:3]) * force_scale(density, dx, dt)
    moment = to_model(fr, ft[3:6]) * moment_scale(density, dx, dt)
    drag, lift, side = float(force @ fr[0]), float(force @ fr[2]), float(force @ fr[1])
    spin = spin_axis is not None
    if speed > 0.0:
        q_ref = 0.5 * density * speed * speed
    else:
        om = float(np.linalg.norm(omega_lat)) / dt if omega_lat is not None else 0.0
        q_ref = 0.5 * density * (om * rotor_radius) ** 2
    cd = drag / (q_ref * frontal_area) if speed > 0.0 and not spin and frontal_area > 0.0 else None
    thrust = torque = power = None
    if spin:
        a = np.asarray(spin_axis, np.float64)
        om = rpm * 2.0 * math.pi / 60.0
        torque = float(moment @ a)
        thrust = float(force @ a)
        power = -torque * om
    return {"force_n": force, "moment_nm": moment, "drag_n": drag, "lift_n": lift, "side_n": side, "q_ref": q_ref,
            "cd": cd, "thrust_n": thrust, "torque_nm": torque, "power_w": power}
