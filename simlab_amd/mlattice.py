"""The refined D3Q27 cumulant solver: Python mirror of `crates/cfd`'s
`MultiLattice` (crates/cfd/src/mlattice.rs:15-948) over the C ABI of
include/simlab_cfd.h (`simlab_cfd_multi_*`).

Names and argument meaning follow the reference: `LevelBox`, `MultiSetup`
(`levels`, `dims`, `cells`, `bytes`), `MultiLattice` with `run`,
`sub_steps`, `level_dims`, `level_cells`, `macros(level)`, `records`,
`reset_samples`, `vertex_means`, `slice_means`, and `voxelize_static27`
(voxel.rs:354-359). Errors raise `SimlabError` with the reference's
messages. The compute path is the HIP library; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import (FaceBc, Geometry, Probes, SimlabError, StaticVoxels, StepIn, Storage, WALL_DTYPE, _Geometry, _Probes,
               _StepIn, _check, _ptr, _surface_arrays, geometry_to_c, load_library, probes_to_c, steps_to_c)

__all__ = ["Q", "REST", "c27", "opp27", "weight27", "LevelBox", "MultiSetup", "MultiLattice", "voxelize_static27",
           "MAX_LEVELS", "RS27"]

MAX_LEVELS = 8
RS27 = 24          # words per wall record (mshader.rs:33)
Q = 27
REST = 13


def c27(i: int):
    """d3q27.rs:9-11: the velocity of direction i."""
    return (i % 3 - 1, (i // 3) % 3 - 1, i // 9 - 1)


def opp27(i: int) -> int:
    """d3q27.rs:13-15."""
    return 26 - i


def weight27(i: int) -> float:
    """d3q27.rs:17-24."""
    s = sum(abs(v) for v in c27(i))
    return (8 / 27, 2 / 27, 1 / 54, 1 / 216)[s]


class _LevelBox(C.Structure):
    _fields_ = [("lo", C.c_int32 * 3), ("n", C.c_uint32 * 3)]


class _MultiSetup(C.Structure):
    _fields_ = [
        ("dims0", C.c_uint32 * 3), ("nboxes", C.c_uint32), ("boxes", _LevelBox * (MAX_LEVELS - 1)),
        ("faces", C.c_int32 * 6), ("nu0", C.c_double), ("smagorinsky", C.c_double),
        ("u_inf", C.c_double * 3), ("u_init", C.c_double * 3), ("body_force0", C.c_double * 3),
        ("sponge", C.c_uint32 * 6), ("sponge_tau", C.c_double), ("sponge_sigma", C.c_double),
        ("absorb_velocity", C.c_int32), ("total_pressure_inlet", C.c_int32), ("storage", C.c_int32),
        ("debug_dump", C.c_int32), ("center", C.c_double * 3),
    ]


class _MultiInfo(C.Structure):
    _fields_ = [
        ("levels", C.c_uint32), ("sub_steps", C.c_uint32), ("fine_steps", C.c_uint64), ("memory_bytes", C.c_uint64),
        ("nrec", C.c_uint32), ("nvert", C.c_uint32), ("npix", C.c_uint32), ("storage", C.c_int32),
        ("dims", (C.c_uint32 * 3) * MAX_LEVELS), ("cells", C.c_uint64 * MAX_LEVELS), ("tau", C.c_double * MAX_LEVELS),
    ]


@dataclass
class LevelBox:
    """mlattice.rs:15-21: `lo` is the parent node of this level's node 0,
    `n` the parent nodes it spans per axis (dims = 2 (n - 1) + 1)."""
    lo: Sequence[int]
    n: Sequence[int]


@dataclass
class MultiSetup:
    """mlattice.rs:23-48 (field meaning as there)."""
    dims0: Sequence[int]
    boxes: List[LevelBox] = field(default_factory=list)
    faces: Sequence[FaceBc] = (FaceBc.PERIODIC,) * 6
    nu0: float = 0.1
    smagorinsky: float = 0.0
    u_inf: Sequence[float] = (0.0, 0.0, 0.0)
    u_init: Sequence[float] = (0.0, 0.0, 0.0)
    body_force0: Sequence[float] = (0.0, 0.0, 0.0)
    sponge: Sequence[int] = (0,) * 6
    sponge_tau: float = 1.0
    sponge_sigma: float = 0.0
    absorb_velocity: bool = False
    total_pressure_inlet: bool = False
    storage: Storage = Storage.FP32
    center: Sequence[float] = (0.0, 0.0, 0.0)
    debug_dump: bool = False

    def levels(self) -> int:
        return len(self.boxes) + 1

    def dims(self, l: int):
        """mlattice.rs:54-61."""
        if l == 0:
            return tuple(int(v) for v in self.dims0)
        b = self.boxes[l - 1]
        return tuple(2 * (int(b.n[k]) - 1) + 1 for k in range(3))

    def cells(self) -> List[int]:
        return [int(np.prod(self.dims(l), dtype=np.int64)) for l in range(self.levels())]

    def bytes(self) -> int:
        """mlattice.rs:65-67: populations A and B, flags and moment buffers."""
        pb = Storage(self.storage).pop_bytes()
        return sum(n * (2 * Q * pb + 4 + 3 * 12 * 4 // 8) for n in self.cells())

    def to_c(self) -> _MultiSetup:
        if len(self.boxes) > MAX_LEVELS - 1:
            raise SimlabError(f"at most {MAX_LEVELS} levels")
        s = _MultiSetup()
        s.dims0[:] = [int(v) for v in self.dims0]
        s.nboxes = len(self.boxes)
        for k, b in enumerate(self.boxes):
            s.boxes[k].lo[:] = [int(v) for v in b.lo]
            s.boxes[k].n[:] = [int(v) for v in b.n]
        s.faces[:] = [int(f) for f in self.faces]
        s.nu0, s.smagorinsky = float(self.nu0), float(self.smagorinsky)
        s.u_inf[:] = [float(v) for v in self.u_inf]
        s.u_init[:] = [float(v) for v in self.u_init]
        s.body_force0[:] = [float(v) for v in self.body_force0]
        s.sponge[:] = [int(v) for v in self.sponge]
        s.sponge_tau, s.sponge_sigma = float(self.sponge_tau), float(self.sponge_sigma)
        s.absorb_velocity = int(bool(self.absorb_velocity))
        s.total_pressure_inlet = int(bool(self.total_pressure_inlet))
        s.storage = int(self.storage)
        s.debug_dump = int(bool(self.debug_dump))
        s.center[:] = [float(v) for v in self.center]
        return s


_sigs_done = False


def _lib():
    global _sigs_done
    lib = load_library()
    if not _sigs_done:
        V, I, U64, P = C.c_void_p, C.c_int, C.c_uint64, C.c_void_p
        sigs = {
            "simlab_cfd_multi_new": (I, [C.POINTER(_MultiSetup), C.POINTER(_Geometry), C.POINTER(_Probes), C.POINTER(V)]),
            "simlab_cfd_multi_free": (None, [V]),
            "simlab_cfd_multi_info": (I, [V, C.POINTER(_MultiInfo)]),
            "simlab_cfd_multi_run": (I, [V, C.POINTER(_StepIn), C.c_size_t, P]),
            "simlab_cfd_multi_run_still": (I, [V, C.c_size_t, P]),
            "simlab_cfd_multi_macros": (I, [V, C.c_uint32, P]),
            "simlab_cfd_multi_records": (I, [V, P]),
            "simlab_cfd_multi_reset_samples": (I, [V]),
            "simlab_cfd_multi_vertex_means": (I, [V, P]),
            "simlab_cfd_multi_slice_means": (I, [V, P]),
            "simlab_cfd_multi_get_pops": (I, [V, C.c_uint32, P]),
            "simlab_cfd_voxelize_static27": (I, [P, U64, P, U64, C.POINTER(U64), C.POINTER(V)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _sigs_done = True
    return lib


def voxelize_static27(verts, tris, dims) -> StaticVoxels:
    """voxel.rs:354-359: solid cells and D3Q27 wall links of a surface in
    lattice coordinates (host CPU)."""
    from . import _voxels_from_handle
    lib = _lib()
    v, t = _surface_arrays(verts, tris)
    d = (C.c_uint64 * 3)(*[int(x) for x in dims])
    h = C.c_void_p()
    _check(lib, lib.simlab_cfd_voxelize_static27(_ptr(v), len(v), _ptr(t), len(t), d, C.byref(h)))
    return _voxels_from_handle(lib, h)


class MultiLattice:
    """mlattice.rs:139-948 on the GPU."""

    def __init__(self, setup: MultiSetup, geom: Optional[Geometry] = None, probes: Optional[Probes] = None):
        lib = _lib()
        self._lib = lib
        self._h = C.c_void_p()
        s = setup.to_c()
        g, keep_g = geometry_to_c(geom)
        p, keep_p = probes_to_c(probes)
        _check(lib, lib.simlab_cfd_multi_new(C.byref(s), C.byref(g), C.byref(p) if p is not None else None,
                                             C.byref(self._h)))
        del keep_g, keep_p
        self.setup = setup
        info = self.info()
        self.setup.storage = Storage(info.storage)
        self.nrec, self.nvert, self.npix = info.nrec, info.nvert, info.npix
        self.memory_bytes = info.memory_bytes

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def close(self):
        """Frees the device buffers now (the Rust `Drop`)."""
        if self._h:
            self._lib.simlab_cfd_multi_free(self._h)
            self._h = C.c_void_p()

    @property
    def storage(self) -> Storage:
        return Storage(self.info().storage)

    def info(self) -> _MultiInfo:
        i = _MultiInfo()
        _check(self._lib, self._lib.simlab_cfd_multi_info(self._h, C.byref(i)))
        return i

    @property
    def fine_steps(self) -> int:
        return int(self.info().fine_steps)

    def sub_steps(self) -> int:
        """mlattice.rs:708-711: finest steps per coarsest step."""
        return int(self.info().sub_steps)

    def level_dims(self):
        i = self.info()
        return [tuple(int(v) for v in i.dims[l]) for l in range(i.levels)]

    def level_cells(self):
        i = self.info()
        return [int(i.cells[l]) for l in range(i.levels)]

    def level_tau(self):
        i = self.info()
        return [float(i.tau[l]) for l in range(i.levels)]

    def run(self, steps: Sequence[StepIn]) -> np.ndarray:
        """mlattice.rs:725-784: force and moment per finest step (nsteps x 6)."""
        n = len(steps)
        arr = steps_to_c(steps)
        out = np.zeros((n, 6), np.float64)
        _check(self._lib, self._lib.simlab_cfd_multi_run(self._h, arr, n, _ptr(out)))
        return out

    def run_still(self, nsteps: int) -> np.ndarray:
        out = np.zeros((int(nsteps), 6), np.float64)
        _check(self._lib, self._lib.simlab_cfd_multi_run_still(self._h, int(nsteps), _ptr(out)))
        return out

    def macros(self, level: int) -> np.ndarray:
        """mlattice.rs:913-930: (rho, u) per cell of a level (needs debug_dump)."""
        n = self.level_cells()[level]
        out = np.zeros((n, 4), np.float32)
        _check(self._lib, self._lib.simlab_cfd_multi_macros(self._h, int(level), _ptr(out)))
        return out

    def records(self) -> np.ndarray:
        out = np.zeros((max(self.nrec, 1), RS27), np.uint32)
        _check(self._lib, self._lib.simlab_cfd_multi_records(self._h, _ptr(out)))
        return out

    def reset_samples(self):
        _check(self._lib, self._lib.simlab_cfd_multi_reset_samples(self._h))

    def vertex_means(self) -> np.ndarray:
        out = np.zeros(self.nvert, np.float32)
        if self.nvert:
            _check(self._lib, self._lib.simlab_cfd_multi_vertex_means(self._h, _ptr(out)))
        return out

    def slice_means(self) -> np.ndarray:
        out = np.zeros((self.npix, 4), np.float32)
        if self.npix:
            _check(self._lib, self._lib.simlab_cfd_multi_slice_means(self._h, _ptr(out)))
        return out

    def get_pops(self, level: int) -> np.ndarray:
        """The populations the level's next step reads, as deviations f - w
        (27 x cells doubles)."""
        n = self.level_cells()[level]
        out = np.zeros((Q, n), np.float64)
        _check(self._lib, self._lib.simlab_cfd_multi_get_pops(self._h, int(level), _ptr(out)))
        return out
