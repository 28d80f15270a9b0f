"""MI355X (gfx950) D3Q19 wind-tunnel lattice: Python mirror of the
`crates/cfd` lattice API over the C ABI of include/simlab_cfd.h.

Names, argument meaning and errors follow the reference's Rust API
(crates/cfd/src/lattice.rs:11-841, voxel.rs:328-352): `LatticeSetup`,
`FaceBc`, `Storage`, `StepIn`, `Geometry`, `Probes`, `Lattice` with
`run / macros / records / reset_samples / vertex_means / slice_means`, and
`voxelize_static`. Errors raise `SimlabError` carrying the library's message
(the reference's `bail!` strings). The compute path is the HIP library
`libsimlab_cfd.so`; there is no fallback: if it is missing or no GPU is
present, the calls fail.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

__all__ = [
    "SimlabError", "Storage", "FaceBc", "LatticeSetup", "StepIn", "IDENTITY", "Geometry", "Probes",
    "SliceDef", "Lattice", "StaticVoxels", "voxelize_static", "inside_grid", "gpu_info", "max_cells",
    "load_library", "LIB_PATH", "Batch", "SpinVoxels", "voxelize_spin",
]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsimlab_cfd.so")


class SimlabError(RuntimeError):
    """A failed library call (the reference's anyhow error)."""


# ---------------------------------------------------------------- C types
class _Setup(C.Structure):
    _fields_ = [
        ("dims", C.c_uint32 * 3), ("faces", C.c_int32 * 6), ("nu", C.c_double), ("smagorinsky", C.c_double),
        ("u_inf", C.c_double * 3), ("u_init", C.c_double * 3), ("body_force", C.c_double * 3),
        ("sponge", C.c_uint32 * 6), ("sponge_tau", C.c_double), ("sponge_sigma", C.c_double),
        ("absorb_velocity", C.c_int32), ("total_pressure_inlet", C.c_int32), ("storage", C.c_int32),
        ("debug_dump", C.c_int32), ("center", C.c_double * 3), ("wall_model", C.c_int32), ("reserved0", C.c_int32),
    ]


class _WallCell(C.Structure):
    _fields_ = [("cell", C.c_uint32), ("mask", C.c_uint32), ("q", C.c_float * 27)]


WALL_DTYPE = np.dtype([("cell", "<u4"), ("mask", "<u4"), ("q", "<f4", (27,))])
assert WALL_DTYPE.itemsize == C.sizeof(_WallCell)


class _Geometry(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("reserved0", C.c_int32),
        ("nsolid", C.c_uint64), ("solid", C.POINTER(C.c_uint32)),
        ("nwall", C.c_uint64), ("walls", C.c_void_p),
        ("occ_lo", C.c_double * 3), ("occ_h", C.c_double), ("occ_dims", C.c_uint32 * 3), ("reserved1", C.c_uint32),
        ("nocc_words", C.c_uint64), ("occ_bits", C.POINTER(C.c_uint32)),
        ("ntri", C.c_uint64), ("tris", C.POINTER(C.c_float)),
        ("box_lo", C.c_int32 * 3), ("box_n", C.c_uint32 * 3),
    ]


class _SliceDef(C.Structure):
    _fields_ = [("origin", C.c_int32 * 3), ("u", C.c_int32 * 3), ("v", C.c_int32 * 3), ("nu", C.c_uint32), ("nv", C.c_uint32)]


class _Probes(C.Structure):
    _fields_ = [("nvert", C.c_uint64), ("verts", C.POINTER(C.c_double)), ("nslice", C.c_uint64),
                ("slices", C.POINTER(_SliceDef)), ("vertex_offset", C.c_double)]


class _StepIn(C.Structure):
    _fields_ = [("rot", C.c_double * 9), ("rot_prev", C.c_double * 9), ("omega", C.c_double * 3),
                ("sample", C.c_int32), ("reserved0", C.c_int32)]


class _Info(C.Structure):
    _fields_ = [
        ("steps_done", C.c_uint64), ("cells", C.c_uint64), ("memory_bytes", C.c_uint64), ("dims", C.c_uint32 * 3),
        ("nrec", C.c_uint32), ("nface", C.c_uint32), ("npix", C.c_uint32), ("nvert", C.c_uint32),
        ("storage", C.c_int32), ("moving", C.c_int32), ("tau0", C.c_double), ("bulk_cells", C.c_uint64),
        ("halfway_cells", C.c_uint64),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libsimlab_cfd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SimlabError(f"{path} is missing: build it with `python simlab_amd/build.py` "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(path)
    V, I, U64, D, P = C.c_void_p, C.c_int, C.c_uint64, C.c_double, C.c_void_p
    sigs = {
        "simlab_cfd_last_error": (C.c_char_p, []),
        "simlab_cfd_gpu_info": (I, [C.c_char_p, C.c_size_t]),
        "simlab_cfd_max_cells": (U64, [C.c_int32]),
        "simlab_cfd_lattice_new": (I, [C.POINTER(_Setup), C.POINTER(_Geometry), C.POINTER(_Probes), C.POINTER(V)]),
        "simlab_cfd_lattice_free": (None, [V]),
        "simlab_cfd_lattice_info": (I, [V, C.POINTER(_Info)]),
        "simlab_cfd_lattice_run": (I, [V, C.POINTER(_StepIn), C.c_size_t, P]),
        "simlab_cfd_lattice_run_still": (I, [V, C.c_size_t, P]),
        "simlab_cfd_lattice_macros": (I, [V, P]),
        "simlab_cfd_lattice_records": (I, [V, P]),
        "simlab_cfd_lattice_reset_samples": (I, [V]),
        "simlab_cfd_lattice_vertex_means": (I, [V, P]),
        "simlab_cfd_lattice_slice_means": (I, [V, P]),
        "simlab_cfd_lattice_mass": (I, [V, C.POINTER(D)]),
        "simlab_cfd_lattice_get_pops": (I, [V, P, C.c_size_t]),
        "simlab_cfd_lattice_set_pops": (I, [V, P, C.c_size_t]),
        "simlab_cfd_lattice_run_noalloc": (I, [V, C.POINTER(_StepIn), C.c_size_t, P]),
        "simlab_cfd_lattice_set_timing": (I, [V, I]),
        "simlab_cfd_lattice_timing": (I, [V, C.POINTER(D), C.POINTER(D), C.POINTER(U64)]),
        "simlab_cfd_lattice_autotune": (I, [V, C.c_uint32, I, P, C.POINTER(D)]),
        "simlab_cfd_lattice_set_tile": (I, [V, C.c_uint32, C.c_uint32]),
        "simlab_cfd_lattice_get_tile": (I, [V, V]),
        "simlab_cfd_lattice_checkpoint_size": (I, [V, V]),
        "simlab_cfd_lattice_checkpoint": (I, [V, V, C.c_size_t]),
        "simlab_cfd_lattice_restore": (I, [V, V, C.c_size_t]),
        "simlab_cfd_batch_new": (I, [P, C.c_size_t, C.POINTER(V)]),
        "simlab_cfd_batch_run_still": (I, [V, C.c_size_t, P]),
        "simlab_cfd_batch_free": (None, [V]),
        "simlab_cfd_lattice_set_solid_fraction": (I, [V, P]),
        "simlab_cfd_adjoint_forward": (I, [V, C.c_size_t, C.c_size_t, C.c_size_t, I, C.POINTER(D), P]),
        "simlab_cfd_adjoint_backward": (I, [V, P]),
        "simlab_cfd_voxelize_static": (I, [P, U64, P, U64, C.POINTER(U64), C.POINTER(V)]),
        "simlab_cfd_voxelize_static_gpu": (I, [P, U64, P, U64, C.POINTER(U64), C.POINTER(V)]),
        "simlab_cfd_voxelize_spin": (I, [P, U64, P, U64, C.POINTER(D), D, C.POINTER(V)]),
        "simlab_cfd_spin_voxels_view": (I, [V, C.POINTER(D), C.POINTER(D), C.POINTER(U64), C.POINTER(U64), C.POINTER(P),
                                            C.POINTER(U64), C.POINTER(P), C.POINTER(D), C.POINTER(D)]),
        "simlab_cfd_spin_voxels_free": (None, [V]),
        "simlab_cfd_voxels_view": (I, [V, C.POINTER(U64), C.POINTER(P), C.POINTER(U64), C.POINTER(P),
                                       C.POINTER(U64), C.POINTER(U64)]),
        "simlab_cfd_voxels_free": (None, [V]),
        "simlab_cfd_inside_grid": (I, [P, U64, P, U64, C.POINTER(D), D, C.POINTER(U64), P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _err(lib) -> str:
    m = lib.simlab_cfd_last_error()
    return m.decode() if m else "unknown error"


def _check(lib, rc):
    if rc != 0:
        raise SimlabError(_err(lib))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- setup types
class Storage(enum.IntEnum):
    """lattice.rs:11-32 (Fp32, Fp16) plus Bf16 and the Fp64 oracle storage."""
    FP32 = 0
    FP16 = 1
    BF16 = 2
    FP64 = 3

    def pop_bytes(self) -> int:
        return {0: 4, 1: 2, 2: 2, 3: 8}[int(self)]

    def bytes_per_cell(self) -> int:
        return 19 * self.pop_bytes() + 4


class FaceBc(enum.IntEnum):
    """lattice.rs:34-43."""
    VELOCITY = 0
    OPEN = 1
    PERIODIC = 2


@dataclass
class LatticeSetup:
    """lattice.rs:45-85 (field meaning and defaults of the reference tests)."""
    dims: Sequence[int]
    faces: Sequence[FaceBc] = (FaceBc.PERIODIC,) * 6
    nu: float = 0.1
    smagorinsky: float = 0.0
    u_inf: Sequence[float] = (0.0, 0.0, 0.0)
    u_init: Sequence[float] = (0.0, 0.0, 0.0)
    body_force: Sequence[float] = (0.0, 0.0, 0.0)
    sponge: Sequence[int] = (0,) * 6
    sponge_tau: float = 0.0
    sponge_sigma: float = 0.0
    absorb_velocity: bool = False
    total_pressure_inlet: bool = False
    storage: Storage = Storage.FP32
    center: Sequence[float] = (0.0, 0.0, 0.0)
    debug_dump: bool = False
    wall_model: bool = False

    def periodic(self):
        return [self.faces[0] == FaceBc.PERIODIC, self.faces[2] == FaceBc.PERIODIC, self.faces[4] == FaceBc.PERIODIC]

    def cells(self) -> int:
        return int(self.dims[0]) * int(self.dims[1]) * int(self.dims[2])

    def to_c(self) -> _Setup:
        s = _Setup()
        s.dims[:] = [int(v) for v in self.dims]
        s.faces[:] = [int(f) for f in self.faces]
        s.nu, s.smagorinsky = float(self.nu), float(self.smagorinsky)
        s.u_inf[:] = [float(v) for v in self.u_inf]
        s.u_init[:] = [float(v) for v in self.u_init]
        s.body_force[:] = [float(v) for v in self.body_force]
        s.sponge[:] = [int(v) for v in self.sponge]
        s.sponge_tau, s.sponge_sigma = float(self.sponge_tau), float(self.sponge_sigma)
        s.absorb_velocity = int(bool(self.absorb_velocity))
        s.total_pressure_inlet = int(bool(self.total_pressure_inlet))
        s.storage = int(self.storage)
        s.debug_dump = int(bool(self.debug_dump))
        s.center[:] = [float(v) for v in self.center]
        s.wall_model = int(bool(self.wall_model))
        return s


IDENTITY = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))


@dataclass
class StepIn:
    """lattice.rs:104-121: rotation now / before (columns = body axes), omega (rad/step)."""
    rot: Sequence[Sequence[float]] = IDENTITY
    rot_prev: Sequence[Sequence[float]] = IDENTITY
    omega: Sequence[float] = (0.0, 0.0, 0.0)
    sample: bool = False

    @staticmethod
    def still(sample: bool = False) -> "StepIn":
        return StepIn(IDENTITY, IDENTITY, (0.0, 0.0, 0.0), sample)


@dataclass
class StaticVoxels:
    """voxel.rs:337-344: solid cell indices and wall cells (sorted by cell)."""
    solid: np.ndarray
    walls: np.ndarray          # WALL_DTYPE records
    lo: tuple = (0, 0, 0)
    hi: tuple = (0, 0, 0)


@dataclass
class Geometry:
    """lattice.rs:87-92: None, Static (solid cells + wall links) or Spin (a
    spinning part's body-frame occupancy and triangles, re-voxelized every
    step inside the swept box box_lo .. box_lo + box_n)."""
    kind: str = "none"
    vox: Optional[object] = None
    box_lo: tuple = (0, 0, 0)
    box_n: tuple = (0, 0, 0)

    @staticmethod
    def static(vox: StaticVoxels) -> "Geometry":
        return Geometry("static", vox)

    @staticmethod
    def spin(vox: "SpinVoxels", box_lo, box_n) -> "Geometry":
        return Geometry("spin", vox, tuple(int(v) for v in box_lo), tuple(int(v) for v in box_n))


@dataclass
class SpinVoxels:
    """voxel.rs:492-505: body-frame occupancy bits at spacing occ_h from
    occ_lo (x fastest), triangles (v0, e1, e2; ntri x 12 float32, body
    frame), largest distance from the spin axis and the axial extent."""
    occ_lo: tuple
    occ_h: float
    occ_dims: tuple
    occ_bits: np.ndarray
    tris: np.ndarray
    radius: float
    axial: tuple


@dataclass
class SliceDef:
    """shader.rs:77-84."""
    origin: Sequence[int]
    u: Sequence[int]
    v: Sequence[int]
    nu: int
    nv: int


@dataclass
class Probes:
    """lattice.rs:94-102: surface vertices (body position, unit normal) and slices."""
    verts: List[tuple] = field(default_factory=list)
    slices: List[SliceDef] = field(default_factory=list)
    vertex_offset: float = 0.0


# ---------------------------------------------------------------- library info
def gpu_info() -> str:
    """gpu.rs:56-63."""
    lib = load_library()
    buf = C.create_string_buffer(256)
    _check(lib, lib.simlab_cfd_gpu_info(buf, 256))
    return buf.value.decode()


def max_cells(storage: Storage) -> int:
    """lattice.rs:223-227."""
    return int(load_library().simlab_cfd_max_cells(int(storage)))


# ---------------------------------------------------------------- voxelizer
def _surface_arrays(verts, tris):
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.uint32).reshape(-1, 3))
    return v, t


def voxelize_static(verts, tris, dims, gpu: bool = False) -> StaticVoxels:
    """voxel.rs:348-352: solid cells and wall links of a surface in lattice
    coordinates; gpu: the inside test and the link search on the GPU
    (simlab_cfd_voxelize_static_gpu, the same output)."""
    lib = load_library()
    v, t = _surface_arrays(verts, tris)
    d = (C.c_uint64 * 3)(*[int(x) for x in dims])
    h = C.c_void_p()
    fn = lib.simlab_cfd_voxelize_static_gpu if gpu else lib.simlab_cfd_voxelize_static
    _check(lib, fn(_ptr(v), len(v), _ptr(t), len(t), d, C.byref(h)))
    return _voxels_from_handle(lib, h)


def _voxels_from_handle(lib, h) -> StaticVoxels:
    """Copies a simlab_static_voxels into a StaticVoxels and frees it."""
    try:
        ns, nw = C.c_uint64(), C.c_uint64()
        ps, pw = C.c_void_p(), C.c_void_p()
        lo, hi = (C.c_uint64 * 3)(), (C.c_uint64 * 3)()
        _check(lib, lib.simlab_cfd_voxels_view(h, C.byref(ns), C.byref(ps), C.byref(nw), C.byref(pw), lo, hi))
        solid = np.ctypeslib.as_array(C.cast(ps, C.POINTER(C.c_uint32)), (ns.value,)).copy() if ns.value else np.zeros(0, np.uint32)
        if nw.value:
            buf = (C.c_char * (nw.value * WALL_DTYPE.itemsize)).from_address(pw.value)
            walls = np.frombuffer(bytes(buf), dtype=WALL_DTYPE).copy()
        else:
            walls = np.zeros(0, WALL_DTYPE)
        return StaticVoxels(solid, walls, tuple(lo), tuple(hi))
    finally:
        lib.simlab_cfd_voxels_free(h)


def voxelize_spin(verts, tris, axis, h: float) -> SpinVoxels:
    """voxel.rs:509-553: a spinning part in body coordinates (lattice units,
    origin at the spin centre): occupancy at spacing h, triangles split to
    edges of at most 1.5 cells."""
    lib = load_library()
    v, t = _surface_arrays(verts, tris)
    ax = (C.c_double * 3)(*[float(a) for a in axis])
    hnd = C.c_void_p()
    _check(lib, lib.simlab_cfd_voxelize_spin(_ptr(v), len(v), _ptr(t), len(t), ax, float(h), C.byref(hnd)))
    try:
        lo = (C.c_double * 3)()
        oh = C.c_double()
        dims = (C.c_uint64 * 3)()
        nw, nt = C.c_uint64(), C.c_uint64()
        bp, tp = C.c_void_p(), C.c_void_p()
        rad = C.c_double()
        axial = (C.c_double * 2)()
        _check(lib, lib.simlab_cfd_spin_voxels_view(hnd, lo, C.byref(oh), dims, C.byref(nw), C.byref(bp), C.byref(nt),
                                                    C.byref(tp), C.byref(rad), axial))
        bits = np.ctypeslib.as_array(C.cast(bp, C.POINTER(C.c_uint32)), shape=(nw.value,)).copy() if nw.value else \
            np.zeros(0, np.uint32)
        tr = np.ctypeslib.as_array(C.cast(tp, C.POINTER(C.c_float)), shape=(nt.value * 12,)).copy().reshape(-1, 12) \
            if nt.value else np.zeros((0, 12), np.float32)
        return SpinVoxels(tuple(lo), oh.value, tuple(int(d) for d in dims), bits, tr, rad.value, (axial[0], axial[1]))
    finally:
        lib.simlab_cfd_spin_voxels_free(hnd)


def inside_grid(verts, tris, origin, h, dims) -> np.ndarray:
    """voxel.rs:73-95: inside flags of the grid points origin + h*(i,j,k), i fastest."""
    lib = load_library()
    v, t = _surface_arrays(verts, tris)
    d = (C.c_uint64 * 3)(*[int(x) for x in dims])
    o = (C.c_double * 3)(*[float(x) for x in origin])
    out = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]), np.uint8)
    _check(lib, lib.simlab_cfd_inside_grid(_ptr(v), len(v), _ptr(t), len(t), o, float(h), d, _ptr(out)))
    return out.astype(bool)


# ---------------------------------------------------------------- geometry marshalling (shared with tests)
def geometry_to_c(geom: Optional[Geometry]):
    """C view of a Geometry; returns (struct, keepalive)."""
    g = _Geometry()
    keep = []
    if geom is None or geom.kind == "none":
        g.kind = 0
        return g, keep
    if geom.kind == "static":
        vox = geom.vox
        solid = np.ascontiguousarray(vox.solid, dtype=np.uint32)
        walls = np.ascontiguousarray(vox.walls, dtype=WALL_DTYPE)
        keep += [solid, walls]
        g.kind = 1
        g.nsolid = len(solid)
        g.solid = solid.ctypes.data_as(C.POINTER(C.c_uint32))
        g.nwall = len(walls)
        g.walls = walls.ctypes.data_as(C.c_void_p)
        return g, keep
    if geom.kind == "spin":
        v = geom.vox
        bits = np.ascontiguousarray(v.occ_bits, dtype=np.uint32)
        tris = np.ascontiguousarray(v.tris, dtype=np.float32).reshape(-1, 12)
        keep += [bits, tris]
        g.kind = 2
        g.occ_lo[:] = [float(x) for x in v.occ_lo]
        g.occ_h = float(v.occ_h)
        g.occ_dims[:] = [int(x) for x in v.occ_dims]
        g.nocc_words = len(bits)
        g.occ_bits = bits.ctypes.data_as(C.POINTER(C.c_uint32))
        g.ntri = len(tris)
        g.tris = tris.ctypes.data_as(C.POINTER(C.c_float))
        g.box_lo[:] = list(geom.box_lo)
        g.box_n[:] = list(geom.box_n)
        return g, keep
    raise SimlabError(f"unknown geometry kind {geom.kind!r}")


def probes_to_c(probes: Optional[Probes]):
    if probes is None:
        return None, []
    p = _Probes()
    verts = np.ascontiguousarray(np.asarray([list(a) + list(b) for a, b in probes.verts], dtype=np.float64).reshape(-1, 6))
    sl = (_SliceDef * max(1, len(probes.slices)))()
    for k, s in enumerate(probes.slices):
        sl[k].origin[:] = list(s.origin)
        sl[k].u[:] = list(s.u)
        sl[k].v[:] = list(s.v)
        sl[k].nu, sl[k].nv = int(s.nu), int(s.nv)
    p.nvert = len(verts)
    p.verts = verts.ctypes.data_as(C.POINTER(C.c_double))
    p.nslice = len(probes.slices)
    p.slices = C.cast(sl, C.POINTER(_SliceDef))
    p.vertex_offset = float(probes.vertex_offset)
    return p, [verts, sl]


def steps_to_c(steps: Sequence[StepIn]):
    arr = (_StepIn * max(1, len(steps)))()
    for k, s in enumerate(steps):
        arr[k].rot[:] = [float(s.rot[r][c]) for r in range(3) for c in range(3)]
        arr[k].rot_prev[:] = [float(s.rot_prev[r][c]) for r in range(3) for c in range(3)]
        arr[k].omega[:] = [float(v) for v in s.omega]
        arr[k].sample = int(bool(s.sample))
    return arr


_POP_DTYPE = {Storage.FP32: np.float32, Storage.FP16: np.uint16, Storage.BF16: np.uint16, Storage.FP64: np.float64}


class Lattice:
    """The MI355X lattice (lattice.rs:148-841)."""

    def __init__(self, setup: LatticeSetup, geometry: Optional[Geometry] = None, probes: Optional[Probes] = None):
        self._lib = load_library()
        self.setup = setup
        cs = setup.to_c()
        g, keep_g = geometry_to_c(geometry)
        p, keep_p = probes_to_c(probes)
        h = C.c_void_p()
        rc = self._lib.simlab_cfd_lattice_new(C.byref(cs), C.byref(g), C.byref(p) if p is not None else None, C.byref(h))
        del keep_g, keep_p
        _check(self._lib, rc)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.simlab_cfd_lattice_free(h)
            self._h = None

    def close(self):
        self.__del__()

    def _info(self) -> _Info:
        i = _Info()
        _check(self._lib, self._lib.simlab_cfd_lattice_info(self._h, C.byref(i)))
        return i

    @property
    def steps_done(self) -> int:
        return int(self._info().steps_done)

    @property
    def nrec(self) -> int:
        return int(self._info().nrec)

    @property
    def nface(self) -> int:
        return int(self._info().nface)

    @property
    def npix(self) -> int:
        return int(self._info().npix)

    @property
    def memory_bytes(self) -> int:
        return int(self._info().memory_bytes)

    @property
    def halfway_cells(self) -> int:
        return int(self._info().halfway_cells)

    @property
    def bulk_cells(self) -> int:
        return int(self._info().bulk_cells)

    @property
    def storage(self) -> Storage:
        return Storage(self._info().storage)

    def run(self, steps: Sequence[StepIn]) -> np.ndarray:
        """lattice.rs:699-753: (len(steps), 6) force and moment after each step."""
        out = np.zeros((len(steps), 6), np.float64)
        arr = steps_to_c(steps)
        _check(self._lib, self._lib.simlab_cfd_lattice_run(self._h, arr, len(steps), _ptr(out)))
        return out

    def run_noalloc(self, steps: Sequence[StepIn]) -> np.ndarray:
        """`run` whose force partials live in the library's device scratch
        array: no device allocation (simlab_cfd_lattice_run_noalloc)."""
        out = np.zeros((len(steps), 6), np.float64)
        arr = steps_to_c(steps)
        _check(self._lib, self._lib.simlab_cfd_lattice_run_noalloc(self._h, arr, len(steps), _ptr(out)))
        return out

    def run_still(self, nsteps: int) -> np.ndarray:
        """`run` with `StepIn::still(false)` steps."""
        out = np.zeros((int(nsteps), 6), np.float64)
        _check(self._lib, self._lib.simlab_cfd_lattice_run_still(self._h, int(nsteps), _ptr(out)))
        return out

    def macros(self) -> np.ndarray:
        """lattice.rs:818-835: (cells, 4) rho, ux, uy, uz; NaN where solid/face."""
        out = np.zeros((self.setup.cells(), 4), np.float32)
        _check(self._lib, self._lib.simlab_cfd_lattice_macros(self._h, _ptr(out)))
        return out

    def records(self) -> np.ndarray:
        """lattice.rs:837-841: (max(nrec,1), 40) u32 words."""
        out = np.zeros((max(self.nrec, 1), 40), np.uint32)
        _check(self._lib, self._lib.simlab_cfd_lattice_records(self._h, _ptr(out)))
        return out

    def reset_samples(self):
        _check(self._lib, self._lib.simlab_cfd_lattice_reset_samples(self._h))

    def vertex_means(self) -> np.ndarray:
        n = int(self._info().nvert)
        out = np.zeros(n, np.float32)
        _check(self._lib, self._lib.simlab_cfd_lattice_vertex_means(self._h, _ptr(out)))
        return out

    def slice_means(self) -> np.ndarray:
        out = np.zeros((self.npix, 4), np.float32)
        _check(self._lib, self._lib.simlab_cfd_lattice_slice_means(self._h, _ptr(out)))
        return out

    def mass(self) -> float:
        m = C.c_double()
        _check(self._lib, self._lib.simlab_cfd_lattice_mass(self._h, C.byref(m)))
        return m.value

    def get_pops(self) -> np.ndarray:
        st = self.storage
        out = np.zeros(19 * self.setup.cells(), _POP_DTYPE[st])
        _check(self._lib, self._lib.simlab_cfd_lattice_get_pops(self._h, _ptr(out), out.nbytes))
        return out

    def set_pops(self, pops: np.ndarray):
        a = np.ascontiguousarray(pops, dtype=_POP_DTYPE[self.storage])
        _check(self._lib, self._lib.simlab_cfd_lattice_set_pops(self._h, _ptr(a), a.nbytes))

    # ---- launch-shape autotuner (no counterpart in the reference)
    def autotune(self, steps: int = 20, use_cache: bool = True):
        """Times each candidate tile of the stream-collide kernel, keeps the
        fastest and restores the state; returns ((bx, by), best step us)."""
        tile = (C.c_uint32 * 2)()
        us = C.c_double()
        _check(self._lib, self._lib.simlab_cfd_lattice_autotune(self._h, int(steps), int(bool(use_cache)), tile, C.byref(us)))
        return (int(tile[0]), int(tile[1])), us.value

    def set_tile(self, bx: int, by: int):
        _check(self._lib, self._lib.simlab_cfd_lattice_set_tile(self._h, int(bx), int(by)))

    def checkpoint(self) -> np.ndarray:
        """The whole evolving state (populations, flags, records, sampler
        sums, steps_done) as bytes."""
        n = C.c_size_t()
        _check(self._lib, self._lib.simlab_cfd_lattice_checkpoint_size(self._h, C.byref(n)))
        out = np.zeros(n.value, np.uint8)
        _check(self._lib, self._lib.simlab_cfd_lattice_checkpoint(self._h, _ptr(out), n.value))
        return out

    def restore(self, ckpt: np.ndarray):
        """Puts back a checkpoint of a lattice with the same setup and geometry."""
        c = np.ascontiguousarray(ckpt, np.uint8)
        _check(self._lib, self._lib.simlab_cfd_lattice_restore(self._h, _ptr(c), c.size))

    def tile(self):
        """(bx, by) of the stream-collide kernel's blocks."""
        t = np.zeros(2, np.uint32)
        _check(self._lib, self._lib.simlab_cfd_lattice_get_tile(self._h, _ptr(t)))
        return int(t[0]), int(t[1])

    # ---- partial bounce-back and the drag adjoint (no counterpart in the reference)
    def set_solid_fraction(self, phi):
        """Solid fraction per cell (x fastest, values in [0, 1]; None: all 0)."""
        if phi is None:
            _check(self._lib, self._lib.simlab_cfd_lattice_set_solid_fraction(self._h, None))
            return
        a = np.ascontiguousarray(np.asarray(phi, np.float32).reshape(-1))
        if a.size != self.setup.cells():
            raise SimlabError("solid fraction needs one value per cell")
        _check(self._lib, self._lib.simlab_cfd_lattice_set_solid_fraction(self._h, _ptr(a)))

    def adjoint_forward(self, nsteps: int, ckpt_every: int = 100, avg_from: int = 0, axis: int = 0):
        """J = mean force component `axis` over steps [avg_from, nsteps) from the
        initial state, with checkpoints for adjoint_backward; returns (J, ft)."""
        j = C.c_double()
        ft = np.zeros((int(nsteps), 6), np.float64)
        _check(self._lib, self._lib.simlab_cfd_adjoint_forward(self._h, int(nsteps), int(ckpt_every), int(avg_from),
                                                               int(axis), C.byref(j), _ptr(ft)))
        return j.value, ft

    def adjoint_backward(self) -> np.ndarray:
        """dJ/dphi per cell of the last adjoint_forward (float64, x fastest)."""
        g = np.zeros(self.setup.cells(), np.float64)
        _check(self._lib, self._lib.simlab_cfd_adjoint_backward(self._h, _ptr(g)))
        return g

    def set_timing(self, enable: bool):
        _check(self._lib, self._lib.simlab_cfd_lattice_set_timing(self._h, int(bool(enable))))

    def timing(self):
        """(step_ms, main_kernel_ms, main_launches) of the last run (HIP events on the lattice stream)."""
        a, b, c = C.c_double(), C.c_double(), C.c_uint64()
        _check(self._lib, self._lib.simlab_cfd_lattice_timing(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, int(c.value)


class Batch:
    """Several lattices stepped together, one launch per kernel and step
    (simlab_cfd_batch_*). The lattices must outlive the batch."""

    def __init__(self, lattices):
        self._lib = load_library()
        self.lattices = list(lattices)
        arr = (C.c_void_p * len(self.lattices))(*[l._h for l in self.lattices])
        h = C.c_void_p()
        _check(self._lib, self._lib.simlab_cfd_batch_new(arr, len(self.lattices), C.byref(h)))
        self._h = h

    def run_still(self, nsteps: int) -> np.ndarray:
        """(lattices, nsteps, 6) forces and moments per step."""
        out = np.zeros((len(self.lattices), int(nsteps), 6), np.float64)
        _check(self._lib, self._lib.simlab_cfd_batch_run_still(self._h, int(nsteps), _ptr(out)))
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.simlab_cfd_batch_free(h)
            self._h = None
