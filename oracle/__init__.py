"""CPU oracle (checker) for the pressure projection — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package. The product library never loads it.

Thin ctypes/numpy wrapper over ``liboracle_elq.so`` (built from
``elq_oracle.c`` by ``oracle/Makefile``). The C functions cite the reference
lines they restate; the scene generators below restate the reference's
emitters (``src/project/3d/fluid/levelset_emitter.cpp:10-32``) and its
simulator steps (``src/project/3d/eulerian_liquid/simulator.cpp:52-76,89-178``)
for the synthetic workloads.

Parity status of each function (how it is pinned) is listed in the test
modules ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ELQ_ORACLE_LIB selects another build of the same source (the sanitizer
# build of the Makefile's `sanitize` target)
_LIB_PATH = os.environ.get("ELQ_ORACLE_LIB") or os.path.join(_HERE, "liboracle_elq.so")
UNKNOWN_INDEX = 2147483647

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class _System(C.Structure):
    _fields_ = [("n", C.c_int)] + [(nm, C.c_void_p) for nm in
                                   ("diag", "ax", "ay", "az", "xp", "yp", "zp", "xm", "ym", "zm", "b")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        L.orc_fraction_inside2.restype = C.c_float
        L.orc_fraction_inside2.argtypes = [C.c_float, C.c_float]
        L.orc_fraction_inside4.restype = C.c_float
        L.orc_fraction_inside4.argtypes = [C.c_float] * 4
        L.orc_room_nodal.argtypes = [C.c_int, C.c_int, C.c_int, _f32p]
        L.orc_liquid_level_set.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                           C.c_void_p, _f32p]
        L.orc_compute_weights.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p, _f32p]
        L.orc_add_gravity.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, C.c_float]
        L.orc_find_valid_and_zero.argtypes = [C.c_int, C.c_int, C.c_int] + [_f32p] * 7
        L.orc_find_valid_velocity.argtypes = [C.c_int, C.c_int, C.c_int] + [_f32p] * 4 + [_u8p] * 3
        L.orc_advect_staggered.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float] + [_f32p] * 9
        L.orc_advect_velocity.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float] + [_f32p] * 6
        L.orc_advect_level_set.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float] + [_f32p] * 5
        L.orc_reinitialize.restype = C.c_long
        L.orc_reinitialize.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, _f32p, _u8p, C.POINTER(C.c_int)]
        L.orc_find_valid_level_set.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _u8p]
        L.orc_set_invalid_zero.argtypes = [C.c_size_t, _u8p, _f32p]
        L.orc_extrapolate.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _u8p, C.c_int]
        L.orc_extrapolate_const.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _u8p, C.c_int]
        L.orc_assign_indices.restype = C.c_int
        L.orc_assign_indices.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p, _f32p, _i32p]
        L.orc_build_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _i32p] + \
            [_f32p] * 7 + [C.POINTER(_System)]
        L.orc_to_csr.restype = C.c_int
        L.orc_to_csr.argtypes = [C.POINTER(_System), _i32p, _i32p, _f64p]
        L.orc_ic_factor.restype = C.c_int
        L.orc_ic_factor.argtypes = [C.POINTER(_System), C.c_double, C.c_double, _f64p]
        L.orc_ic_apply.argtypes = [C.POINTER(_System), _f64p, _f64p, _f64p]
        L.orc_spmv.argtypes = [C.POINTER(_System), _f64p, _f64p]
        L.orc_pcg.restype = C.c_int
        L.orc_pcg.argtypes = [C.POINTER(_System), _f64p, C.c_double, C.c_int, _f64p,
                              C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_build_grid_from_vector.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, _i32p, C.c_float, _f32p]
        L.orc_update_velocity.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float] + [_f32p] * 5
        L.orc_solve_pressure.restype = C.c_int
        L.orc_solve_pressure.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float] + [_f32p] * 7 + \
            [C.c_double, C.c_int, C.c_double, C.c_double, _i32p, _f32p,
             C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        L.orc_max_threads.restype = C.c_int
        L.orc_set_krylov_parallel.argtypes = [C.c_int]
        L.orc_set_krylov_parallel.restype = None
        L.orc_set_progress.argtypes = [C.c_int]
        L.orc_set_progress.restype = None
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_uniform01.restype = C.c_double
        L.orc_uniform01.argtypes = [C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


# --------------------------------------------------------------------------- threads
def set_threads(n: int, krylov_parallel: bool = False):
    """OpenMP threads of the oracle: the scatter and update loops run in
    parallel as in the reference (projection.cpp:85-134); with krylov_parallel
    also the SpMV, axpys and dots (a timing companion; the IC sweeps and the
    assembly stay serial)."""
    lib().orc_set_threads(int(n))
    lib().orc_set_krylov_parallel(1 if krylov_parallel else 0)


def set_progress(every: int):
    """orc_pcg reports its iteration count to stderr every `every`
    iterations (0: silent)."""
    lib().orc_set_progress(int(every))


def max_threads() -> int:
    return int(lib().orc_max_threads())


# --------------------------------------------------------------------------- scalar helpers
def fraction_inside2(a: float, b: float) -> float:
    return lib().orc_fraction_inside2(a, b)


def fraction_inside4(bl: float, br: float, tl: float, tr: float) -> float:
    return lib().orc_fraction_inside4(bl, br, tl, tr)


# --------------------------------------------------------------------------- scene
@dataclass
class Scene:
    """Inputs of one projection, reference layout (i fastest, float32)."""
    ni: int
    nj: int
    nk: int
    dx: float
    dt: float
    phi: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    uw: np.ndarray
    vw: np.ndarray
    ww: np.ndarray

    def copy(self) -> "Scene":
        return Scene(self.ni, self.nj, self.nk, self.dx, self.dt, self.phi.copy(), self.u.copy(),
                     self.v.copy(), self.w.copy(), self.uw.copy(), self.vw.copy(), self.ww.copy())


class Rng:
    """SplitMix64 stream (same definition as the GPU package's scene generator)."""

    def __init__(self, seed: int):
        self.state = C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def uniform(self) -> float:
        return lib().orc_uniform01(C.byref(self.state))

    def draw(self, lo: float, hi: float) -> np.float32:
        return np.float32(lo + (hi - lo) * self.uniform())


def _face_shapes(ni, nj, nk):
    return (nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)


def make_scene(ni, nj, nk, slab, spheres, dt=0.01, width=1.0):
    """create_grid (simulator.cpp:52-70) + emitters + init_scene (:72-76) +
    compute_weights (:187-189). Velocities start at zero."""
    dx = np.float32(np.float32(width) / np.float32(ni))
    L = lib()
    phi = np.empty(ni * nj * nk, np.float32)
    sph = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1))
    L.orc_liquid_level_set(ni, nj, nk, dx, np.float32(slab if slab is not None else -1e38),
                           len(sph) // 4, sph.ctypes.data if len(sph) else None, phi)
    nodal = np.empty((ni + 1) * (nj + 1) * (nk + 1), np.float32)
    L.orc_room_nodal(ni, nj, nk, nodal)
    uw = np.empty((ni + 1) * nj * nk, np.float32)
    vw = np.empty(ni * (nj + 1) * nk, np.float32)
    ww = np.empty(ni * nj * (nk + 1), np.float32)
    L.orc_compute_weights(ni, nj, nk, nodal, uw, vw, ww)
    z = lambda n: np.zeros(n, np.float32)
    return Scene(ni, nj, nk, float(dx), float(dt), phi, z(uw.size), z(vw.size), z(ww.size), uw, vw, ww)


def add_gravity(s: Scene, dt: float):
    lib().orc_add_gravity(s.ni, s.nj, s.nk, s.v, np.float32(dt))


def find_valid_and_zero(s: Scene):
    lib().orc_find_valid_and_zero(s.ni, s.nj, s.nk, s.phi, s.uw, s.vw, s.ww, s.u, s.v, s.w)


def find_valid_velocity(s: Scene):
    """0/1 flags (uint8) of the u, v and w faces."""
    out = tuple(np.empty(a.size, np.uint8) for a in (s.u, s.v, s.w))
    lib().orc_find_valid_velocity(s.ni, s.nj, s.nk, s.phi, s.uw, s.vw, s.ww, *out)
    return out


def extrapolate(grid: np.ndarray, valid: np.ndarray, dims, layers: int = 10):
    """extrapolation.cpp:5-46 in place on a flat float32 grid of dims
    (ni, nj, nk) and its uint8 0/1 flags."""
    ni, nj, nk = dims
    assert grid.size == valid.size == ni * nj * nk
    lib().orc_extrapolate(ni, nj, nk, grid, valid, layers)


def extrapolate_const(grid: np.ndarray, valid: np.ndarray, dims, layers: int = 10):
    ni, nj, nk = dims
    assert grid.size == valid.size == ni * nj * nk
    lib().orc_extrapolate_const(ni, nj, nk, grid, valid, layers)


def advect_velocity(s: Scene, dt: float):
    """Simulator::advect_velocity (simulator.cpp:79-82) in place on s.u/v/w."""
    u0, v0, w0 = s.u.copy(), s.v.copy(), s.w.copy()
    lib().orc_advect_velocity(s.ni, s.nj, s.nk, np.float32(s.dx), np.float32(dt), u0, v0, w0, s.u, s.v, s.w)


def advect_level_set(s: Scene, dt: float):
    """Simulator::advect_level_set (simulator.cpp:84-87) in place on s.phi."""
    phi0 = s.phi.copy()
    lib().orc_advect_level_set(s.ni, s.nj, s.nk, np.float32(s.dx), np.float32(dt), phi0, s.u, s.v, s.w, s.phi)


def reinitialize(phi: np.ndarray, valid: np.ndarray, dims, dx: float):
    """reinitialization.cpp:19-156 in place on a flat float32 level set;
    returns (FAR cells left, marching passes)."""
    ni, nj, nk = dims
    passes = C.c_int(0)
    left = lib().orc_reinitialize(ni, nj, nk, np.float32(dx), phi, valid, C.byref(passes))
    return int(left), passes.value


def random_velocity(s: Scene, rng: Rng, amp: float):
    for arr in (s.u, s.v, s.w):
        for c in range(arr.size):
            arr[c] = rng.draw(-amp, amp)


def scene_drop(n: int, dt: float = 0.01) -> Scene:
    """Slab y<0.4 plus a sphere r=0.15 at (0.5,0.7,0.5); zero velocity, one
    gravity step, validity + zeroing (apps/eulerian_liquid_3d.cpp:115-118)."""
    s = make_scene(n, n, n, 0.4, [(0.5, 0.7, 0.5, 0.15)], dt)
    add_gravity(s, dt)
    find_valid_and_zero(s)
    return s


def random_spheres(rng: Rng, count, c_lo, c_hi, r_lo, r_hi, lengths=(1.0, 1.0, 1.0)):
    out = []
    for _ in range(count):
        cx = rng.draw(c_lo * lengths[0], c_hi * lengths[0])
        cy = rng.draw(c_lo * lengths[1], c_hi * lengths[1])
        cz = rng.draw(c_lo * lengths[2], c_hi * lengths[2])
        r = rng.draw(r_lo, r_hi)
        out.append((cx, cy, cz, r))
    return out


def scene_spheres(n: int, seed: int, n_spheres: int = 32, dt: float = 0.01, vel_amp: float = 0.5) -> Scene:
    """Slab y<0.3 plus seeded spheres (centres U[0.15,0.85]^3, radii
    U[0.03,0.08]); face velocities U[-amp,amp] drawn u, v, w in linear order;
    then validity + zeroing. Python loops: small n only."""
    rng = Rng(seed)
    sph = random_spheres(rng, n_spheres, 0.15, 0.85, 0.03, 0.08)
    s = make_scene(n, n, n, 0.3, sph, dt)
    random_velocity(s, rng, vel_amp)
    find_valid_and_zero(s)
    return s


def clamp_geometry(n: int):
    """Slab and sphere of the clamp scene: the slab top lies 1e-5 above the
    centres of cell row j0 = floor(0.35 n), the sphere (centre at the centre
    of cell (n/2, n/2, n/2)) 1e-5 beyond the centres of the cells three cells
    from it on each axis, so the faces from those cells outwards have theta
    ~ 1e-4 < 0.01 (clamped) on every axis."""
    dx = np.float32(np.float32(1.0) / np.float32(n))
    c = np.float32((n // 2 + 0.5) * dx)
    slab = np.float32((int(0.35 * n) + 0.5) * dx + 1e-5)
    r = np.float32(3 * dx + 1e-5)
    return float(slab), [(c, c, c, r)]


def scene_clamp(n: int = 16, seed: int = 5, dt: float = 0.01) -> Scene:
    """Ghost-fluid theta clamp scene (projection.cpp:59-63,100-131): the
    clamp_geometry level set, face velocities U[-0.5, 0.5], validity + zeroing."""
    slab, sph = clamp_geometry(n)
    s = make_scene(n, n, n, slab, sph, dt)
    random_velocity(s, Rng(seed), 0.5)
    find_valid_and_zero(s)
    return s


# --------------------------------------------------------------------------- linear system
@dataclass
class System:
    n: int
    diag: np.ndarray
    ax: np.ndarray
    ay: np.ndarray
    az: np.ndarray
    b: np.ndarray
    xp: np.ndarray
    yp: np.ndarray
    zp: np.ndarray
    xm: np.ndarray
    ym: np.ndarray
    zm: np.ndarray

    def c(self) -> _System:
        st = _System()
        st.n = self.n
        for nm in ("diag", "ax", "ay", "az", "xp", "yp", "zp", "xm", "ym", "zm", "b"):
            setattr(st, nm, getattr(self, nm).ctypes.data)
        return st


def assign_indices(s: Scene):
    idx = np.empty(s.ni * s.nj * s.nk, np.int32)
    n = lib().orc_assign_indices(s.ni, s.nj, s.nk, s.phi, s.uw, s.vw, s.ww, idx)
    return n, idx


def build_matrix(s: Scene, indices: np.ndarray, n: int) -> System:
    f = lambda: np.zeros(n, np.float64)
    g = lambda: np.full(n, -1, np.int32)
    sys_ = System(n, f(), f(), f(), f(), f(), g(), g(), g(), g(), g(), g())
    st = sys_.c()
    lib().orc_build_matrix(s.ni, s.nj, s.nk, np.float32(s.dx), np.float32(s.dt), indices, s.phi,
                           s.u, s.v, s.w, s.uw, s.vw, s.ww, C.byref(st))
    return sys_


def to_csr(sys_: System):
    rp = np.zeros(sys_.n + 1, np.int32)
    col = np.zeros(7 * sys_.n, np.int32)
    val = np.zeros(7 * sys_.n, np.float64)
    st = sys_.c()
    nnz = lib().orc_to_csr(C.byref(st), rp, col, val)
    return rp, col[:nnz].copy(), val[:nnz].copy()


def ic_factor(sys_: System, tau: float = 0.0, sigma: float = 0.25):
    inv_e = np.zeros(sys_.n, np.float64)
    st = sys_.c()
    fixed = lib().orc_ic_factor(C.byref(st), tau, sigma, inv_e)
    return inv_e, fixed


def ic_apply(sys_: System, inv_e, r):
    z = np.zeros(sys_.n, np.float64)
    st = sys_.c()
    lib().orc_ic_apply(C.byref(st), inv_e, np.ascontiguousarray(r, np.float64), z)
    return z


def spmv(sys_: System, x):
    y = np.zeros(sys_.n, np.float64)
    st = sys_.c()
    lib().orc_spmv(C.byref(st), np.ascontiguousarray(x, np.float64), y)
    return y


def pcg(sys_: System, inv_e, tol=1e-4, max_iter=1000):
    x = np.zeros(sys_.n, np.float64)
    it = C.c_int(0)
    err = C.c_double(0)
    st = sys_.c()
    lib().orc_pcg(C.byref(st), inv_e, tol, max_iter, x, C.byref(it), C.byref(err))
    return x, it.value, err.value


def build_grid_from_vector(s: Scene, x, indices):
    p = np.empty(s.ni * s.nj * s.nk, np.float32)
    lib().orc_build_grid_from_vector(s.ni, s.nj, s.nk, np.ascontiguousarray(x, np.float64), indices,
                                     np.float32(0), p)
    return p


def update_velocity(s: Scene, pressure):
    lib().orc_update_velocity(s.ni, s.nj, s.nk, np.float32(s.dx), np.float32(s.dt), pressure, s.phi,
                              s.u, s.v, s.w)


@dataclass
class Projection:
    n: int
    indices: np.ndarray
    pressure: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    iterations: int
    error: float
    true_rel_res: float


def solve_pressure(s: Scene, tol=1e-4, max_iter=1000, tau=0.0, sigma=0.25) -> Projection:
    """Simulator::solve_pressure (simulator.cpp:96-112) on a copy of the scene."""
    t = s.copy()
    idx = np.empty(s.ni * s.nj * s.nk, np.int32)
    p = np.zeros(s.ni * s.nj * s.nk, np.float32)
    it = C.c_int(0)
    err = C.c_double(0)
    trr = C.c_double(0)
    n = lib().orc_solve_pressure(t.ni, t.nj, t.nk, np.float32(t.dx), np.float32(t.dt), t.phi, t.u, t.v, t.w,
                                 t.uw, t.vw, t.ww, tol, max_iter, tau, sigma, idx, p,
                                 C.byref(it), C.byref(err), C.byref(trr))
    return Projection(n, idx, p, t.u, t.v, t.w, it.value, err.value, trr.value)
