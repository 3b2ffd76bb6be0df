"""Python host binding of libelq.so (include/elq.h) — ctypes + numpy.

Mirrors the reference's projection interface
(``include/nama/project/3d/fluid/projection.h:10-22`` and
``Simulator::solve_pressure``, ``src/project/3d/eulerian_liquid/simulator.cpp:96-112``):

* :func:`solve_pressure` — the whole projection; updates ``u, v, w`` in place,
  returns ``(pressure, indices, stats)``; raises ``RuntimeError("No fluid")``
  like ``projection.cpp:24-26``.
* :meth:`Context.assign_indices`, :meth:`Context.build_matrix`,
  :meth:`Context.update_velocity_using_pressure` — the free functions.

Every numeric step runs on the GPU through the C ABI. Loading the library or
creating a context fails loudly (``RuntimeError``) when ``libelq.so`` is
missing or no gfx950 device is present; there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libelq.so")
UNKNOWN_INDEX = 2147483647

OK, NO_FLUID, NOT_CONVERGED, BREAKDOWN, DEVICE, ADMISSION, INVALID_ARG, BOUNDARY, BUSY = range(9)
STATUS_NAMES = {0: "OK", 1: "NO_FLUID", 2: "NOT_CONVERGED", 3: "BREAKDOWN", 4: "DEVICE", 5: "ADMISSION",
                6: "INVALID_ARG", 7: "BOUNDARY", 8: "BUSY"}
FP64, MIXED = 0, 1
NOALLOC_SCRATCH_BYTES = 64 << 20  # ELQ_NOALLOC_SCRATCH_BYTES
IC0, MIC0 = 0, 1
WAVEFRONT, LEVELS, WAVEFRONT_SPLIT, WAVEFRONT_FUSED = 0, 1, 2, 3


class Grid(C.Structure):
    _fields_ = [("ni", C.c_int32), ("nj", C.c_int32), ("nk", C.c_int32), ("dx", C.c_float),
                ("level_set", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
                ("u_weight", C.c_void_p), ("v_weight", C.c_void_p), ("w_weight", C.c_void_p),
                ("pressure", C.c_void_p), ("indices", C.c_void_p)]


class Opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int32), ("precision", C.c_int32),
                ("precond", C.c_int32), ("tau", C.c_double), ("sigma", C.c_double),
                ("inner_tol", C.c_double), ("schedule", C.c_int32), ("check_every", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n_unknowns", C.c_int64), ("n_bricks", C.c_int64), ("iterations", C.c_int32),
                ("outer_iterations", C.c_int32), ("error", C.c_double), ("true_rel_residual", C.c_double),
                ("status", C.c_int32), ("converged", C.c_int32), ("t_setup_ms", C.c_double),
                ("t_solve_ms", C.c_double), ("t_update_ms", C.c_double), ("t_total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_slots", C.c_int32), ("max_queue", C.c_int32),
                ("opts", Opts), ("batched", C.c_int32), ("batch_wait_us", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("iterations", C.c_int32),
                ("outer_iterations", C.c_int32), ("slot", C.c_int32), ("n_unknowns", C.c_int64),
                ("error", C.c_double), ("true_rel_residual", C.c_double), ("t_submit_ms", C.c_double),
                ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load():
    """Loads libelq.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libelq.so is missing at {LIB_PATH}: build it "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f32, f64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_size_t
    L.elq_default_opts.argtypes = [C.POINTER(Opts)]
    L.elq_default_opts.restype = None
    L.elq_ctx_create.argtypes = [C.c_int, sz, C.POINTER(vp)]
    L.elq_ctx_destroy.argtypes = [vp]
    L.elq_ctx_destroy.restype = None
    L.elq_ctx_last_error.argtypes = [vp]
    L.elq_ctx_last_error.restype = C.c_char_p
    L.elq_project.argtypes = [vp, C.POINTER(Grid), f32, C.POINTER(Opts), C.POINTER(Stats)]
    L.elq_project_device.argtypes = [vp, C.POINTER(Grid), f32, C.POINTER(Opts), C.POINTER(Stats)]
    L.elq_project_device_noalloc.argtypes = [vp, C.POINTER(Grid), f32, C.POINTER(Opts), C.POINTER(Stats)]
    L.elq_assign_indices.argtypes = [vp, C.POINTER(Grid), vp, C.POINTER(i64)]
    L.elq_build_matrix.argtypes = [vp, C.POINTER(Grid), f32, vp, vp, vp, vp, vp]
    L.elq_update_velocity.argtypes = [vp, C.POINTER(Grid), f32]
    L.elq_ic_factor.argtypes = [vp, C.POINTER(Grid), f32, C.POINTER(Opts), vp]
    L.elq_device_alloc.argtypes = [vp, sz, C.POINTER(vp)]
    L.elq_device_free.argtypes = [vp, vp]
    L.elq_memcpy_h2d.argtypes = [vp, vp, vp, sz]
    L.elq_memcpy_d2h.argtypes = [vp, vp, vp, sz]
    L.elq_memcpy_d2d.argtypes = [vp, vp, vp, sz]
    L.elq_compute_weights_device.argtypes = [vp, i32, i32, i32, vp, vp, vp, vp]
    L.elq_add_gravity_device.argtypes = [vp, i32, i32, i32, vp, f32]
    L.elq_valid_and_zero_device.argtypes = [vp, C.POINTER(Grid)]
    L.elq_find_valid_velocity_device.argtypes = [vp, C.POINTER(Grid), vp, vp, vp]
    L.elq_set_invalid_velocity_zero_device.argtypes = [vp, C.POINTER(Grid), vp, vp, vp]
    L.elq_find_valid_level_set_device.argtypes = [vp, i32, i32, i32, vp, vp]
    L.elq_extrapolate_device.argtypes = [vp, i32, i32, i32, vp, vp, i32]
    L.elq_extrapolate_const_device.argtypes = [vp, i32, i32, i32, vp, vp, i32]
    L.elq_extrapolate_velocity_device.argtypes = [vp, C.POINTER(Grid), vp, vp, vp, i32]
    L.elq_advect_staggered_device.argtypes = [vp, i32, i32, i32, f32, f32] + [vp] * 9
    L.elq_advect_collocated_device.argtypes = [vp, i32, i32, i32, f32, f32] + [vp] * 5
    L.elq_advect_velocity_device.argtypes = [vp, C.POINTER(Grid), f32]
    L.elq_advect_level_set_device.argtypes = [vp, i32, i32, i32, f32, f32, vp, vp, vp, vp]
    L.elq_reinitialize_device.argtypes = [vp, i32, i32, i32, f32, vp, vp, C.POINTER(i64), C.POINTER(i32)]
    L.elq_set_profiling_filter.argtypes = [vp, C.c_char_p]
    L.elq_set_profiling_filter.restype = None
    L.elq_last_kernel_times.argtypes = [vp, C.c_char_p, sz, C.POINTER(f64), C.POINTER(i64), C.c_int]
    L.elq_set_profiling.argtypes = [vp, C.c_int]
    L.elq_set_profiling.restype = None
    L.elq_set_max_groups.argtypes = [vp, C.c_int]
    L.elq_set_wait_limit.argtypes = [vp, f64]
    L.elq_set_reinit_rebase.argtypes = [vp, i32]
    L.elq_set_graph_capture.argtypes = [vp, C.c_int]
    L.elq_server_create.argtypes = [C.c_int, C.POINTER(ServerOpts), C.POINTER(vp)]
    L.elq_server_destroy.argtypes = [vp]
    L.elq_server_destroy.restype = None
    L.elq_server_submit.argtypes = [vp, C.POINTER(Grid), f32, C.POINTER(i64)]
    L.elq_server_wait.argtypes = [vp, C.c_int]
    L.elq_server_poll.argtypes = [vp, C.POINTER(Result), C.c_int]
    L.elq_server_in_flight.argtypes = [vp]
    L.elq_server_device_bytes.argtypes = [vp]
    L.elq_server_device_bytes.restype = sz
    L.elq_server_batch_stats.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.elq_projection_footprint.argtypes = [i32, i32, i32, i32, vp]
    L.elq_projection_footprint.restype = sz
    L.elq_ctx_device_bytes.argtypes = [vp]
    L.elq_ctx_device_bytes.restype = sz
    L.elq_scene_level_set.argtypes = [i32, i32, i32, f32, f32, i32, vp, vp]
    L.elq_scene_weights.argtypes = [i32, i32, i32, vp, vp, vp]
    L.elq_scene_room_nodal.argtypes = [i32, i32, i32, vp]
    L.elq_scene_add_gravity.argtypes = [i32, i32, i32, vp, f32]
    L.elq_scene_valid_and_zero.argtypes = [C.POINTER(Grid)]
    L.elq_scene_random_spheres.argtypes = [C.POINTER(C.c_uint64), i32, f64, f64, f64, f64, f64, f64, f64, vp]
    L.elq_scene_random_velocity.argtypes = [C.POINTER(C.c_uint64), i32, i32, i32, f64, vp, vp, vp]
    L.elq_uniform01.argtypes = [C.POINTER(C.c_uint64)]
    L.elq_uniform01.restype = f64
    _lib = L
    return L


def default_opts(**kw) -> Opts:
    o = Opts()
    load().elq_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def _ptr(a):
    return a.ctypes.data if a is not None else None


# --------------------------------------------------------------------------- scenes (host set-up steps)
@dataclass
class Scene:
    """Inputs of one projection in the reference layout (float32, i fastest)."""
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
    meta: dict = field(default_factory=dict)

    def copy(self) -> "Scene":
        return Scene(self.ni, self.nj, self.nk, self.dx, self.dt, self.phi.copy(), self.u.copy(),
                     self.v.copy(), self.w.copy(), self.uw.copy(), self.vw.copy(), self.ww.copy(),
                     dict(self.meta))

    def grid(self, pressure=None, indices=None) -> Grid:
        g = Grid()
        g.ni, g.nj, g.nk, g.dx = self.ni, self.nj, self.nk, self.dx
        g.level_set, g.u, g.v, g.w = _ptr(self.phi), _ptr(self.u), _ptr(self.v), _ptr(self.w)
        g.u_weight, g.v_weight, g.w_weight = _ptr(self.uw), _ptr(self.vw), _ptr(self.ww)
        g.pressure, g.indices = _ptr(pressure), _ptr(indices)
        return g

    @property
    def cells(self):
        return self.ni * self.nj * self.nk


def make_scene(ni, nj, nk, slab, spheres, dt=0.01, width=1.0) -> Scene:
    """create_grid + emitters + init_scene + compute_weights; zero velocity."""
    L = load()
    dx = float(np.float32(np.float32(width) / np.float32(ni)))
    phi = np.empty(ni * nj * nk, np.float32)
    sph = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1))
    rc = L.elq_scene_level_set(ni, nj, nk, dx, np.float32(slab if slab is not None else -1e38),
                               len(sph) // 4, _ptr(sph) if len(sph) else None, _ptr(phi))
    _check_rc(rc, None)
    uw = np.empty((ni + 1) * nj * nk, np.float32)
    vw = np.empty(ni * (nj + 1) * nk, np.float32)
    ww = np.empty(ni * nj * (nk + 1), np.float32)
    _check_rc(L.elq_scene_weights(ni, nj, nk, _ptr(uw), _ptr(vw), _ptr(ww)), None)
    z = lambda n: np.zeros(n, np.float32)
    return Scene(ni, nj, nk, dx, dt, phi, z(uw.size), z(vw.size), z(ww.size), uw, vw, ww,
                 {"slab": slab, "spheres": [tuple(map(float, s)) for s in np.asarray(spheres).reshape(-1, 4)]})


def add_gravity(s: Scene, dt: float):
    _check_rc(load().elq_scene_add_gravity(s.ni, s.nj, s.nk, _ptr(s.v), np.float32(dt)), None)


def valid_and_zero(s: Scene):
    g = s.grid()
    _check_rc(load().elq_scene_valid_and_zero(C.byref(g)), None)


class Rng:
    """SplitMix64 stream used by every synthetic workload."""

    def __init__(self, seed: int):
        self.state = C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def uniform(self) -> float:
        return load().elq_uniform01(C.byref(self.state))

    def spheres(self, count, c_lo, c_hi, r_lo, r_hi, lengths=(1.0, 1.0, 1.0)):
        out = np.zeros(4 * count, np.float32)
        _check_rc(load().elq_scene_random_spheres(C.byref(self.state), count, c_lo, c_hi, r_lo, r_hi,
                                                  lengths[0], lengths[1], lengths[2],
                                                  _ptr(out) if count else None), None)
        return out.reshape(-1, 4)

    def velocity(self, s: Scene, amp: float):
        _check_rc(load().elq_scene_random_velocity(C.byref(self.state), s.ni, s.nj, s.nk, amp, _ptr(s.u),
                                                   _ptr(s.v), _ptr(s.w)), None)


def scene_drop(n: int, dt: float = 0.01) -> Scene:
    """Slab y < 0.4 plus a sphere of radius 0.15 at (0.5, 0.7, 0.5); zero
    velocity, one gravity step, then find_valid_velocity and
    set_invalid_velocity_zero (apps/eulerian_liquid_3d.cpp:115-118)."""
    s = make_scene(n, n, n, 0.4, [(0.5, 0.7, 0.5, 0.15)], dt)
    add_gravity(s, dt)
    valid_and_zero(s)
    s.meta["workload"] = f"drop{n}"
    return s


def scene_spheres(n: int, seed: int, n_spheres: int = 32, dt: float = 0.01, vel_amp: float = 0.5) -> Scene:
    """Slab y < 0.3 plus seeded spheres (centres U[0.15,0.85]^3, radii
    U[0.03,0.08]); face velocities U[-amp, amp]; validity + zeroing."""
    rng = Rng(seed)
    sph = rng.spheres(n_spheres, 0.15, 0.85, 0.03, 0.08)
    s = make_scene(n, n, n, 0.3, sph, dt)
    rng.velocity(s, vel_amp)
    valid_and_zero(s)
    s.meta["workload"] = f"spheres{n}_s{seed}_k{n_spheres}"
    return s


def scene_clamp(n: int = 16, seed: int = 5, dt: float = 0.01) -> Scene:
    """A slab whose top lies 1e-5 above the centres of cell row floor(0.35 n)
    and a sphere 1e-5 beyond the centres of the cells three cells from its
    centre cell (n/2, n/2, n/2) on each axis: the faces from those cells
    outwards have theta ~ 1e-4, below the ghost-fluid clamp of 0.01
    (projection.cpp:59-63,100-131), on every axis. Face velocities
    U[-0.5, 0.5]; validity + zeroing."""
    dx = np.float32(np.float32(1.0) / np.float32(n))
    c = np.float32((n // 2 + 0.5) * dx)
    slab = np.float32((int(0.35 * n) + 0.5) * dx + 1e-5)
    r = np.float32(3 * dx + 1e-5)
    s = make_scene(n, n, n, float(slab), [(c, c, c, r)], dt)
    Rng(seed).velocity(s, 0.5)
    valid_and_zero(s)
    s.meta["workload"] = f"clamp{n}"
    return s


SERVE_DIMS = (32, 48, 64, 96, 128)


def scene_request(seed: int, dt: float = 0.01) -> Scene:
    """One serving request: dims per axis from {32,48,64,96,128}, slab height
    U[0.2,0.45] of the y extent, 1-8 spheres (centres U[0.2,0.8] of each
    extent, radii U[0.05,0.2]), face velocities U[-1,1]."""
    rng = Rng(seed)
    dims = [SERVE_DIMS[min(4, int(rng.uniform() * 5))] for _ in range(3)]
    ni, nj, nk = dims
    lengths = (1.0, nj / ni, nk / ni)
    nsph = 1 + min(7, int(rng.uniform() * 8))
    slab = np.float32(0.2 * lengths[1] + (0.45 - 0.2) * lengths[1] * rng.uniform())
    sph = rng.spheres(nsph, 0.2, 0.8, 0.05, 0.2, lengths)
    s = make_scene(ni, nj, nk, float(slab), sph, dt)
    rng.velocity(s, 1.0)
    valid_and_zero(s)
    s.meta["workload"] = f"request_s{seed}_{ni}x{nj}x{nk}"
    return s


# --------------------------------------------------------------------------- teardown order
# Library objects still alive when the interpreter exits are destroyed here,
# in an explicit order, while the HIP runtime is still intact: servers first
# (elq_server_destroy joins their slot threads, which run HIP calls and hold
# per-thread runtime state), then contexts. Python runs this hook before it
# tears down modules, and long before exit() runs the runtime's own static
# destructors; left to the garbage collector, a server could outlive the
# interpreter with its threads still attached to the runtime, or be destroyed
# at an arbitrary point of the shutdown. __del__ does nothing once the hook
# has started.
_live_servers: "weakref.WeakSet" = weakref.WeakSet()
_live_contexts: "weakref.WeakSet" = weakref.WeakSet()
_shutting_down = False


def _shutdown():
    global _shutting_down
    _shutting_down = True
    for srv in list(_live_servers):
        srv.close()
    for ctx in list(_live_contexts):
        ctx.close()


atexit.register(_shutdown)


# --------------------------------------------------------------------------- context
def _check_rc(rc, ctx):
    if rc == OK:
        return
    msg = STATUS_NAMES.get(rc, str(rc))
    if ctx is not None:
        msg += ": " + load().elq_ctx_last_error(ctx).decode()
    if rc == NO_FLUID:
        raise RuntimeError("No fluid")
    raise ElqError(rc, msg)


class ElqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Context:
    """One device context (owns the library's device buffers)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = load()
        h = C.c_void_p()
        rc = L.elq_ctx_create(device, hbm_budget, C.byref(h))
        if rc != OK:
            raise RuntimeError(f"elq_ctx_create failed ({STATUS_NAMES.get(rc, rc)}): no usable gfx950 device")
        self.h = h
        self.L = L
        _live_contexts.add(self)

    def close(self):
        if getattr(self, "h", None):
            self.L.elq_ctx_destroy(self.h)
            self.h = None
        _live_contexts.discard(self)

    def __del__(self):
        if _shutting_down:
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self) -> str:
        return self.L.elq_ctx_last_error(self.h).decode()

    def device_bytes(self) -> int:
        """Device bytes the context's workspace holds."""
        return int(self.L.elq_ctx_device_bytes(self.h))

    # ---- projection
    def solve_pressure(self, s: Scene, opts: Opts | None = None, raise_on_error=True):
        """Projects s in place (u, v, w updated); returns (pressure, indices, stats)."""
        p = np.zeros(s.cells, np.float32)
        idx = np.zeros(s.cells, np.int32)
        st = Stats()
        g = s.grid(p, idx)
        rc = self.L.elq_project(self.h, C.byref(g), np.float32(s.dt), C.byref(opts) if opts else None, C.byref(st))
        if raise_on_error and rc not in (OK, NOT_CONVERGED):
            _check_rc(rc, self.h)
        return p, idx, st

    def project_device(self, g: Grid, dt: float, opts: Opts | None = None) -> tuple[int, Stats]:
        st = Stats()
        rc = self.L.elq_project_device(self.h, C.byref(g), np.float32(dt), C.byref(opts) if opts else None,
                                       C.byref(st))
        return rc, st

    def project_device_noalloc(self, g: Grid, dt: float, opts: Opts | None = None) -> tuple[int, Stats]:
        """elq_project_device_noalloc: the same projection with its whole
        workspace in the library's static scratch array (no device allocation)."""
        st = Stats()
        rc = self.L.elq_project_device_noalloc(self.h, C.byref(g), np.float32(dt), C.byref(opts) if opts else None,
                                               C.byref(st))
        return rc, st

    def assign_indices(self, s: Scene):
        idx = np.zeros(s.cells, np.int32)
        n = C.c_int64(0)
        g = s.grid()
        rc = self.L.elq_assign_indices(self.h, C.byref(g), _ptr(idx), C.byref(n))
        _check_rc(rc, self.h)
        return n.value, idx

    def build_matrix(self, s: Scene, n: int):
        out = [np.zeros(n, np.float64) for _ in range(5)]
        g = s.grid()
        rc = self.L.elq_build_matrix(self.h, C.byref(g), np.float32(s.dt), *[_ptr(a) for a in out])
        _check_rc(rc, self.h)
        return dict(zip(("diag", "ax", "ay", "az", "b"), out))

    def ic_factor(self, s: Scene, n: int, opts: Opts | None = None):
        """IC(0)/MIC(0) pivot inverses 1/e (FP64) in the reference's unknown numbering."""
        out = np.zeros(n, np.float64)
        g = s.grid()
        rc = self.L.elq_ic_factor(self.h, C.byref(g), np.float32(s.dt), C.byref(opts) if opts else None, _ptr(out))
        _check_rc(rc, self.h)
        return out

    def update_velocity_using_pressure(self, s: Scene, pressure: np.ndarray):
        pressure = np.ascontiguousarray(pressure, np.float32)
        g = s.grid(pressure)
        _check_rc(self.L.elq_update_velocity(self.h, C.byref(g), np.float32(s.dt)), self.h)

    # ---- device residency
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        _check_rc(self.L.elq_device_alloc(self.h, nbytes, C.byref(p)), self.h)
        return p.value

    def free(self, ptr: int):
        _check_rc(self.L.elq_device_free(self.h, ptr), self.h)

    def h2d(self, dst: int, a: np.ndarray):
        _check_rc(self.L.elq_memcpy_h2d(self.h, dst, _ptr(a), a.nbytes), self.h)

    def d2h(self, a: np.ndarray, src: int):
        _check_rc(self.L.elq_memcpy_d2h(self.h, _ptr(a), src, a.nbytes), self.h)

    def d2d(self, dst: int, src: int, nbytes: int):
        _check_rc(self.L.elq_memcpy_d2d(self.h, dst, src, nbytes), self.h)

    def extrapolate(self, grid: np.ndarray, valid: np.ndarray, dims, layers: int = 10, const: bool = False):
        """fluid::extrapolate (or extrapolate_const) of a host float32 grid of
        dims (ni, nj, nk) and its uint8 0/1 flags, in place, through device
        copies and elq_extrapolate[_const]_device."""
        ni, nj, nk = dims
        assert grid.dtype == np.float32 and valid.dtype == np.uint8 and grid.size == valid.size == ni * nj * nk
        pg, pv = self.alloc(max(grid.nbytes, 4)), self.alloc(max(valid.nbytes, 4))
        try:
            self.h2d(pg, grid)
            self.h2d(pv, valid)
            fn = self.L.elq_extrapolate_const_device if const else self.L.elq_extrapolate_device
            _check_rc(fn(self.h, ni, nj, nk, pg, pv, layers), self.h)
            self.d2h(grid, pg)
            self.d2h(valid, pv)
        finally:
            self.free(pg)
            self.free(pv)

    def reinitialize(self, phi: np.ndarray, valid: np.ndarray, dims, dx: float):
        """fluid::reinitialize_fast_marching of a host float32 level set of
        dims (ni, nj, nk) with uint8 0/1 flags, in place, through device copies
        and elq_reinitialize_device; returns (unreached cells, passes)."""
        ni, nj, nk = dims
        assert phi.dtype == np.float32 and valid.dtype == np.uint8 and phi.size == valid.size == ni * nj * nk
        pp, pv = self.alloc(max(phi.nbytes, 4)), self.alloc(max(valid.nbytes, 4))
        try:
            self.h2d(pp, phi)
            self.h2d(pv, valid)
            left, passes = C.c_int64(0), C.c_int32(0)
            _check_rc(self.L.elq_reinitialize_device(self.h, ni, nj, nk, np.float32(dx), pp, pv, C.byref(left),
                                                     C.byref(passes)), self.h)
            self.d2h(phi, pp)
        finally:
            self.free(pp)
            self.free(pv)
        return int(left.value), int(passes.value)

    def set_max_groups(self, max_groups: int):
        """Caps the persistent kernels' grid (0: whole device)."""
        _check_rc(self.L.elq_set_max_groups(self.h, int(max_groups)), self.h)

    def set_graph_capture(self, on: bool):
        """Graph replay of the Krylov loop's poll intervals (default off)."""
        _check_rc(self.L.elq_set_graph_capture(self.h, 1 if on else 0), self.h)

    def set_wait_limit(self, seconds: float):
        """Bound of every device-side wait (default 4 s); see elq_set_wait_limit."""
        _check_rc(self.L.elq_set_wait_limit(self.h, float(seconds)), self.h)

    def set_reinit_rebase(self, passes: int):
        """Stamp rebase interval of the reinitialisation's tile route
        (default 60000); see elq_set_reinit_rebase."""
        _check_rc(self.L.elq_set_reinit_rebase(self.h, int(passes)), self.h)

    def set_profiling(self, on: bool, only: str = ""):
        """Per-kernel event timing; only: comma-separated kernel classes to time (all if empty)."""
        self.L.elq_set_profiling_filter(self.h, only.encode())
        self.L.elq_set_profiling(self.h, 1 if on else 0)

    def kernel_times(self) -> dict:
        names = C.create_string_buffer(4096)
        ms = (C.c_double * 64)()
        cnt = (C.c_int64 * 64)()
        n = self.L.elq_last_kernel_times(self.h, names, 4096, ms, cnt, 64)
        keys = names.value.decode().split(",") if n else []
        return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}


class DeviceScene:
    """A scene resident in device memory (inputs already in HBM)."""

    FIELDS = ("phi", "u", "v", "w", "uw", "vw", "ww")

    def __init__(self, ctx: Context, s: Scene):
        self.ctx, self.s = ctx, s
        self.ptr = {}
        for f in self.FIELDS:
            a = getattr(s, f)
            self.ptr[f] = ctx.alloc(a.nbytes)
            ctx.h2d(self.ptr[f], a)
        self.ptr["pressure"] = ctx.alloc(s.cells * 4)
        self.ptr["indices"] = ctx.alloc(s.cells * 4)

    def reset_velocity(self):
        for f in ("u", "v", "w"):
            self.ctx.h2d(self.ptr[f], getattr(self.s, f))

    def grid(self) -> Grid:
        s, p = self.s, self.ptr
        g = Grid()
        g.ni, g.nj, g.nk, g.dx = s.ni, s.nj, s.nk, s.dx
        g.level_set, g.u, g.v, g.w = p["phi"], p["u"], p["v"], p["w"]
        g.u_weight, g.v_weight, g.w_weight = p["uw"], p["vw"], p["ww"]
        g.pressure, g.indices = p["pressure"], p["indices"]
        return g

    def fetch(self, name: str, dtype=np.float32, count=None):
        n = count if count is not None else (self.s.cells if name in ("pressure", "indices", "phi")
                                             else getattr(self.s, name).size)
        a = np.empty(n, dtype)
        self.ctx.d2h(a, self.ptr[name])
        return a

    def close(self):
        for p in self.ptr.values():
            self.ctx.free(p)
        self.ptr = {}


def footprint(s: Scene, precision: int = FP64, bound_by_level_set: bool = True) -> int:
    """Device bytes a projection of s requests on a fresh context (upper bound)."""
    phi = _ptr(s.phi) if bound_by_level_set else None
    return int(load().elq_projection_footprint(s.ni, s.nj, s.nk, precision, phi))


class Server:
    """Continuous batching of projections (elq_server_*). Each submitted scene
    is projected in place (u, v, w) and its pressure returned at poll time; the
    scene must not be touched until then."""

    def __init__(self, device: int = 0, max_slots: int = 4, max_queue: int = 4096, hbm_budget: int = 0,
                 opts: Opts | None = None, batched: int = 0, batch_wait_us: int = 0):
        self.L = load()
        so = ServerOpts()
        so.hbm_budget_bytes = hbm_budget
        so.max_slots = max_slots
        so.max_queue = max_queue
        so.opts = opts if opts is not None else default_opts()
        so.batched = int(batched)
        so.batch_wait_us = int(batch_wait_us)
        h = C.c_void_p()
        rc = self.L.elq_server_create(device, C.byref(so), C.byref(h))
        if rc != OK:
            raise ElqError(rc, f"elq_server_create: {STATUS_NAMES.get(rc, rc)}")
        self.h = h
        self.live = {}
        _live_servers.add(self)

    def submit(self, s: Scene) -> int:
        p = np.zeros(s.cells, np.float32)
        g = s.grid(p, None)
        rid = C.c_int64()
        rc = self.L.elq_server_submit(self.h, C.byref(g), np.float32(s.dt), C.byref(rid))
        if rc != OK:
            raise ElqError(rc, f"elq_server_submit: {STATUS_NAMES.get(rc, rc)}")
        self.live[rid.value] = (s, p, g)
        return rid.value

    def wait(self, timeout_ms: int = -1) -> int:
        return int(self.L.elq_server_wait(self.h, int(timeout_ms)))

    def poll(self, cap: int = 256) -> list:
        """Finished requests: dicts with the result fields plus scene and pressure."""
        buf = (Result * cap)()
        n = self.L.elq_server_poll(self.h, buf, cap)
        out = []
        for q in range(n):
            d = buf[q].as_dict()
            s, p, _ = self.live.pop(d["id"])
            d["scene"], d["pressure"] = s, p
            out.append(d)
        return out

    def in_flight(self) -> int:
        return int(self.L.elq_server_in_flight(self.h))

    def device_bytes(self) -> int:
        return int(self.L.elq_server_device_bytes(self.h))

    def batch_stats(self) -> tuple[int, int]:
        """(batched Krylov chunks launched, requests they advanced)."""
        b, j = C.c_int64(), C.c_int64()
        _check_rc(self.L.elq_server_batch_stats(self.h, C.byref(b), C.byref(j)), None)
        return b.value, j.value

    def close(self):
        if getattr(self, "h", None) is not None:
            self.L.elq_server_destroy(self.h)
            self.h = None
            self.live = {}
        _live_servers.discard(self)

    def __del__(self):
        if _shutting_down:
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def solve_pressure(s: Scene, opts: Opts | None = None, ctx: Context | None = None):
    """Simulator::solve_pressure(dt) on the GPU (simulator.cpp:96-112)."""
    own = ctx is None
    ctx = ctx or Context()
    try:
        return ctx.solve_pressure(s, opts)
    finally:
        if own:
            ctx.close()
