"""Benchmark of the pressure projection on one MI355X.

Workload (default): the 256^3 seeded scene — slab y < 0.3 plus 32 spheres
(centres U[0.15,0.85]^3, radii U[0.03,0.08], SplitMix64 seed 1), face
velocities U[-0.5,0.5], validity + zeroing; projected to relative residual
1e-8 in mixed precision (FP32 SpMV/preconditioner/vectors, FP64 dots/scalars,
FP64 iterative refinement), natural-order IC(0), wavefront-scheduled sweeps.
One step = one full projection (indices, assembly, factorisation, PCG,
pressure scatter, velocity update) on inputs already resident in HBM.

Prints one JSON line: solves/s, the dominant kernel's roofline (HIP events on
the library's stream, algorithmic bytes per launch), the FP64 GPU drop-in on
the same inputs, max|p - p_FP64| and max|p - p_oracle|, the largest weighted
divergence left in a fluid cell relative to the input's largest (both
precisions), the serving stream, the other configurations (`configs`: 64^3
and 128^3 against the oracle's 1e-12 solve, 512^3 against the FP64 drop-in),
and the CPU oracle timed on full projections (`cpu_baseline`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "eulerian-liquid_amd"))

import elq  # noqa: E402

PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per unknown per launch of each kernel class: s_m bytes per
# stored coefficient (diag, ax, ay, az, inv_e), s_v bytes per Krylov element.
#   spmv (+ p update, p.q):           4 s_m + 2 s_v + 3 s_v
#   fwd sweep (+ x, r axpy, |r|^2):   4 s_m + 2 s_v + 6 s_v
#   bwd sweep (+ r.z):                4 s_m + 3 s_v
BYTES_PER_UNKNOWN = {
    "spmv": lambda sm, sv: 4 * sm + 5 * sv,
    "fwd_sweep_update": lambda sm, sv: 4 * sm + 8 * sv,
    "fwd_sweep": lambda sm, sv: 4 * sm + 2 * sv,
    "bwd_sweep": lambda sm, sv: 4 * sm + 3 * sv,
}
B_ITER = lambda sm, sv: 12 * sm + 16 * sv  # per PCG iteration and unknown
B_OUTER = 64  # per refinement step and unknown (FP64 residual + x update)


WORKLOADS = {
    "spheres256": lambda seed: elq.scene_spheres(256, seed=seed, n_spheres=32),
    "drop128": lambda seed: elq.scene_drop(128),
    "drop64": lambda seed: elq.scene_drop(64),
    "spheres512": lambda seed: elq.scene_spheres(512, seed=seed, n_spheres=128),
}


def workload(name: str, seed: int):
    if name not in WORKLOADS:
        raise SystemExit(f"unknown workload {name} (one of {', '.join(WORKLOADS)})")
    return WORKLOADS[name](seed), name


def project(ctx, ds, opts):
    ds.reset_velocity()
    t0 = time.perf_counter()
    rc, st = ctx.project_device(ds.grid(), ds.s.dt, opts)
    dt_ms = (time.perf_counter() - t0) * 1e3
    if rc != elq.OK:
        raise SystemExit(f"projection failed: {elq.STATUS_NAMES.get(rc)} {ctx.last_error()}")
    return dt_ms, st


def weighted_divergence(s, u, v, w):
    """Cell-centred weighted divergence sum over faces of +/- weight * u / dx
    (the right-hand side of build_matrix up to its -1/dx factor), FP64."""
    ni, nj, nk = s.ni, s.nj, s.nk
    uu = (u * s.uw).astype(np.float64).reshape(nk, nj, ni + 1)
    vv = (v * s.vw).astype(np.float64).reshape(nk, nj + 1, ni)
    ww = (w * s.ww).astype(np.float64).reshape(nk + 1, nj, ni)
    return (((uu[:, :, 1:] - uu[:, :, :-1]) + (vv[:, 1:, :] - vv[:, :-1, :]) + (ww[1:] - ww[:-1])) / s.dx).reshape(-1)


def max_divergence(s, u, v, w, mask=None, slab=32):
    """Largest |weighted divergence| over the cells (where mask, if given),
    computed in k slabs so that large grids need little extra host memory."""
    ni, nj, nk = s.ni, s.nj, s.nk
    uu, vv, ww = u.reshape(nk, nj, ni + 1), v.reshape(nk, nj + 1, ni), w.reshape(nk + 1, nj, ni)
    uw, vw, wwt = s.uw.reshape(nk, nj, ni + 1), s.vw.reshape(nk, nj + 1, ni), s.ww.reshape(nk + 1, nj, ni)
    m = mask.reshape(nk, nj, ni) if mask is not None else None
    best = 0.0
    for k0 in range(0, nk, slab):
        k1 = min(nk, k0 + slab)
        a = (uu[k0:k1] * uw[k0:k1]).astype(np.float64)
        b = (vv[k0:k1] * vw[k0:k1]).astype(np.float64)
        c = (ww[k0:k1 + 1] * wwt[k0:k1 + 1]).astype(np.float64)
        d = np.abs(((a[:, :, 1:] - a[:, :, :-1]) + (b[:, 1:, :] - b[:, :-1, :]) + (c[1:] - c[:-1])) / s.dx)
        if m is not None:
            d = d[m[k0:k1]]
        if d.size:
            best = max(best, float(d.max()))
    return best


def fluid_divergence(ds):
    """After the last projection on ds: the largest |weighted divergence| over
    the fluid cells (those with an index), relative to the largest over all
    cells of the input velocity."""
    s = ds.s
    fluid = ds.fetch("indices", np.int32) != elq.UNKNOWN_INDEX
    before = max_divergence(s, s.u, s.v, s.w)
    after = max_divergence(s, ds.fetch("u"), ds.fetch("v"), ds.fetch("w"), fluid)
    return after / before if before else 0.0


def run_steps(ctx, ds, opts, steps, warmup):
    """The warm-up steps (the first one also allocates the context's device
    buffers), one untimed step with HIP events around every kernel class (the
    per-class breakdown), then the timed steps with events around the
    dominant kernel class only (its average launch duration gives the
    roofline; the other classes run without events)."""
    for _ in range(warmup):
        project(ctx, ds, opts)
    ctx.set_profiling(True)
    project(ctx, ds, opts)
    breakdown = ctx.kernel_times()
    ctx.set_profiling(False)
    dominant = max((k for k in breakdown if k in BYTES_PER_UNKNOWN), key=lambda k: breakdown[k][0])
    ctx.set_profiling(True, only=dominant)
    times, stats, dom = [], [], [0.0, 0]
    for _ in range(steps):
        dt_ms, st = project(ctx, ds, opts)
        times.append(dt_ms)
        stats.append(st.as_dict())
        ms, n = ctx.kernel_times().get(dominant, (0.0, 0))
        dom[0] += ms
        dom[1] += n
    ctx.set_profiling(False)
    return times, stats, breakdown, (dominant, dom[0], dom[1])


def class_rates(breakdown, n_unknowns, sm, sv):
    """Per kernel class of the profiled (breakdown) step: launches, average
    launch time and the algorithmic bytes per launch over it, in GB/s and as
    a fraction of peak. Averages cover every launch of the class, as the
    roofline's do, including the few queued past convergence that exit at
    once."""
    out = {}
    for k, (ms, launches) in sorted(breakdown.items()):
        if k not in BYTES_PER_UNKNOWN or launches <= 0 or ms <= 0:
            continue
        avg_s = ms / launches / 1e3
        gbs = BYTES_PER_UNKNOWN[k](sm, sv) * n_unknowns / avg_s / 1e9
        out[k] = {"launches": int(launches), "avg_us": round(avg_s * 1e6, 2), "achieved_gbs": round(gbs, 1),
                  "frac": round(gbs / PEAK_HBM_GBS, 4)}
    return out


def roofline(dom, n_unknowns, sm, sv):
    kernel, ms, launches = dom
    avg_s = ms / launches / 1e3
    bytes_per_launch = BYTES_PER_UNKNOWN[kernel](sm, sv) * n_unknowns
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None, "kernel": kernel,
            "avg_launch_us": round(avg_s * 1e6, 2), "launches_timed": launches,
            "algorithmic_bytes_per_launch": int(bytes_per_launch)}


def profiled_traffic(wname, precision, kernel):
    """HBM bytes per launch of a kernel class on this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (counters scaled by a
    plain copy kernel of known size); None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"pcg_{wname}_{precision}_hbm_traffic.txt")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch "):
                v = json.loads(line.split(" ", 1)[1]).get(kernel)
                return v, os.path.relpath(path, ROOT)
    return None, None


def frame_steps(ctx, ds, reps=5):
    """The frame steps around the projection on the workload's grids, device
    resident, each timed with HIP events on the library's stream (median of
    reps calls; the inputs are restored between calls, untimed):
    compute_weights from the room() solid, add_gravity, find_valid_velocity,
    set_invalid_velocity_zero, and after the projection extrapolate_velocity
    (10 layers, the three face grids), find_valid_level_set + extrapolate_const
    of the level set, and the semi-Lagrangian advection of the level set and
    of the velocity (kernel time), and the reinitialisation of the level set
    (events over its kernels; wall_ms adds the host's pass-count reads).
    Algorithmic bytes: each input and
    output array once (a cached neighbour read counts once); for the
    extrapolation, every flag read once plus, per extrapolated cell, its
    value and flag written and one neighbour value read (4 + 1 + 4 B); for
    the reinitialisation, the level set and flags read and the level set
    written once (9 B per cell)."""
    s, L, h = ds.s, ctx.L, ctx.h
    ni, nj, nk = s.ni, s.nj, s.nk
    cells = s.cells
    nf = [(ni + 1) * nj * nk, ni * (nj + 1) * nk, ni * nj * (nk + 1)]
    faces = sum(nf)
    nodes = (ni + 1) * (nj + 1) * (nk + 1)
    nodal = np.empty(nodes, np.float32)
    L.elq_scene_room_nodal(ni, nj, nk, nodal.ctypes.data)
    p_nodal = ctx.alloc(nodal.nbytes)
    ctx.h2d(p_nodal, nodal)
    p_w = [ctx.alloc(4 * n) for n in nf]
    p_flags = [ctx.alloc(n) for n in nf]
    p_save = [ctx.alloc(4 * n) for n in nf]
    p_fsave = [ctx.alloc(n) for n in nf]
    p_ls = ctx.alloc(cells)
    p_phi = ctx.alloc(4 * cells)
    g = ds.grid()
    out = {}

    def timed(name, call, restore=None, classes=None):
        ms = []
        for _ in range(reps):
            if restore:
                restore()
            ctx.set_profiling(True)
            if call() != elq.OK:
                raise SystemExit(f"{name} failed: {ctx.last_error()}")
            kt = ctx.kernel_times()
            ctx.set_profiling(False)
            ms.append(sum(v[0] for k, v in kt.items() if classes is None or k in classes))
        return float(np.median(ms)), kt

    def entry(name, ms, nbytes, kt=None):
        e = {"ms": round(ms, 4), "algorithmic_bytes": int(nbytes), "gbs": round(nbytes / (ms / 1e3) / 1e9, 1),
             "hbm_frac": round(nbytes / (ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4)}
        if kt and len(kt) > 1:
            e["kernels_ms"] = {k: round(v[0], 4) for k, v in kt.items()}
        out[name] = e

    ms, _ = timed("weights", lambda: L.elq_compute_weights_device(h, ni, nj, nk, p_nodal, *p_w))
    entry("compute_weights", ms, 3 * 4 * nodes + 4 * faces)
    v_save = ctx.alloc(4 * nf[1])
    ctx.d2d(v_save, ds.ptr["v"], 4 * nf[1])
    ms, _ = timed("gravity", lambda: L.elq_add_gravity_device(h, ni, nj, nk, ds.ptr["v"], np.float32(s.dt)),
                  restore=lambda: ctx.d2d(ds.ptr["v"], v_save, 4 * nf[1]))
    ctx.d2d(ds.ptr["v"], v_save, 4 * nf[1])
    ctx.free(v_save)
    entry("add_gravity", ms, 8 * nf[1])
    ms, _ = timed("find_valid", lambda: L.elq_find_valid_velocity_device(h, g, *p_flags))
    entry("find_valid_velocity", ms, 4 * faces + 3 * 4 * cells + faces)
    flags = [np.empty(n, np.uint8) for n in nf]
    for f, p in zip(flags, p_flags):
        ctx.d2h(f, p)
    n_invalid = sum(int(f.size - np.count_nonzero(f)) for f in flags)
    ms, _ = timed("zero", lambda: L.elq_set_invalid_velocity_zero_device(h, g, *p_flags))
    entry("set_invalid_velocity_zero", ms, faces + 4 * n_invalid)
    # the velocity after the projection, as the next frame's extrapolation sees it
    for k, p, n in zip("uvw", p_save, nf):
        ctx.d2d(p, ds.ptr[k], 4 * n)
    for p, q, n in zip(p_fsave, p_flags, nf):
        ctx.d2d(p, q, n)

    def restore_vel():
        for k, p, n in zip("uvw", p_save, nf):
            ctx.d2d(ds.ptr[k], p, 4 * n)
        for q, p, n in zip(p_flags, p_fsave, nf):
            ctx.d2d(q, p, n)

    ms, kt = timed("extrapolate", lambda: L.elq_extrapolate_velocity_device(h, g, *p_flags, 10), restore_vel)
    after = [np.empty(n, np.uint8) for n in nf]
    for f, p in zip(after, p_flags):
        ctx.d2h(f, p)
    n_new = sum(int(np.count_nonzero(a)) - int(np.count_nonzero(b)) for a, b in zip(after, flags))
    entry("extrapolate_velocity", ms, faces + 9 * n_new, kt)
    out["extrapolate_velocity"]["faces"] = faces
    out["extrapolate_velocity"]["extrapolated_faces"] = n_new
    restore_vel()
    ms, _ = timed("ls_valid", lambda: L.elq_find_valid_level_set_device(h, ni, nj, nk, p_nodal, p_ls))
    entry("find_valid_level_set", ms, 4 * nodes + cells)
    ctx.d2d(p_phi, ds.ptr["phi"], 4 * cells)
    ms, kt = timed("extrapolate_ls", lambda: L.elq_extrapolate_const_device(h, ni, nj, nk, p_phi, p_ls, 10),
                   restore=lambda: ctx.d2d(p_phi, ds.ptr["phi"], 4 * cells))
    ls = np.empty(cells, np.uint8)
    ctx.d2h(ls, p_ls)
    entry("extrapolate_level_set", ms, cells + 9 * int(cells - np.count_nonzero(ls)), kt)
    # advection of the level set, then of the velocity (apps/eulerian_liquid_3d.cpp:111-112),
    # kernel time only (the copy of the advected grid taken before the step is not counted)
    restore_vel()
    ms, _ = timed("advect_ls", lambda: L.elq_advect_level_set_device(h, ni, nj, nk, np.float32(s.dx),
                                                                     np.float32(s.dt), p_phi, ds.ptr["u"],
                                                                     ds.ptr["v"], ds.ptr["w"]),
                  restore=lambda: ctx.d2d(p_phi, ds.ptr["phi"], 4 * cells), classes=("advect_cells",))
    entry("advect_level_set", ms, 8 * cells + 4 * faces)
    ms, _ = timed("advect_vel", lambda: L.elq_advect_velocity_device(h, g, np.float32(s.dt)), restore=restore_vel,
                  classes=("advect_faces",))
    entry("advect_velocity", ms, 8 * faces)
    restore_vel()
    # reinitialisation of the level set with the room's flags (simulator.cpp:181-183)
    import ctypes as C
    left, passes = C.c_int64(0), C.c_int32(0)
    wall = []

    def rei():
        t0 = time.perf_counter()
        rc = L.elq_reinitialize_device(h, ni, nj, nk, np.float32(s.dx), p_phi, p_ls, C.byref(left), C.byref(passes))
        wall.append((time.perf_counter() - t0) * 1e3)
        return rc

    ms, kt = timed("reinit", rei, restore=lambda: ctx.d2d(p_phi, ds.ptr["phi"], 4 * cells))
    entry("reinitialize", ms, 9 * cells, kt)
    out["reinitialize"].update({"wall_ms": round(float(np.median(wall)), 4), "passes": passes.value,
                                "unreached": left.value})
    for p in [p_nodal, p_ls, p_phi] + p_w + p_flags + p_save + p_fsave:
        ctx.free(p)
    return out


def shim_frame(n=256, frames=4):
    """One frame of the reference's loop through the C++ Simulator shim
    (elq_sample3d: extrapolate_velocity, extrapolate_level_set, advect_level_set,
    advect_velocity, reinitialize, add_gravity, find_valid_velocity,
    set_invalid_velocity_zero, solve_pressure) on the n^3 drop scene (slab
    y < 0.4 plus a sphere of radius 0.15), mixed precision to rel.res 1e-8,
    grids resident in device memory for the whole run. frame_ms: host wall
    clock of a frame (median of frames 1.., frame 0 warms up); device_sum_ms:
    a second run with HIP-event timing of every step's kernels, the sum over
    the frame (median of frames 1..), which is what a device-resident frame
    costs without any host work."""
    import subprocess
    exe = os.path.join(ROOT, "eulerian-liquid_amd", "elq_sample3d")

    def run(extra):
        r = subprocess.run([exe, str(n), str(frames), "1e-8", "1"] + extra, capture_output=True, text=True,
                           timeout=300)
        if r.returncode != 0:
            raise SystemExit(f"elq_sample3d failed: {r.stderr.strip()}")
        return r.stdout.splitlines()

    def steps(line):   # "<label>: name=ms name=ms ..."
        return {k: float(v) for k, v in (kv.split("=") for kv in line.split(":", 1)[1].split())}

    out = run([])
    fl = [ln for ln in out if ln.startswith("frame")]
    wall = [ln for ln in out if ln.strip().startswith("wall_ms:")]
    frame_ms = [float(ln.split("frame_ms=")[1]) for ln in fl]
    iters = [int(ln.split("iterations=")[1].split()[0]) for ln in fl]
    walls = [steps(ln) for ln in wall]
    prof = run(["--profile"])
    dev = [ln.strip() for ln in prof if ln.strip().startswith("device_ms")]
    dev_sum = [float(ln.split("(sum ")[1].split(")")[0]) for ln in dev]
    dev_steps = [steps(ln.split(")", 1)[1]) for ln in dev]
    med = lambda xs: float(np.median(xs[1:] if len(xs) > 1 else xs))
    names = list(walls[0].keys())
    e = {"workload": f"drop{n} (slab y<0.4 + sphere r=0.15), mixed, rel.res 1e-8",
         "frames": len(fl), "frame_ms": round(med(frame_ms), 3),
         "iterations": iters,
         "wall_ms": {k: round(med([w[k] for w in walls]), 3) for k in names},
         "device_sum_ms": round(med(dev_sum), 3),
         "device_ms": {k: round(med([d[k] for d in dev_steps]), 3) for k in names}}
    e["steps_outside_projection_ms"] = round(e["frame_ms"] - e["wall_ms"]["solve_pressure"], 3)
    e["frame_over_device_sum"] = round(e["frame_ms"] / e["device_sum_ms"], 4)
    return e


def whole_solve_frac(cells, n, iters, outer, ms, mixed):
    """Algorithmic bytes of a whole projection / its time, as a fraction of
    the HBM peak: iterations x unknowns x B_iter, plus the refinement steps
    (mixed) and the set-up (indices, assembly, factor, scatter, update)."""
    sm = sv = 4 if mixed else 8
    total = iters * n * B_ITER(sm, sv) + (outer * n * B_OUTER if mixed else 0) + 92 * cells + n * (9 * sm + 16)
    return round(total / (ms / 1e3) / 1e9 / PEAK_HBM_GBS, 4)


def gpu_config(ctx, ds, opts, reps):
    """One warm-up projection, then reps timed ones (median); the outputs of
    the last one."""
    project(ctx, ds, opts)
    ts, st = [], None
    for _ in range(reps):
        dt_ms, st = project(ctx, ds, opts)
        ts.append(dt_ms)
    ms = float(np.median(ts))
    mixed = opts.precision == elq.MIXED
    e = {"ms_per_solve": round(ms, 2), "iterations": st.iterations, "outer_iterations": st.outer_iterations,
         "true_rel_residual": st.true_rel_residual, "max_fluid_divergence_rel": fluid_divergence(ds),
         "hbm_frac_whole_solve": whole_solve_frac(ds.s.cells, st.n_unknowns, st.iterations, st.outer_iterations,
                                                  ms, mixed)}
    return e, ds.fetch("pressure"), st


def configs(base, seed, threads, cfg5_oracle=False):
    """The other configurations of the workload, on the GPU (both
    precisions), each with its error: cfg1 (64^3 drop) and cfg2 (128^3 drop)
    against the oracle's FP64 solve to rel.res 1e-12, cfg1 also at the
    reference's solver defaults (tol 1e-4, 1000 iterations) against the
    oracle's iteration count and error(); cfg5 (512^3, 128 spheres, 48.7 M
    unknowns) against the FP64 GPU drop-in (the oracle would need about five
    minutes of CPU there; its parity is shown at the smaller sizes)."""
    import oracle as O
    O.set_threads(threads, False)
    out = {}
    ctx = elq.Context(0)
    for name, sc in (("cfg1_drop64", elq.scene_drop(64)), ("cfg2_drop128", elq.scene_drop(128))):
        ref = O.solve_pressure(oracle_scene(sc), tol=1e-12, max_iter=20000)
        scale = float(np.abs(ref.pressure).max())
        ds = elq.DeviceScene(ctx, sc)
        e = {"workload": f"{name[5:]}: slab y<0.4 + sphere r=0.15, one gravity step, rel.res 1e-8",
             "unknowns": ref.n, "oracle": f"FP64 IC(0)-PCG to rel.res 1e-12 ({ref.iterations} iterations)"}
        ps = {}
        for key, prec in (("fp64", elq.FP64), ("mixed", elq.MIXED)):
            e[key], ps[key], _ = gpu_config(ctx, ds, elq.default_opts(precision=prec, **base), 3)
            e[key]["max_rel_p_minus_p_oracle"] = float(np.abs(ps[key] - ref.pressure).max()) / scale
        e["speedup_mixed_vs_fp64"] = round(e["fp64"]["ms_per_solve"] / e["mixed"]["ms_per_solve"], 3)
        e["max_rel_p_mixed_minus_p_fp64"] = float(np.abs(ps["mixed"] - ps["fp64"]).max()) / scale
        if name.startswith("cfg1"):
            _, st = project(ctx, ds, elq.default_opts())
            ref4 = O.solve_pressure(oracle_scene(sc), tol=1e-4, max_iter=1000)
            e["reference_defaults_tol1e-4"] = {"gpu_iterations": st.iterations, "gpu_error": st.error,
                                               "oracle_iterations": ref4.iterations, "oracle_error": ref4.error}
        ds.close()
        out[name] = e
        print(f"# {name}: fp64 {e['fp64']['ms_per_solve']} ms, mixed {e['mixed']['ms_per_solve']} ms, "
              f"err {e['fp64']['max_rel_p_minus_p_oracle']:.2e}/{e['mixed']['max_rel_p_minus_p_oracle']:.2e}",
              file=sys.stderr, flush=True)
    ctx.close()
    t0 = time.perf_counter()
    sc = elq.scene_spheres(512, seed=seed, n_spheres=128)
    print(f"# cfg5: 512^3 scene generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    ctx = elq.Context(0)
    ds = elq.DeviceScene(ctx, sc)
    e = {"workload": f"spheres512 seed {seed}: slab y<0.3 + 128 spheres, velocities U[-0.5,0.5], rel.res 1e-8"}
    ps = {}
    for key, prec in (("mixed", elq.MIXED), ("fp64", elq.FP64)):
        e[key], ps[key], st = gpu_config(ctx, ds, elq.default_opts(precision=prec, **base), 1)
        e["unknowns"] = st.n_unknowns
        print(f"# cfg5 {key}: {e[key]['ms_per_solve']} ms, {e[key]['iterations']} iterations", file=sys.stderr,
              flush=True)
    scale = float(np.abs(ps["fp64"]).max())
    e["speedup_mixed_vs_fp64"] = round(e["fp64"]["ms_per_solve"] / e["mixed"]["ms_per_solve"], 3)
    e["max_rel_p_mixed_minus_p_fp64"] = float(np.abs(ps["mixed"] - ps["fp64"]).max()) / scale
    e["error_reference"] = "FP64 GPU drop-in (oracle parity shown at cfg1/cfg2 and the golden fixtures)"
    ds.close()
    ctx.close()
    if cfg5_oracle:
        # the checker at this size: the oracle's FP64 solve to 1e-10 with
        # OpenMP over SpMV, axpys and dots (the IC sweeps stay serial);
        # minutes of host time, so off by default
        import oracle as O
        O.set_threads(threads, True)
        O.set_progress(50)
        t0 = time.perf_counter()
        ref = O.solve_pressure(oracle_scene(sc), tol=1e-10, max_iter=20000)
        O.set_progress(0)
        O.set_threads(threads, False)
        sref = float(np.abs(ref.pressure).max())
        e["oracle"] = {"solve": "FP64 IC(0)-PCG to rel.res 1e-10", "iterations": ref.iterations,
                       "host_s": round(time.perf_counter() - t0, 1), "same_unknowns": ref.n == e["unknowns"]}
        for key in ("mixed", "fp64"):
            e[key]["max_rel_p_minus_p_oracle"] = float(np.abs(ps[key] - ref.pressure).max()) / sref
        e["error_reference"] = "the oracle (and the FP64 GPU drop-in)"
        print(f"# cfg5 vs oracle: {e['oracle']}, mixed {e['mixed']['max_rel_p_minus_p_oracle']:.2e}, "
              f"fp64 {e['fp64']['max_rel_p_minus_p_oracle']:.2e}", file=sys.stderr, flush=True)
    out["cfg5_spheres512"] = e
    return out


def host_info():
    """The host the CPU baseline runs on: CPU model, logical CPUs, the CPUs
    this process may use, and the OpenMP threads it is given."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": affinity, "omp_threads": threads}


def oracle_scene(scene):
    import oracle as O
    return O.Scene(scene.ni, scene.nj, scene.nk, scene.dx, scene.dt, scene.phi, scene.u.copy(), scene.v.copy(),
                   scene.w.copy(), scene.uw, scene.vw, scene.ww)


def time_oracle(scene, reps, threads, krylov_parallel=False, tol=1e-8):
    """Full oracle projections (indices, assembly, IC(0) factor, PCG to tol,
    scatter, update) of scene: one warm-up, then the median of reps."""
    import oracle as O
    os_ = oracle_scene(scene)
    O.set_threads(threads, krylov_parallel)
    ts, res = [], None
    for q in range(reps + 1):
        t0 = time.perf_counter()
        res = O.solve_pressure(os_, tol=tol, max_iter=20000)
        if q:
            ts.append(time.perf_counter() - t0)
    O.set_threads(threads, False)
    return float(np.median(ts)), res


def cpu_baseline(scene, threads):
    """The reference path on the host's cores: the oracle (the FP64 CPU
    restatement of projection.cpp + Eigen's CG loop, natural-order IC(0);
    built -O3 -DNDEBUG -fopenmp like the reference's Release build) timed on
    full projections to rel.res 1e-8, with the reference's threading: serial
    assembly, factorisation and PCG, OpenMP over the pressure scatter and the
    velocity update (projection.cpp:85-134).
    - value: one full projection of the benchmark workload (256^3), OpenMP
      threads = the CPUs the process may use;
    - cfg1 (64^3 drop) and cfg2 (128^3 drop): median of 5 after a warm-up,
      with 1 thread and with all threads;
    - companion (labelled, not the reference's threading): cfg2 with OpenMP
      SpMV, axpys and dots as well (the IC sweeps stay serial)."""
    info = host_info()
    out = {"host": info}
    cfgs = {}
    for name, sc, reps in (("cfg1_drop64", elq.scene_drop(64), 5), ("cfg2_drop128", elq.scene_drop(128), 5)):
        e = {}
        for th in sorted({1, threads}):
            t, res = time_oracle(sc, reps, th)
            e[f"threads_{th}"] = {"s_per_solve": round(t, 4), "solves_per_s": round(1.0 / t, 3)}
            e["iterations"] = res.iterations
            print(f"# cpu {name} threads {th}: {t:.3f} s ({res.iterations} iterations)", file=sys.stderr, flush=True)
        cfgs[name] = e
    t, res = time_oracle(elq.scene_drop(128), 3, threads, krylov_parallel=True)
    cfgs["cfg2_drop128"]["companion_openmp_krylov"] = {"threads": threads, "s_per_solve": round(t, 4),
                                                       "note": "OpenMP SpMV/axpy/dot added; not the reference's "
                                                               "threading"}
    import oracle as O
    O.set_threads(threads, False)
    os_ = oracle_scene(scene)
    t0 = time.perf_counter()
    res = O.solve_pressure(os_, tol=1e-8, max_iter=20000)
    t = time.perf_counter() - t0
    print(f"# cpu workload: {t:.1f} s ({res.iterations} iterations)", file=sys.stderr, flush=True)
    out["_pressure"] = res.pressure
    out.update({"value": round(1.0 / t, 5), "unit": "solves/s", "cores": threads, "kind": "port",
                "sample": f"one full oracle projection of the benchmark workload to rel.res 1e-8 "
                          f"({res.iterations} IC(0)-PCG iterations, {t:.1f} s) with {threads} OpenMP threads on "
                          f"{info['cpu_model']}; serial assembly/factor/PCG as in the reference, OpenMP scatter "
                          f"and update",
                "s_per_solve": round(t, 3), "iterations": res.iterations, "true_rel_residual": res.true_rel_res,
                "configs": cfgs})
    return out


def cpu_baseline_frame(ctx):
    """The oracle's frame steps (single thread) on a bounded sample, the
    128^3 version of the workload's scene (about 4 s of CPU work), beside the
    GPU's times for the same sample: extrapolate_velocity (10 layers),
    advect_level_set, advect_velocity and reinitialize."""
    import oracle as O
    s = elq.scene_spheres(128, seed=1, n_spheres=32)
    n = s.ni
    dims = ((n + 1, n, n), (n, n + 1, n), (n, n, n + 1))
    cpu = {}
    flags = O.find_valid_velocity(s)
    vel = [a.copy() for a in (s.u, s.v, s.w)]
    t0 = time.perf_counter()
    for a, f, d in zip(vel, flags, dims):
        O.extrapolate(a, f.copy(), d, 10)
    cpu["extrapolate_velocity"] = time.perf_counter() - t0
    phi = np.empty_like(s.phi)
    t0 = time.perf_counter()
    O.lib().orc_advect_level_set(n, n, n, np.float32(s.dx), np.float32(s.dt), s.phi, s.u, s.v, s.w, phi)
    cpu["advect_level_set"] = time.perf_counter() - t0
    out = [np.empty_like(a) for a in (s.u, s.v, s.w)]
    t0 = time.perf_counter()
    O.lib().orc_advect_velocity(n, n, n, np.float32(s.dx), np.float32(s.dt), s.u, s.v, s.w, *out)
    cpu["advect_velocity"] = time.perf_counter() - t0
    nodal = np.empty((n + 1) ** 3, np.float32)
    O.lib().orc_room_nodal(n, n, n, nodal)
    ls = np.empty(n ** 3, np.uint8)
    O.lib().orc_find_valid_level_set(n, n, n, nodal, ls)
    phi = s.phi.copy()
    t0 = time.perf_counter()
    O.reinitialize(phi, ls, (n, n, n), s.dx)
    cpu["reinitialize"] = time.perf_counter() - t0
    ds = elq.DeviceScene(ctx, s)
    gpu = frame_steps(ctx, ds, reps=3)
    ds.close()
    return {"sample": "spheres128 seed 1 (the workload's generator at 128^3), oracle single thread vs the GPU "
                      "on the same sample",
            "cores": 1, "kind": "port",
            "steps": {k: {"cpu_s": round(v, 3), "gpu_ms": gpu[k]["ms"],
                          "speedup": round(v * 1e3 / gpu[k]["ms"], 1)} for k, v in cpu.items()}}


def copy_calibration(nbytes, reps):
    """Copies nbytes device-to-device with the library's streaming kernel
    (reads nbytes, writes nbytes) so that per-dispatch FETCH_SIZE/WRITE_SIZE
    counters can be scaled against a kernel of known traffic."""
    ctx = elq.Context(0)
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    ctx.h2d(src, np.ones(nbytes // 4, np.float32))
    ts = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        ctx.d2d(dst, src, nbytes)
        ts.append(time.perf_counter() - t0)
    ctx.free(src)
    ctx.free(dst)
    ctx.close()
    t = min(ts[1:])
    print(json.dumps({"copy_bytes": nbytes, "dispatches": reps + 1, "best_s": round(t, 6),
                      "copy_gbs_incl_sync": round(2 * nbytes / t / 1e9, 1)}), flush=True)


def read_calibration(nbytes, reps):
    """Sums nbytes of float32 with torch's reduction kernel (reads nbytes,
    writes a few bytes): the read-only companion of copy_calibration for
    scaling per-dispatch read counters."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        v = float(x.sum())
        ts.append(time.perf_counter() - t0)
    assert v == nbytes // 4
    t = min(ts[1:])
    print(json.dumps({"read_bytes": nbytes, "dispatches": reps + 1, "best_s": round(t, 6),
                      "read_gbs_incl_sync": round(nbytes / t / 1e9, 1)}), flush=True)


def serve_run(scenes, arrivals, slots, opts, keep_pressure=False, batched=0, stats=None):
    """Submits fresh copies of the scenes at the given arrival times (s from
    the start) to a server with `slots` slots (batched > 0: Krylov chunks of the
    running requests advance in shared launches); returns per-request results
    (with keep_pressure, each keeps its pressure and the index of its scene).
    stats (a dict) receives the server's batch counts."""
    work = [sc.copy() for sc in scenes]
    n = len(work)
    got = []
    index = {}
    with elq.Server(0, max_slots=slots, max_queue=n, opts=opts, batched=batched) as srv:
        t0 = time.perf_counter()
        i = 0
        while len(got) < n:
            now = time.perf_counter() - t0
            while i < n and arrivals[i] <= now:
                index[srv.submit(work[i])] = i
                i += 1
            nxt = (arrivals[i] - now) if i < n else 1.0
            if srv.in_flight() == 0 and i < n:
                time.sleep(max(nxt, 0.0))
            else:
                srv.wait(int(max(nxt, 0.0) * 1e3) + 1)
            got += srv.poll()
        if stats is not None:
            stats["batches"], stats["batched_requests"] = srv.batch_stats()
    for r in got:
        r.pop("scene")
        r["index"] = index[r["id"]]
        if not keep_pressure:
            r.pop("pressure")
    return got


def serving_oracle_errors(scenes, got, threads):
    """Every served request against the oracle's FP64 solve of its scene to
    rel.res 1e-12 (the checker; solves in parallel host threads, one OpenMP
    thread each): max|p - p_oracle| / max|p_oracle| per request."""
    import concurrent.futures as cf
    import oracle as O
    O.set_threads(1, False)

    def err(r):
        ref = O.solve_pressure(oracle_scene(scenes[r["index"]]), tol=1e-12, max_iter=20000)
        scale = float(np.abs(ref.pressure).max())
        return float(np.abs(r["pressure"] - ref.pressure).max()) / scale if scale else 0.0, ref.n == r["n_unknowns"]

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(err, got))
    O.set_threads(threads, False)
    e = np.array([x[0] for x in res])
    return {"requests": len(res), "max_rel_p_minus_p_oracle": float(e.max()),
            "median_rel_p_minus_p_oracle": float(np.median(e)), "same_unknowns": all(x[1] for x in res),
            "oracle": "FP64 IC(0)-PCG to rel.res 1e-12", "host_s": round(time.perf_counter() - t0, 1)}


def serve_summary(got):
    lat = np.array([r["t_done_ms"] - r["t_submit_ms"] for r in got])
    span_s = (max(r["t_done_ms"] for r in got) - min(r["t_submit_ms"] for r in got)) / 1e3
    return {"solves_per_s": round(len(got) / span_s, 2), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "max_ms": round(float(lat.max()), 2),
            "all_ok": all(r["status"] == elq.OK for r in got),
            "max_true_rel_residual": max(r["true_rel_residual"] for r in got),
            "mean_iterations": round(float(np.mean([r["iterations"] for r in got])), 1)}


def serving_batched(scenes, slots, o_mixed, groups, seed):
    """The same stream on a batched server (elq_server_opts.batched = groups):
    the Krylov chunks of the running requests advance in lockstep, one launch
    per kernel class for a whole group of slots, replayed from one captured
    HIP graph. Saturation and Poisson arrivals at 0.8x its own saturation
    rate; batches launched and requests per batch."""
    n = len(scenes)
    serve_run(scenes[:min(8, n)], [0.0] * min(8, n), slots, o_mixed, batched=groups)  # warm-up, graph capture
    st = {}
    sat = serve_summary(serve_run(scenes, [0.0] * n, slots, o_mixed, batched=groups, stats=st))
    lam = 0.8 * sat["solves_per_s"]
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / lam, n)).tolist()
    poisson = serve_summary(serve_run(scenes, arrivals, slots, o_mixed, batched=groups))
    out = {"groups": groups, "slots": slots, "saturation": sat, "poisson_rate_per_s": round(lam, 2),
           "poisson": poisson, "batches": st["batches"],
           "requests_per_batch": round(st["batched_requests"] / max(st["batches"], 1), 2)}
    print(f"# serving, batched lockstep: {out}", file=sys.stderr, flush=True)
    return out


def serving(n_req, slots, seed, base, fp64=True, check_threads=0, batch_groups=0):
    """Serving stream: n_req independent requests, dims per axis from
    {32,48,64,96,128}, 1-8 spheres (radii 0.05-0.2), slab, face velocities
    U[-1,1] (request seeds seed*100000 + i). Runs: saturation (all submitted
    at t=0) with `slots` slots; Poisson arrivals at 0.8x that saturation rate;
    the plain FP64 drop-in one request at a time (1 slot) at saturation."""
    t = time.perf_counter()
    scenes = [elq.scene_request(seed * 100000 + i) for i in range(n_req)]
    cells = int(sum(sc.cells for sc in scenes))
    print(f"# serving: {n_req} requests generated in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    o_mixed = elq.default_opts(precision=elq.MIXED, **base)
    o_fp64 = elq.default_opts(precision=elq.FP64, **base)
    serve_run(scenes[:min(8, n_req)], [0.0] * min(8, n_req), slots, o_mixed)  # warm-up
    sat = serve_summary(serve_run(scenes, [0.0] * n_req, slots, o_mixed))
    lam = 0.8 * sat["solves_per_s"]
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req)).tolist()
    got = serve_run(scenes, arrivals, slots, o_mixed, keep_pressure=check_threads > 0)
    poisson = serve_summary(got)
    print(f"# serving: saturation {sat}, poisson {poisson}", file=sys.stderr, flush=True)
    out = {"workload": f"{n_req} requests, dims per axis from (32,48,64,96,128), {cells} cells in all, "
                       f"mixed precision to rel.res 1e-8, {slots} slots",
           "poisson_rate_per_s": round(lam, 2), "poisson": poisson, "saturation": sat}
    if check_threads > 0:
        out["poisson_vs_oracle"] = serving_oracle_errors(scenes, got, check_threads)
        print(f"# serving vs oracle: {out['poisson_vs_oracle']}", file=sys.stderr, flush=True)
    del got
    if batch_groups > 0:
        out["batched_lockstep"] = serving_batched(scenes, slots, o_mixed, batch_groups, seed)
    if not fp64:
        return out
    seq64 = serve_summary(serve_run(scenes, [0.0] * n_req, 1, o_fp64))
    print(f"# serving: fp64 one slot {seq64}", file=sys.stderr, flush=True)
    return {**out, "fp64_drop_in_one_slot_saturation": seq64,
            "speedup_vs_fp64_one_slot": round(sat["solves_per_s"] / seq64["solves_per_s"], 2)}


def serving_child(a):
    """The same serving stream with more slots than hardware queues allow
    by default: a child process with GPU_MAX_HW_QUEUES=16 (the runtime reads
    it at start-up, so it cannot change inside this process) and
    --throughput-slots slots. More slots raise the saturated rate; the
    Poisson run at 0.8x that rate then has longer queues."""
    import subprocess
    env = dict(os.environ, GPU_MAX_HW_QUEUES="16")
    cmd = [sys.executable, os.path.abspath(__file__), "--serve-only", "--serve", str(a.serve), "--slots",
           str(a.throughput_slots), "--seed", str(a.seed), "--precond", a.precond, "--schedule", a.schedule]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        raise SystemExit(f"serving child failed: {r.stderr.strip()[-500:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])["serving"]
    out["hw_queues"] = 16
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="spheres256", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precond", default="ic0", choices=["ic0", "mic0"])
    ap.add_argument("--schedule", default="wavefront", choices=["wavefront", "levels"])
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true", help="time only the FP64 drop-in (profiling runs)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the MIC(0), frame-step and host-buffer measurements")
    ap.add_argument("--frame-only", action="store_true",
                    help="one untimed projection, then only the frame steps (profiling runs)")
    ap.add_argument("--serve", type=int, default=512, metavar="N",
                    help="serving stream of N requests after the single-solve benchmark (0: skip)")
    ap.add_argument("--slots", type=int, default=4, help="serving slots")
    ap.add_argument("--batch-groups", type=int, default=2, metavar="G",
                    help="also serve the stream batched in G lockstep groups of slots (0: skip)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving stream (mixed precision), as one JSON object (child runs)")
    ap.add_argument("--throughput-slots", type=int, default=12, metavar="N",
                    help="slots of the throughput-oriented serving run in a child process with "
                         "GPU_MAX_HW_QUEUES=16 (0: skip)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (64^3, 128^3 against the oracle; 512^3)")
    ap.add_argument("--cfg5-oracle", action="store_true",
                    help="also check the 512^3 configuration against the oracle (minutes of host time)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the library's plain copy kernel on MIB MiB (HBM counter calibration)")
    ap.add_argument("--read-calibration", type=int, default=0, metavar="MIB",
                    help="only sum MIB MiB of float32 with torch (read-only HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "the projection runs on one GPU here; --gpus must be 1"}))
        raise SystemExit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration << 20, max(a.steps, 1))
        return
    if a.read_calibration:
        read_calibration(a.read_calibration << 20, max(a.steps, 1))
        return
    if a.serve_only:
        base = dict(tol=1e-8, max_iter=20000, precond=elq.IC0 if a.precond == "ic0" else elq.MIC0,
                    schedule=elq.WAVEFRONT if a.schedule == "wavefront" else elq.LEVELS, check_every=32)
        print(json.dumps({"serving": serving(a.serve, a.slots, a.seed, base, fp64=False)}), flush=True)
        return

    t_gen = time.perf_counter()
    scene, wname = workload(a.workload, a.seed)
    t_gen = time.perf_counter() - t_gen
    print(f"# workload {wname}: {scene.ni}x{scene.nj}x{scene.nk}, generated in {t_gen:.1f}s", file=sys.stderr,
          flush=True)
    precond = elq.IC0 if a.precond == "ic0" else elq.MIC0
    schedule = elq.WAVEFRONT if a.schedule == "wavefront" else elq.LEVELS
    base = dict(tol=1e-8, max_iter=20000, precond=precond, schedule=schedule, check_every=32)
    o_mixed = elq.default_opts(precision=elq.MIXED, **base)
    o_fp64 = elq.default_opts(precision=elq.FP64, **base)

    ctx = elq.Context(0)
    ds = elq.DeviceScene(ctx, scene)
    if a.frame_only:
        project(ctx, ds, o_mixed)
        print(json.dumps({"workload": wname, "frame_steps": frame_steps(ctx, ds)}), flush=True)
        ds.close()
        ctx.close()
        return
    out = {}
    if not a.fp64_only:
        res = run_steps(ctx, ds, o_mixed, a.steps, a.warmup)
        p_mixed = ds.fetch("pressure")
        div_mixed = fluid_divergence(ds)
        print(f"# mixed: {[round(t, 1) for t in res[0]]} ms, iters {res[1][-1]['iterations']} "
              f"outer {res[1][-1]['outer_iterations']}", file=sys.stderr, flush=True)
        out["mixed"] = res
    if not a.no_fp64 or a.fp64_only:
        steps64 = a.steps if a.fp64_only else max(1, min(a.steps, 2))
        res = run_steps(ctx, ds, o_fp64, steps64, a.warmup if a.fp64_only else 1)
        p_fp64 = ds.fetch("pressure")
        div_fp64 = fluid_divergence(ds)
        print(f"# fp64: {[round(t, 1) for t in res[0]]} ms, iters {res[1][-1]['iterations']}", file=sys.stderr,
              flush=True)
        out["fp64"] = res

    key = "fp64" if a.fp64_only else "mixed"
    times, stats, kt, dom = out[key]
    st = stats[-1]
    n = st["n_unknowns"]
    sm = sv = 8 if key == "fp64" else 4
    ms_step = float(np.mean(times))
    iters, outer = st["iterations"], st["outer_iterations"]
    setup_bytes = 92 * scene.cells + n * (9 * sm + 16)
    total_bytes = iters * n * B_ITER(sm, sv) + (outer * n * B_OUTER if key == "mixed" else 0) + setup_bytes
    roof = roofline(dom, n, sm, sv)
    roof["traffic"], src = profiled_traffic(wname, key, roof["kernel"])
    if src:
        roof["traffic_source"] = src
    line = {
        "metric": "PCG pressure solves/s & HBM GB/s at 256^3 (rel.res 1e-8); p99 batched latency",
        "value": round(1000.0 / ms_step, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if key == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{wname} seed {a.seed} ({scene.ni}^3 slab+spheres, rel.res 1e-8)",
                   "precision": key, "precond": a.precond, "schedule": a.schedule, "unknowns": n,
                   "bricks": st["n_bricks"], "iterations": iters, "outer_iterations": outer,
                   "true_rel_residual": st["true_rel_residual"],
                   "max_fluid_divergence_rel": div_fp64 if key == "fp64" else div_mixed},
        "hbm_gbs_whole_solve": round(total_bytes / (ms_step / 1e3) / 1e9, 1),
        "hbm_frac_whole_solve": round(total_bytes / (ms_step / 1e3) / 1e9 / PEAK_HBM_GBS, 4),
        "roofline": roof,
        "kernel_ms_per_step": {k: round(v[0], 3) for k, v in sorted(kt.items())},
        "pcg_ms_per_solve": round(float(np.median([x["t_solve_ms"] for x in stats])), 2),
        "kernel_classes": class_rates(kt, n, sm, sv),
    }
    if "fp64" in out and key == "mixed":
        t64, s64 = out["fp64"][0], out["fp64"][1]
        ms64 = float(np.mean(t64))
        scale = float(np.abs(p_fp64).max())
        line["fp64_drop_in"] = {"solves_per_s": round(1000.0 / ms64, 4), "ms_per_solve": round(ms64, 2),
                                "pcg_ms_per_solve": round(float(np.median([x["t_solve_ms"] for x in s64])), 2),
                                "iterations": s64[-1]["iterations"],
                                "true_rel_residual": s64[-1]["true_rel_residual"],
                                "max_fluid_divergence_rel": div_fp64,
                                "kernel_classes": class_rates(out["fp64"][2], n, 8, 8)}
        line["speedup_mixed_vs_fp64"] = round(ms64 / ms_step, 3)
        line["max_abs_p_minus_p_fp64"] = float(np.abs(p_mixed - p_fp64).max())
        line["max_rel_p_minus_p_fp64"] = float(np.abs(p_mixed - p_fp64).max()) / scale if scale else 0.0
    if key == "mixed" and a.precond == "ic0" and not a.no_extra:
        # the same inputs with the modified IC(0) preconditioner (tau from the
        # default options): same accuracy target, fewer iterations
        o_mic = elq.default_opts(precision=elq.MIXED, **dict(base, precond=elq.MIC0))
        t_mic, st_mic = [], None
        for _ in range(3):
            dt_ms, st_mic = project(ctx, ds, o_mic)
            t_mic.append(dt_ms)
        line["mic0"] = {"tau": o_mic.tau, "ms_per_solve": round(float(np.mean(t_mic[1:])), 2),
                        "solves_per_s": round(1e3 / float(np.mean(t_mic[1:])), 3),
                        "iterations": st_mic.iterations, "outer_iterations": st_mic.outer_iterations,
                        "true_rel_residual": st_mic.true_rel_residual}
    if not a.no_extra or a.frame_only:
        line["frame_steps"] = frame_steps(ctx, ds)
    ds.close()
    if key == "mixed" and not a.no_extra:
        # the same projection through elq_project on host buffers (inputs copied
        # in, u/v/w and pressure copied back): the PCIe-inclusive rate
        t_host = []
        for _ in range(2):
            sc = scene.copy()
            t0 = time.perf_counter()
            ctx.solve_pressure(sc, o_mixed)
            t_host.append((time.perf_counter() - t0) * 1e3)
        line["host_buffers"] = {"ms_per_solve": round(min(t_host), 2), "solves_per_s": round(1e3 / min(t_host), 3),
                                "note": "elq_project on pageable host arrays, copies included; not the headline"}
    if not a.no_extra:
        line["shim_frame"] = shim_frame()
    ctx.close()
    if a.serve > 0:
        line["serving"] = serving(a.serve, a.slots, a.seed, base,
                                  check_threads=0 if a.no_cpu else host_info()["omp_threads"],
                                  batch_groups=a.batch_groups)
        if a.throughput_slots > 0:
            line["serving"]["throughput_config"] = serving_child(a)
    threads = host_info()["omp_threads"]
    if not a.no_configs:
        line["configs"] = configs(base, a.seed, threads, a.cfg5_oracle)
        line["configs"]["cfg3_" + wname] = "the headline: value, config, roofline, fp64_drop_in"
        line["configs"]["cfg4_serving"] = "serving" if a.serve > 0 else "not run (--serve 0)"
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(scene, threads)
        p_oracle = line["cpu_baseline"].pop("_pressure")
        scale = float(np.abs(p_oracle).max())
        if "mixed" in out:
            line["config"]["max_rel_p_minus_p_oracle"] = float(np.abs(p_mixed - p_oracle).max()) / scale
        if "fp64" in out and "fp64_drop_in" in line:
            line["fp64_drop_in"]["max_rel_p_minus_p_oracle"] = float(np.abs(p_fp64 - p_oracle).max()) / scale
        if not a.no_extra:
            ctx2 = elq.Context(0)
            line["cpu_baseline"]["frame_steps"] = cpu_baseline_frame(ctx2)
            ctx2.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
