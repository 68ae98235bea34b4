"""Benchmark of the particle-DEM step on one MI355X.

Workload: 1,048,576 monodisperse spheres (r = 1/128) on a jittered cubic
lattice in [-1,1]^3, 256^3 hash grid (cell = 2r), spring .5, damping .02,
shear .1, boundary damping -.5, gravity -0.0003, dt 0.5.  A step is one
ParticleSystem::update: integrate -> hash -> sort -> reorder -> collide.

The lattice is first advanced --settle untimed steps (default 1500), by
which time it has fallen and settled into a bed; then W warm-up steps, then
K timed steps.  The timed region is K steps of the multi-step run
(graph-launched) with the state resident in HBM.

Prints ONE JSON line (the last line of stdout) with the throughput in
particle-steps/s, the HBM roofline of the dominant kernel of the timed path
(algorithmic bytes of the reference stages it performs over its
HIP-event-timed average duration; HBM traffic from the committed rocprofv3
counter summary), the FP32-vs-FP64 force error on the timed state, the CPU
baseline (the oracle, the CPU restatement of the same step, timed on this
host's available threads and on one) and the surrogate-boundary leg.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--settle S] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s at 1M spheres, % of HBM peak; SDF-MLP vs distance-map speedup"
HBM_PEAK_GBS = 8000.0
N = 1 << 20
GRID = 256
SORT_PASSES = 3  # 8-bit digits of the 24-bit cell hashes of a 256^3 grid
RADIUS = 1.0 / 128.0
DT = 0.5
SEED = 1973
JITTER = 0.1
VAMP = 0.01


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class ProfilerRegion:
    """roctxProfilerResume / roctxProfilerPause around the measured region,
    so that `rocprofv3 --selected-regions` traces exactly the timed steps
    (and the copy probe that calibrates its byte counters); no-ops without a
    profiler or without the roctx library."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "libroctx64.so.4"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            self.lib.roctxProfilerResume.argtypes = [ctypes.c_uint64]
            self.lib.roctxProfilerPause.argtypes = [ctypes.c_uint64]

    def resume(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)

    def pause(self):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)


REGION = ProfilerRegion()
PROFILE_LEG = "dem"  # the leg whose measured region is bracketed (--profile-leg)


def region_resume(leg: str):
    if leg == PROFILE_LEG:
        REGION.resume()


def region_pause(leg: str):
    if leg == PROFILE_LEG:
        REGION.pause()


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cell_counts(cs: np.ndarray, grid: int) -> tuple[int, int]:
    """C_occ (occupied cells) and C_touch (distinct cells in the 27-cell
    neighbourhoods of occupied cells, with the grid's wrap)."""
    occ = np.nonzero(cs != 0xFFFFFFFF)[0].astype(np.int64)
    g = grid
    x, y, z = occ % g, (occ // g) % g, occ // (g * g)
    mark = np.zeros(g ** 3, bool)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                mark[(((z + dz) % g) * g + ((y + dy) % g)) * g + ((x + dx) % g)] = True
    return int(len(occ)), int(mark.sum())


def stage_bytes(name: str, n: int, c_occ: int, c_touch: int, path: str = "run") -> int | None:
    """Algorithmic HBM bytes of one launch of a stage: every array it must
    touch read or written once, cell tables counted per occupied (or
    neighbourhood-touched) cell, no dense table clears.  `path` is "run"
    (the multi-step run, which keeps the particle state in sorted-slot order
    between fused steps) or "update" (single updates)."""
    run = path == "run"
    if name == "collide":  # own pos/vel 32 B, fixed-point coordinates 8 B, index 4 B,
        # newVel 16 B (the neighbours' hashes ride in sortedVel.w); + the positions
        # scattered back to id order by a run's last step
        return 60 * n + (16 * n if run else 0) + 8 * c_touch
    if name == "collide_integrate":  # own pos/vel 32, coordinates 8, index 4; pos/vel 32,
        return 84 * n + 8 * c_touch  # key 4 at the slot, id of slot 4 (newVel fused away)
    if name == "cell_bounds":  # sorted hashes 4 B, cellStart/cellEnd per occupied cell; in
        # a run also the ids 4 B (cells whose ids are out of order, a data-dependent
        # share, are rewritten on top)
        return (8 if run else 4) * n + 8 * c_occ
    if name.startswith("sort_pass"):  # keys+values read and written once (the first pass
        # reads keys only: its values are the keys' positions); the last pass also gathers pos/vel into
        # sorted order (32 B read, 32 B written), writes the fixed-point coordinates
        # (8 B) and, in a run, the ids through the slot map (4 B read, 4 B written)
        p = int(name[len("sort_pass"):])
        last = p == SORT_PASSES - 1
        return (12 if p == 0 else 16) * n + ((72 * n + (8 * n if run else 0)) if last else 0)
    if name == "integrate_hash":  # pos/vel read+write 64 B, key/value 8 B
        return 72 * n
    if name == "row_changers":  # next keys and current hashes 8 B (the few changers' bucket
        return 8 * n  # entries and the emptied cells' table entries on top)
    if name == "row_sort":  # keys and ids 8 B, pos/vel gathered 32 B; hash, id 8 B, sortedPos/
        return 88 * n + 8 * c_occ  # sortedVel 32 B and fixed-point coordinates 8 B written; tables
    return None


def reference_stage_bytes(name: str, n: int, c_occ: int, c_touch: int) -> int | None:
    """Algorithmic bytes of one launch by the reference-stage accounting:
    integrate 64 N, calcHash 24 N, sort 52 N (one histogram read + three
    passes of 8 B read and 8 B written), reorder 72 N + 16 C_occ, collide
    52 N + 8 C_touch.  A fused launch counts every stage it performs."""
    collide, integrate, calc_hash = 52 * n + 8 * c_touch, 64 * n, 24 * n
    if name == "collide_integrate":
        return collide + integrate + calc_hash
    if name == "collide":
        return collide
    if name == "integrate_hash":
        return integrate + calc_hash
    return None


def profiled_traffic(kernel: str) -> int | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size)."""
    f = ROOT / "profiles" / "traffic_dem-1M-spheres.md"
    try:
        for ln in f.read_text().splitlines():
            parts = ln.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(parts[2])
    except OSError:
        pass
    return None


def gpu_leg(steps: int, warmup: int, timing_steps: int, settle: int,
            progress_every: int = 0) -> dict:
    from fullc0ntact_amd import ParticleSystem, make_params

    ps = ParticleSystem(N, (GRID, GRID, GRID))
    ps.m_params = make_params(N, GRID, RADIUS)
    ps.reset(ps.CONFIG_GRID, seed=SEED, jitter_frac=JITTER, vel_amp=VAMP)
    ps.sync()
    # The lattice falls, hits the floor and settles into a bed within about
    # 1200 steps; the timed steps are taken from the settled bed so that the
    # figure does not depend on how many warm-up steps the caller asks for.
    log(f"[bench] settling the bed: {settle} untimed steps")
    chunk = progress_every if progress_every > 0 else max(settle, 1)
    done = 0
    while done < settle:
        k = min(chunk, settle - done)
        ps.run(DT, k)
        ps.sync()
        done += k
        if progress_every > 0:
            log(f"[bench] untimed step {done}")
    region_resume("dem")
    probe_bytes = 1 << 30  # 1 GiB per buffer: far beyond the 256 MB last-level cache
    probe_ms = ps.ctx.copy_probe(probe_bytes, 10)
    copy_gbs = 2 * probe_bytes / (probe_ms * 1e-3) / 1e9
    log(f"[bench] copy probe: {copy_gbs:.0f} GB/s (read + write of 1 GiB)")
    # the warm-up steps come after the probe, which evicts the bed's state
    # from the last-level cache
    if warmup > 0:
        log(f"[bench] warmup {warmup} steps")
        ps.run(DT, warmup)
        ps.sync()
    log(f"[bench] timed {steps} steps")
    t0 = time.perf_counter()
    ps.run(DT, steps)
    ps.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0

    # Kernel durations: HIP events around every launch on the system's
    # stream, over a further run of the same multi-step path (launched
    # eagerly so that each launch is bracketed), then over single updates.
    log(f"[bench] per-kernel event timing over {timing_steps} steps")
    ctx = ps.ctx
    ctx.enable_timing(True)
    ctx.reset_timing()
    ps.run(DT, timing_steps)
    run_stages = ctx.timing()
    ctx.reset_timing()
    for _ in range(timing_steps):
        ps.update(DT)
    update_stages = ctx.timing()
    ctx.enable_timing(False)
    ps.sync()
    region_pause("dem")
    cs = ps.download("cellStart")
    c_occ, c_touch = cell_counts(cs, GRID)
    state = (ps.getArray(ps.POSITION), ps.getArray(ps.VELOCITY))
    drop_in = drop_in_leg(ps, timing_steps)
    ps.close()
    return dict(elapsed=elapsed, run_stages=run_stages, update_stages=update_stages,
                c_occ=c_occ, c_touch=c_touch, copy_gbs=copy_gbs, state=state,
                step_index=settle + warmup, drop_in=drop_in)


def drop_in_leg(ps, steps: int) -> dict:
    """The per-stage path an unchanged reference caller takes through the
    particleSystem.cuh-style entry points (ParticleSystem::update,
    particleSystem.cpp:222-284, over the C ABI): integrate, calcHash,
    sortParticles, reorderDataAndFindCellStart (which clears both cell tables
    densely, as the reference's memset does), collide, on the system's own
    device arrays, from the bed the run reached.  Per-stage times with HIP
    events on the context stream."""
    import ctypes as C
    from fullc0ntact_amd import _lib
    lib, check = _lib.lib, _lib.check
    ctx = ps.ctx
    ctx.set_parameters(ps.m_params)
    d = ps.device_arrays()
    C_ = ps.m_params.numCells
    h = ctx.handle

    def step():
        check(lib.fc_integrate_system(h, d["pos"], d["vel"], C.c_float(DT), N))
        check(lib.fc_calc_hash(h, d["hash"], d["index"], d["pos"], N))
        check(lib.fc_sort_particles(h, d["hash"], d["index"], N))
        check(lib.fc_reorder_data_and_find_cell_start(h, d["cellStart"], d["cellEnd"],
                                                      d["sortedPos"], d["sortedVel"], d["hash"],
                                                      d["index"], d["pos"], d["vel"], N, C_))
        check(lib.fc_collide(h, d["vel"], d["sortedPos"], d["sortedVel"], d["index"],
                             d["cellStart"], d["cellEnd"], N, C_))

    step()  # first touch of the caller-style sequence
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    el = time.perf_counter() - t0
    ctx.enable_timing(True)
    ctx.reset_timing()
    for _ in range(steps):
        step()
    t = ctx.timing()
    ctx.enable_timing(False)
    us = {k: round(v[1] / max(v[0], 1) * 1e3, 3) for k, v in t.items()}
    log(f"[bench] per-stage drop-in path: {el / steps * 1e3:.4f} ms/step {us}")
    return {"value": round(N * steps / el, 1), "unit": "particle-steps/s",
            "ms_per_step": round(el / steps * 1e3, 5), "steps": steps, "stages_us": us}


def force_error(state, threads: int) -> dict:
    """Error of the FP32 collide (bit-identical to the GPU's) against the
    binary64 restatement of the same sum, on the bed state the GPU reached:
    both over the same neighbour pairs (the cells are the FP32 ones)."""
    from oracle import oracle as orc
    orc.set_threads(threads)
    P = orc.make_params(N, GRID, RADIUS)
    pos, vel = state
    h, i = orc.calc_hash(pos, P)
    h, i = orc.sort_pairs(h, i)
    cs, ce, sp, sv = orc.reorder(h, i, pos, vel, P.numCells)
    f32 = orc.collide(sp, sv, i, cs, ce, P, newVel4=vel)[:, :3].astype(np.float64) - vel[:, :3]
    f64 = orc.collide_f64(sp, sv, i, cs, ce, P) - vel[:, :3]
    orc.set_threads(1)
    # spheres clamped into the same box corner coincide and get a NaN force
    # on both sides (0 * inf, as in the reference); they are counted, not compared
    finite = np.all(np.isfinite(f32), axis=1) & np.all(np.isfinite(f64), axis=1)
    f32, f64 = f32[finite], f64[finite]
    d = np.abs(f32 - f64)
    mag = np.linalg.norm(f64, axis=1)
    touching = mag > 0
    rel = np.linalg.norm(f32 - f64, axis=1)[touching] / mag[touching]
    return {"max_abs_err_over_max_force": float(d.max() / np.abs(f64).max()),
            "median_rel_err": float(np.median(rel)) if rel.size else 0.0,
            "p99_rel_err": float(np.quantile(rel, 0.99)) if rel.size else 0.0,
            "max_rel_err": float(rel.max()) if rel.size else 0.0,
            "spheres_with_force": int(touching.sum()),
            "spheres_nan_both": int((~finite).sum()),
            "what": "per-sphere force of one collide on the timed bed state: FP32 (the GPU's, "
                    "bit-identical to the oracle's) against the binary64 restatement"}


def host_threads() -> tuple[int, dict]:
    """Host threads the CPU baseline may use: the CPUs this process may run
    on, capped by OMP_NUM_THREADS where the environment sets it."""
    avail = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    cap = int(env) if env and env.isdigit() and int(env) > 0 else avail
    info = {"nproc": os.cpu_count(), "affinity_cpus": avail, "OMP_NUM_THREADS": env,
            "model": cpu_model()}
    return max(1, min(avail, cap)), info


SDF_BODIES_LATTICE = 4        # 64 rigid bodies on a 4^3 lattice of [-1,1]^3
SDF_SHAPES = None             # every body its own seeded shape and network
SDF_SPHERES = 1 << 21         # 2,097,152 spheres
SDF_RADIUS = 1.0 / 256
SDF_MAP_CELLS = 256
SDF_SAMPLES = 1 << 20         # exact-distance training samples per network
SDF_EPOCHS = 30
SDF_BATCH = 16384


def sdf_leg(reps: int) -> dict:
    """Boundary contacts of 2M spheres with 64 rigid bodies: the exact
    distance-map query for every sphere in a body box, against the
    surrogate path (BF16 MFMA network cull, exact query only in the band).
    Spheres are uniform in [-1,1]^3 in random order, as the integrate stage
    meets them once a bed has mixed."""
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.boundary import build_scene

    ctx = _lib.Context()
    log(f"[bench] sdf: building {SDF_BODIES_LATTICE ** 3} shapes with {SDF_MAP_CELLS}^3 maps "
        f"and networks on {SDF_SAMPLES} samples each")
    sc = build_scene(ctx, lattice=SDF_BODIES_LATTICE, cells=SDF_MAP_CELLS, samples=SDF_SAMPLES,
                     epochs=SDF_EPOCHS, batch=SDF_BATCH, seed=SEED, radius=SDF_RADIUS,
                     shapes=SDF_SHAPES, knot_segments=(128, 8), ellipsoid_segments=(32, 16),
                     log=lambda m: log(f"[bench] sdf: {m}"))
    log(f"[bench] sdf: maps {sc.build_s:.1f} s, training + calibration {sc.train_s:.1f} s")
    # spheres in the free space around the bodies: uniform draws whose
    # centre lies inside a body (map distance < 0) are dropped
    rng = np.random.default_rng(SEED)
    draw = np.ones((SDF_SPHERES + SDF_SPHERES // 4, 4), np.float32)
    draw[:, :3] = rng.uniform(-1.0, 1.0, (len(draw), 3))
    dd = _lib.DeviceArray.from_host(ctx, draw)
    inside, _ = sc.bodies.contacts(dd, len(draw), 0.0)
    keep = inside.download(np.uint8, len(draw)) == 0
    pos = np.ascontiguousarray(draw[keep][:SDF_SPHERES])
    assert len(pos) == SDF_SPHERES
    dp = _lib.DeviceArray.from_host(ctx, pos)
    contact = _lib.DeviceArray(ctx, SDF_SPHERES)
    cand = _lib.DeviceArray(ctx, SDF_SPHERES)
    fe_dev, ce = sc.bodies.contacts(dp, SDF_SPHERES, SDF_RADIUS)
    fe = fe_dev.download(np.uint8, SDF_SPHERES)
    fs_dev, cs = sc.bodies.contacts(dp, SDF_SPHERES, SDF_RADIUS, sc.surrogate, sc.margins)
    fs = fs_dev.download(np.uint8, SDF_SPHERES)
    agreement = float(np.mean(fe == fs))
    inter = int(np.sum(fe & fs))
    union = int(np.sum(fe | fs))
    times = {}
    region_resume("sdf")
    for name, sur in (("map", None), ("sdf", sc.surrogate)):
        for _ in range(3):  # warm-up
            sc.bodies.contacts(dp, SDF_SPHERES, SDF_RADIUS, sur, sc.margins if sur else None,
                               cand, contact, counts=False)
        ctx.sync()
        ctx.enable_timing(True)
        ctx.reset_timing()
        for _ in range(reps):
            sc.bodies.contacts(dp, SDF_SPHERES, SDF_RADIUS, sur, sc.margins if sur else None,
                               cand, contact, counts=False)
        t = ctx.timing()
        ctx.enable_timing(False)
        times[name] = {k: v[1] / max(v[0], 1) * 1e3 for k, v in t.items()}
    region_pause("sdf")
    t_map = times["map"]["boundary_exact"]
    ts = times["sdf"]
    t_sdf = ts["boundary_bucket"] + ts["boundary_cull"] + ts["boundary_exact_band"]
    log(f"[bench] sdf: map path {t_map:.1f} us, surrogate path {t_sdf:.1f} us "
        f"(bucket {ts['boundary_bucket']:.1f} + cull {ts['boundary_cull']:.1f} + band "
        f"{ts['boundary_exact_band']:.1f}), agreement {agreement:.6f}")
    out = {"bodies": SDF_BODIES_LATTICE ** 3, "networks": len(sc.shape_maps),
           "samples_per_network": SDF_SAMPLES, "epochs": SDF_EPOCHS, "spheres": SDF_SPHERES,
           "radius": SDF_RADIUS, "map_cells": SDF_MAP_CELLS,
           "margins": "lattice bound (boundary.lattice_margins)",
           "margin_range": [round(float(np.min(sc.margins)), 6),
                            round(float(np.max(sc.margins)), 6)],
           "map_path_us": round(t_map, 2), "surrogate_path_us": round(t_sdf, 2),
           "bucket_us": round(ts["boundary_bucket"], 2),
           "cull_us": round(ts["boundary_cull"], 2),
           "band_exact_us": round(ts["boundary_exact_band"], 2),
           "speedup": round(t_map / t_sdf, 3), "contact_agreement": agreement,
           "contact_set_jaccard": (inter / union) if union else 1.0,
           "in_box": int(ce[0]), "exact_queries_surrogate": int(cs[1]), "contacts": int(ce[2]),
           "map_build_s": round(sc.build_s, 1), "train_s": round(sc.train_s, 1)}
    sc.bodies.close()
    sc.surrogate.close()
    for m in sc.shape_maps:
        m.close()
    ctx.close()
    return out


BATCH_SCENES, BATCH_LO, BATCH_HI, BATCH_GRID = 256, 2000, 200000, 64
HOPPER_N, HOPPER_GRID, HOPPER_MEDIAN = 1 << 22, 512, 1.0 / 512


def hopper_leg(settle: int, steps: int) -> dict:
    """The polydisperse hopper: 4,194,304 spheres with log-normal radii
    (median 1/512, sigma 0.25, clipped to a 3x ratio) on the jittered
    lattice, the cell the largest diameter (a 512^3 grid), the default
    parameters, and a (2, 3) torus-knot tube of 65,536 triangles as the
    rigid boundary through an FP32 256^3 distance map.  `settle` untimed
    steps of the fall, then a graph-launched run of `steps` steps timed.
    Only the early fall is timed: at this radius the reference's spring
    cannot hold the spheres apart at dt 0.5, and within a few hundred steps
    the pile collapses to over a thousand spheres per cell, whose cost per
    step is the collapse's, not the step's."""
    from fullc0ntact_amd import ParticleSystem, make_params, meshes, statistics
    from fullc0ntact_amd.distance_map import DistanceMap

    rmax = float(np.float32(HOPPER_MEDIAN * np.sqrt(3.0)))
    ps = ParticleSystem(HOPPER_N, (HOPPER_GRID,) * 3)
    t0 = time.perf_counter()
    v, f = meshes.torus_knot(2, 3, 1024, 32, radius=0.12, scale=0.25, center=(0.0, 0.0, -0.55))
    gm = DistanceMap.from_mesh(ps.ctx, v, f, (256,) * 3, 2.0 / 256, (-1.0, -1.0, -1.0))
    map_s = time.perf_counter() - t0
    ps.m_params = make_params(HOPPER_N, HOPPER_GRID, rmax)
    ps.reset(ps.CONFIG_GRID, seed=SEED, jitter_frac=JITTER, vel_amp=VAMP)
    pos = ps.getArray(ps.POSITION).copy()
    pos[:, 3] = statistics.polydisperse_radii(HOPPER_N, HOPPER_MEDIAN, 0.25, 3.0, seed=5)
    ps.setArray(ps.POSITION, pos, 0, HOPPER_N)
    ps.setPolydisperse(True)
    ps.setBoundary(gm)
    ps.run(DT, settle)
    ps.sync()
    t0 = time.perf_counter()
    ps.run(DT, steps)
    ps.sync()
    el = time.perf_counter() - t0
    ps.setBoundary(None)
    ps.close()
    gm.close()
    log(f"[bench] hopper: {HOPPER_N * steps / el:.3e} particle-steps/s "
        f"({el / steps * 1e3:.3f} ms per step; map built in {map_s:.1f} s)")
    return {"particles": HOPPER_N, "grid": [HOPPER_GRID] * 3, "radius_median": HOPPER_MEDIAN,
            "triangles": int(len(f)), "map_cells": [256] * 3, "timed_from_step": settle,
            "steps": steps, "value": round(HOPPER_N * steps / el, 1),
            "unit": "particle-steps/s", "ms_per_step": round(el / steps * 1e3, 4),
            "map_build_s": round(map_s, 2)}


def batch_leg(steps: int) -> dict:
    """The batched ensemble: 256 independent scenes, sizes log-uniform in
    [2000, 200000] (about 11M spheres), spring and damping jittered +-20%
    per scene, r = 1/64 on a 64^3 grid each, one launch per stage with a
    scene-segmented hash, graph-launched, launch shape autotuned on the
    batch; plus the measured cost of a grid barrier in one persistent
    (cooperative) launch against a kernel boundary in a graph."""
    from fullc0ntact_amd import _lib, make_params
    from fullc0ntact_amd.batch import Batch, ensemble_scene_counts, jittered_params
    ctx = _lib.Context()
    counts = ensemble_scene_counts(BATCH_SCENES, BATCH_LO, BATCH_HI, seed=SEED)
    params = jittered_params(make_params(0, BATCH_GRID, 1.0 / 64), BATCH_SCENES, 0.2, seed=SEED)
    for k, p in enumerate(params):
        p.numBodies = int(counts[k])
    b = Batch(counts, (BATCH_GRID,) * 3, context=ctx)
    b.set_params(params)
    b.reset(1, SEED, 0.1, 0.01)
    log(f"[bench] batch: {BATCH_SCENES} scenes, {b.n} spheres; autotuning")
    b.run(DT, 2)
    best, ms = b.autotune(DT, 3)
    b.run(DT, 10)  # graph capture of the run's steps with the chosen shape
    ctx.sync()
    region_resume("batch")
    t0 = time.perf_counter()
    b.run(DT, steps)
    ctx.sync()
    el = time.perf_counter() - t0
    region_pause("batch")
    e = b.stats_fp64()
    assert np.all(np.isfinite(e))
    n = b.n
    b.close()
    barrier_us, kernel_us = ctx.sync_probe(256, 1000)
    ctx.close()
    log(f"[bench] batch: {n * steps / el:.3e} particle-steps/s, {el / steps * 1e3:.3f} ms/step; "
        f"grid barrier {barrier_us:.2f} us vs graph kernel boundary {kernel_us:.2f} us")
    return {"scenes": BATCH_SCENES, "particles": int(n), "grid": [BATCH_GRID] * 3,
            "steps": steps, "value": round(n * steps / el, 1), "unit": "particle-steps/s",
            "ms_per_step": round(el / steps * 1e3, 4),
            "tuning": {"integrate_per_thread": int(best.integratePerThread),
                       "collide_block": int(best.collideBlock),
                       "candidate_ms": [round(float(x), 4) for x in ms]},
            "grid_barrier_us": round(barrier_us, 3),
            "graph_kernel_boundary_us": round(kernel_us, 3)}


def cpu_leg(target_seconds: float, threads: int) -> dict:
    from oracle import oracle as orc

    orc.set_threads(threads)
    P = orc.make_params(N, GRID, RADIUS)
    pos, vel = orc.reset(N, 1, SEED, RADIUS, JITTER, VAMP)
    st = orc.State(pos, vel, P.numCells)
    t0 = time.perf_counter()
    st.update(DT, P)  # warm-up step (first touch of every array)
    first = time.perf_counter() - t0
    k = int(max(1, min(500, target_seconds / max(first, 1e-3))))
    t0 = time.perf_counter()
    st.run(DT, k, P)
    el = time.perf_counter() - t0
    orc.set_threads(1)
    return dict(value=N * k / el, steps=k, seconds=el)


SMALL_N, SMALL_GRID, SMALL_RADIUS, SMALL_STEPS = 4096, 64, 1.0 / 64, 1000


def small_case_leg(threads: int) -> dict:
    """The 4096-sphere case (r = 1/64, 64^3 grid, 1000 steps of dt 0.5):
    GPU graph-launched run, and the CPU restatement on `threads` threads and
    on one."""
    from fullc0ntact_amd import ParticleSystem, make_params
    from oracle import oracle as orc

    ps = ParticleSystem(SMALL_N, (SMALL_GRID,) * 3)
    ps.m_params = make_params(SMALL_N, SMALL_GRID, SMALL_RADIUS)
    ps.reset(ps.CONFIG_GRID, seed=SEED, jitter_frac=JITTER, vel_amp=VAMP)
    ps.run(DT, 20)  # warm-up (graph capture)
    ps.reset(ps.CONFIG_GRID, seed=SEED, jitter_frac=JITTER, vel_amp=VAMP)
    ps.sync()
    t0 = time.perf_counter()
    ps.run(DT, SMALL_STEPS)
    ps.sync()
    gpu_s = time.perf_counter() - t0
    # per-kernel event times of the same run path (launched eagerly)
    ps.ctx.enable_timing(True)
    ps.ctx.reset_timing()
    ps.run(DT, 100)
    st = ps.ctx.timing()
    ps.ctx.enable_timing(False)
    ps.close()
    out = {"particles": SMALL_N, "grid": [SMALL_GRID] * 3, "radius": SMALL_RADIUS,
           "steps": SMALL_STEPS, "gpu_value": round(SMALL_N * SMALL_STEPS / gpu_s, 1),
           "gpu_us_per_step": round(gpu_s / SMALL_STEPS * 1e6, 2),
           "gpu_stages_us": {k: round(v[1] / max(v[0], 1) * 1e3, 3) for k, v in st.items()}}
    P = orc.make_params(SMALL_N, SMALL_GRID, SMALL_RADIUS)
    for key, t in (("cpu_value", threads), ("cpu_value_1_core", 1)):
        orc.set_threads(t)
        pos, vel = orc.reset(SMALL_N, 1, SEED, SMALL_RADIUS, JITTER, VAMP)
        st = orc.State(pos, vel, P.numCells)
        t0 = time.perf_counter()
        st.run(DT, SMALL_STEPS, P)
        out[key] = round(SMALL_N * SMALL_STEPS / (time.perf_counter() - t0), 1)
    orc.set_threads(1)
    out["cpu_cores"] = threads
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--settle", type=int, default=1500,
                    help="untimed steps before the warmup, to reach the settled bed")
    ap.add_argument("--timing-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-sdf", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--sdf-reps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-hopper", action="store_true", help="skip the polydisperse hopper leg")
    ap.add_argument("--hopper-only", action="store_true",
                    help="run only the polydisperse hopper leg and print its result")
    ap.add_argument("--hopper-settle", type=int, default=10)
    ap.add_argument("--hopper-steps", type=int, default=20)
    ap.add_argument("--profile-leg", choices=["dem", "sdf", "batch"], default="dem",
                    help="leg whose measured region is bracketed for rocprofv3 --selected-regions")
    ap.add_argument("--progress-every", type=int, default=0,
                    help="run the warmup in chunks of this many steps, logging each")
    ap.add_argument("--small-case-only", action="store_true",
                    help="run only the 4096-sphere case (GPU and CPU) and print its result, "
                         "for profiling that case on its own")
    a = ap.parse_args(argv)
    global PROFILE_LEG
    PROFILE_LEG = a.profile_leg
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {a.gpus} is not supported "
              "(this build has no multi-GPU path).", file=sys.stderr)
        return 2
    if a.small_case_only:
        print(json.dumps({"small_case": small_case_leg(host_threads()[0])}), flush=True)
        return 0
    if a.hopper_only:
        print(json.dumps({"polydisperse_hopper": hopper_leg(a.hopper_settle, a.hopper_steps)}),
              flush=True)
        return 0

    g = gpu_leg(a.steps, a.warmup, a.timing_steps, a.settle, a.progress_every)
    ms = g["elapsed"] / a.steps * 1e3
    value = N * a.steps / g["elapsed"]

    def per_launch_us(stages):
        return {k: round(v[1] / max(v[0], 1) * 1e3, 3) for k, v in stages.items()}

    run_us = per_launch_us(g["run_stages"])
    upd_us = per_launch_us(g["update_stages"])
    for k, v in g["run_stages"].items():
        log(f"[bench] run    {k:18s} {run_us[k]:10.2f} us x {v[0]}")
    for k, v in upd_us.items():
        log(f"[bench] update {k:18s} {v:10.2f} us")

    # The dominant kernel of the timed path: the largest total event time.
    dom = max(g["run_stages"].items(), key=lambda kv: kv[1][1])[0]
    nbytes = reference_stage_bytes(dom, N, g["c_occ"], g["c_touch"])
    impl_bytes = stage_bytes(dom, N, g["c_occ"], g["c_touch"])
    roof = None
    if nbytes:
        us = run_us[dom]
        achieved = nbytes / (us * 1e-6) / 1e9
        traffic = profiled_traffic(dom)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom, "bytes_per_launch": nbytes,
                "bytes_accounting": "reference stages fused in the launch: collide 52N + 8C_touch"
                                    " + integrate 64N + calcHash 24N",
                "implementation_bytes_per_launch": impl_bytes, "avg_launch_us": us,
                "copy_probe_GBs": round(g["copy_gbs"], 1)}
    threads, host = host_threads()
    ferr = None

    sdf = None
    if not a.no_sdf:
        sdf = sdf_leg(a.sdf_reps)
    batch = None
    if not a.no_batch:
        batch = batch_leg(a.batch_steps)
    hopper = None
    if not a.no_hopper:
        hopper = hopper_leg(a.hopper_settle, a.hopper_steps)

    cpu = None
    if not a.no_cpu:
        # the CPU leg: the oracle's binary64 restatement states the FP32
        # force error on the timed bed, then the CPU step is timed
        log(f"[bench] FP32 vs FP64 force error on the timed bed ({threads} threads)")
        ferr = force_error(g["state"], threads)
        log(f"[bench] CPU baseline with {threads} threads on {host}")
        c = cpu_leg(a.cpu_seconds, threads)
        log("[bench] CPU baseline on one core")
        c1 = cpu_leg(a.cpu_seconds / 2, 1)
        log("[bench] 4096-sphere case, GPU and CPU")
        small = small_case_leg(threads)
        cpu = {"value": round(c["value"], 1), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"{c['steps']} updates of the same 1,048,576-sphere lattice after one "
                         f"warm-up update, {threads} OpenMP threads on {host['model']} "
                         f"({c['seconds']:.1f} s)",
               "host": host,
               "value_1_core": round(c1["value"], 1),
               "sample_1_core": f"{c1['steps']} updates on one thread ({c1['seconds']:.1f} s)",
               "small_case": small}

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "dem-1M-spheres", "particles": N, "grid": [GRID] * 3,
                   "radius": RADIUS, "dt": DT, "seed": SEED, "C_occ": g["c_occ"],
                   "C_touch": g["c_touch"], "timed_from_step": g["step_index"]},
        "force_error_fp32_vs_fp64": ferr,
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_us": {"run": run_us, "update": upd_us},
        "drop_in_per_stage_path": g["drop_in"],
        "sdf_vs_distance_map": sdf,
        "batched_ensemble": batch,
        "polydisperse_hopper": hopper,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
