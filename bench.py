#!/usr/bin/env python3
"""Benchmark of the LJ hot path on one MI355X.

Workload (the value line): LJ fluid, N = 1,048,576 (FCC 64^3 x 4), reduced
density 0.8442, T = 0.72, rc = 2.5 sigma (shifted energies), velocity Verlet
dt = 0.005, Verlet list rebuilt and particles re-sorted into cell order every
20 steps; FP32 pair forces on FP64 displacements with FP64 integration and
reductions.  A "step" is one force + integration pass over all particles.

Also reported on the same line: the plain FP64 form of the same run, the
mixed form's force error against FP64 and its energy drift, the batched
replica workload (256 LJ replicas, N log-uniform in [500, 20000]) as one
batched launch per step against each replica run alone in FP64, the
roofline of the pair kernel (HIP events on the library's stream) and the
CPU restatement of the reference loop timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import particlebase_amd as pb  # noqa: E402
from particlebase_amd import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
ALG_BYTES = 98.0          # FP64 r, v read + written (96 B) + u16 type read
RC2 = 6.25
DT = 0.005
METRIC = "LJ atom-steps/s at 1M atoms (roofline %) + batched-vs-plain speedup at 1e-5 err"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_system(ctx, w_or_batch):
    if isinstance(w_or_batch, workloads.Batch):
        r, v, m, t, cells = w_or_batch.concat()
        s = pb.System(ctx, w_or_batch.counts)
        s.set_cell(cells)
        s.import_particles(r, m, v=v, type=t)
    else:
        w = w_or_batch
        s = pb.System(ctx, [w.n])
        s.set_cell(w.cell)
        s.import_particles(w.r, w.m, v=w.v, type=w.type)
    s.lj_set_params([1.0, 1.0], 1)
    s.lj_set_cutoff(RC2, shifted=True)
    s.clear_force(True)
    return s


# Timed runs time every TIMING_STRIDE-th step kernel, the step after each
# rebuild and every rebuild with HIP events (the kernel's own dispatch
# timestamps on the library's stream); a timed dispatch idles the GPU about
# 9 us before it, so timing every step would slow the run it measures.
TIMING_STRIDE = 4


def timed_run(ctx, s, steps, precision, skin, rebuild_every, energy_every, timing=TIMING_STRIDE):
    ctx.sync()
    t0 = time.perf_counter()
    e, rc = s.lj_run(DT, steps, precision=precision, skin=skin, rebuild_every=rebuild_every,
                     energy_every=energy_every, timing=timing)
    ctx.sync()
    return time.perf_counter() - t0, e


def choose_tiles(s, a, precision):
    """The chosen tile shape: the autotuner's (pb_lj_autotune, timed from the
    warmed-up state, which it restores), or the one given on the command
    line (profiling runs, so the kernel trace holds only timed launches)."""
    if a.tile_fill > 0:
        s.lj_set_tiles(a.tile_w or 2, a.tile_fill)
        return {"tile_w": a.tile_w or 2, "tile_fill": a.tile_fill, "autotuned": False}
    t = s.lj_autotune(DT, steps=20, precision=precision, skin=a.skin,
                      rebuild_every=a.rebuild_every)
    return {"tile_w": t["tile_w"], "tile_fill": t["tile_fill"], "autotuned": True}


def drift_per_1e4(e, n, energy_every):
    et = (e[:, 0, 0] + e[:, 0, 1]) / n
    if et.size < 4:
        return None
    t = np.arange(et.size) * energy_every
    slope = np.polyfit(t, et, 1)[0]
    return float(slope * 1e4)


def drift_direct(ctx, precision, skin, rebuild_every, steps=10000, warm=1000):
    """The direct energy-drift figure: the N=4000 LJ fluid (FCC 10^3 x 4,
    rho 0.8442, T 0.72, rc 2.5 shifted energy, dt 0.005) melted for `warm`
    fused steps, then `steps` more with the total energy recorded every 100
    steps: (E_end - E_start) / N and the fitted slope per 10^4 steps."""
    w = workloads.lj_config1()
    s = make_system(ctx, w)
    s.lj_run(DT, warm, precision=precision, skin=skin, rebuild_every=rebuild_every,
             energy_every=0)
    e, _ = s.lj_run(DT, steps, precision=precision, skin=skin, rebuild_every=rebuild_every,
                    energy_every=100)
    et = (e[:, 0, 0] + e[:, 0, 1]) / w.n
    del s
    return {"delta_e_per_atom": float(et[-1] - et[0]), "fit_per_1e4": drift_per_1e4(e, w.n, 100),
            "records": int(et.size), "steps": steps, "e_per_atom_start": float(et[0])}


def bench_gr(sb, reps):
    """g(r) of every replica (200 bins to L_k/2, the batched config's
    histogram), 3 timed calls after one warm-up; integer totals checked."""
    rsteps = np.array([np.float32(x.L / 400.0) for x in reps.systems])
    sb.gr_hist(rsteps, 200)
    t0 = time.perf_counter()
    for _ in range(3):
        hist = sb.gr_hist(rsteps, 200)
    elg = (time.perf_counter() - t0) / 3
    pairs = int(sum(int(n) * (int(n) - 1) // 2 for n in reps.counts))
    assert int(hist.sum()) == pairs
    return {"pairs": pairs, "ms": 1e3 * elg, "pairs_per_s": pairs / elg, "bins": 200}


def host_info():
    """This host's CPU: model name, logical CPUs, the CPUs this process may
    use (affinity) and the thread count of the lease (OMP_NUM_THREADS)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    lease = min(aff, int(omp)) if omp and omp.isdigit() else aff
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff,
            "omp_num_threads": omp, "lease_threads": lease}


def cpu_baseline(w, budget_s=20.0, cores=0):
    """The oracle's restatement of the reference loop (rebin, 27-cell forces,
    VV, wrap; FP64, OpenMP) on this host, on a bounded sample of steps;
    `cores` threads (0: every thread of the lease, host_info()["lease_threads"])."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pb_oracle as orc
    cores = cores or host_info()["lease_threads"]
    p = orc.lj_params([1.0, 1.0], 1, RC2, True)
    minv = 1.0 / w.m
    t0 = time.perf_counter()
    orc.lj_run(w.r, w.v, np.zeros_like(w.r), minv, w.type, w.cell, 2.5, RC2, p, DT, 1, cores)
    t1 = time.perf_counter() - t0
    steps = max(1, min(50, int(budget_s / max(t1, 1e-3)) - 1))
    t0 = time.perf_counter()
    orc.lj_run(w.r, w.v, np.zeros_like(w.r), minv, w.type, w.cell, 2.5, RC2, p, DT, steps, cores)
    el = time.perf_counter() - t0
    return {"value": w.n * steps / el, "unit": "atom-steps/s", "cores": cores, "kind": "port",
            "sample": f"{steps} reference-loop steps of the N={w.n} LJ fluid "
                      f"({el:.1f} s, FP64, OpenMP {cores} threads)"}


def bench_sph(ctx, a, config=False):
    """SPH dam break, the reference driver's step: bin, density,
    acceleration, gravity, leap-frog, inspectVelocity, absolute boundary.
    Reference physics (Lucy, support h = 1.7 dx) or the config physics
    (cubic spline of support 2h, h = 1.3 dx, Tait, artificial viscosity)."""
    dx = 0.08 if a.quick else 0.4 / 63
    w = workloads.sph_dam_break_config(dx) if config else workloads.sph_dam_break(dx)
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(w.r, w.m, is_fixed=w.fixed)
    if config:
        s.sph_set_wcsph(w.h, w.rho0, w.c[0], 7.0, 0.1, 0.0)
        sup = 2 * w.h
    else:
        s.sph_set_properties(w.mu0, w.c, w.h)
        sup = w.h
    s.setup_cells(sup, periodic=False)
    dt = 0.25 * w.h / w.c[0]

    def step():
        s.bin()
        s.sph_density()
        s.sph_acceleration()
        s.add_acceleration(2, -9.81)
        s.lf_step(dt)
        s.inspect_velocity(2 * sup / dt, 0.04, 0.0025)
        s.treat_absolute()

    for _ in range(3):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.sph_steps):
        step()
    ctx.sync()
    el = time.perf_counter() - t0
    v = w.n * a.sph_steps / el
    kind = "cubic spline h=1.3dx, Tait, artificial viscosity" if config else "Lucy h=1.7dx"
    log(f"sph ({kind}): N={w.n} ({int((w.fixed == 0).sum())} fluid), {v:.4g} particle-steps/s, "
        f"{1e3 * el / a.sph_steps:.3f} ms/step")
    return {"workload": f"SPH dam break, {kind}, dx={w.dx:.6g}, {int((w.fixed == 0).sum())} "
                        f"fluid + {int(w.fixed.sum())} wall particles, FP64",
            "value": v, "unit": "particle-steps/s", "ms_per_step": 1e3 * el / a.sph_steps,
            "roofline_frac": SPH_ALG_BYTES * v / 1e9 / HBM_PEAK_GBS,
            "alg_bytes": f"{SPH_ALG_BYTES:g} B per particle-step",
            "steps": a.sph_steps}


SPH_ALG_BYTES = 121.0        # per particle-step: FP64 r, v read + written, type/move, rho/1/rho/mu
DEM_ALG_BYTES = 156.0        # per particle-step: FP64 r, v, w/L read + written, radius, mass
DEM_ALG_BYTES_CONTACT = 52.0  # per contact: column index + 3-double history in and out


def write_dem_input(w, path):
    """The dem_settle driver's input file: n (uint32), then r, radius, mass
    and v as FP64 arrays."""
    with open(path, "wb") as f:
        f.write(np.array([w.n], dtype=np.uint32).tobytes())
        for x in (w.r, w.rad, w.m, np.zeros_like(w.r)):
            f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())


def bench_dem(a, law="linear"):
    """DEM settle as the workload states it: 200,000 spheres, radii U[0.9,
    1.1] mm, random non-overlapping insertion in 0.1 x 0.1 x 0.4 m, at rest;
    linear spring-dashpot contacts (kn 1e4 N/m, restitution 0.8) with a
    tangential spring (kt = 2/7 kn) and Coulomb friction mu 0.5, rotation,
    gravity, reflecting walls, adaptive time step (AdaptiveTime with
    rollback; under the linear law held at 0.2 sqrt(meff_min / kn)), 5000
    steps.  Run by the class-layer driver in the reference driver's call
    order (testDEM5.cu:181-234) from the workload written to a file; the
    driver times its step loop.  law="hertz": the reference's Hertz law on
    the same spheres."""
    import subprocess
    import tempfile
    w = workloads.dem_settle_config(20000 if a.quick else 200_000)
    exe = os.path.join(ROOT, "particlebase_amd", "bin", "dem_settle")
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        path = f.name
    write_dem_input(w, path)
    try:
        p = subprocess.run([exe, str(w.n), str(a.dem_steps), law, path], capture_output=True,
                           text=True, timeout=600)
    finally:
        os.unlink(path)
    if p.returncode != 0:
        raise RuntimeError(f"dem_settle failed: {p.stderr[-2000:]}")
    line = [x for x in p.stdout.splitlines() if x.startswith("done")][-1]
    d = {k: float(v) for k, v in (t.split("=") for t in line.split()[1:])}
    iters = d["iterations"]
    v = d["particle_steps_per_s"]
    alg = (DEM_ALG_BYTES * w.n + DEM_ALG_BYTES_CONTACT * d["mean_contacts"]) * iters / d["time_s"]
    log(f"dem ({law}): N={w.n}, {v:.4g} particle-steps/s, {1e3 * d['time_s'] / iters:.3f} ms/step,"
        f" {line}")
    lawtxt = ("linear spring-dashpot kn 1e4 N/m e 0.8, tangential spring + Coulomb mu 0.5"
              if law == "linear" else "reference Hertz law (E 2.11e9 Pa, sigma 0.29, e 0.8)")
    return {"workload": f"DEM settle, {w.n} spheres r in [0.9, 1.1] mm, random non-overlapping "
                        f"insertion in 0.1x0.1x0.4 m at rest, {lawtxt}, rotation, gravity, "
                        f"reflecting walls, adaptive dt, FP64",
            "value": v, "unit": "particle-steps/s", "ms_per_step": 1e3 * d["time_s"] / iters,
            "steps": int(d["steps"]), "iterations": int(iters), "rollbacks": int(d["rollbacks"]),
            "final_dt": d["dt"], "sim_time_s": d["t"], "mean_contacts": d["mean_contacts"],
            "max_contacts": int(d["max_contacts"]),
            "roofline_frac": alg / 1e9 / HBM_PEAK_GBS,
            "alg_bytes": f"{DEM_ALG_BYTES:g} B per particle-step + "
                         f"{DEM_ALG_BYTES_CONTACT:g} B per contact"}


def load_traffic():
    """HBM bytes per k_lj_step launch from the committed rocprofv3 counter
    passes (FETCH_SIZE / WRITE_SIZE, calibrated by a copy of known size):
    a stored profile figure, not measured in this run, with its source."""
    path = os.path.join(ROOT, "profiles", "lj_step_1M_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        t = d.get("tiles", {})
        src = (f"profiles/lj_step_1M_traffic.json: {d.get('measured', 'rocprofv3 counters')}, "
               f"tiles {t.get('tile_w')}/{t.get('tile_fill')}, "
               f"{d.get('kernel_avg_us', 0):.1f} us a launch there")
        return d.get("hbm_bytes_per_launch"), src
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--skin", type=float, default=0.3)
    ap.add_argument("--rebuild-every", type=int, default=20)
    ap.add_argument("--tile-w", type=int, default=0,
                    help="fixed tile shape (with --tile-fill) instead of the autotuner")
    ap.add_argument("--tile-fill", type=float, default=0.0)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--plain-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-sph", action="store_true")
    ap.add_argument("--no-dem", action="store_true")
    ap.add_argument("--no-drift", action="store_true",
                    help="skip the 10^4-step N=4000 energy-drift runs")
    ap.add_argument("--sph-steps", type=int, default=50)
    ap.add_argument("--dem-steps", type=int, default=5000)
    ap.add_argument("--quick", action="store_true", help="small sizes (debug)")
    ap.add_argument("--profile", action="store_true",
                    help="only the mixed 1M run (for rocprofv3 kernel traces / counters)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only a device copy of 256 MiB, 5 times (HBM counter calibration)")
    ap.add_argument("--gr-only", action="store_true",
                    help="only g(r) of the 256 batched replicas, 3 times (for counter runs)")
    ap.add_argument("--sph-only", action="store_true",
                    help="only the two SPH dam-break legs (for kernel traces / counter runs)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only the FP32 / FP64 FMA and INT32 issue-rate probe, 3 launches each "
                         "(calibration of counter-based VALU utilisation)")
    ap.add_argument("--dem-input", metavar="PATH",
                    help="write the DEM settle workload for the dem_settle driver to PATH and "
                         "exit (no GPU; profiles run the driver on it directly)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {a.gpus} is not supported "
              f"(the multi-GPU decomposition is a design, not part of this build)",
              file=sys.stderr)
        sys.exit(2)

    if a.dem_input:
        w = workloads.dem_settle_config(20000 if a.quick else 200_000)
        write_dem_input(w, a.dem_input)
        log(f"wrote the DEM settle workload (N={w.n}) to {a.dem_input}")
        return

    if a.valu_probe:
        ctx = pb.Context(0)
        _, cus = ctx.info()
        for kind, name in ((0, "FP32 FMA"), (1, "FP64 FMA"), (2, "INT32 xor+add"),
                           (3, "INT32 cmp+cndmask+add")):
            ctx.valu_probe(kind, 256)                     # warm-up
            for _ in range(3):
                ms, wi = ctx.valu_probe(kind, 4096)
                rate = wi / (ms * 1e-3)
                log(f"valu probe {name}: {wi:.4g} wave instructions in "
                    f"{ms:.3f} ms = {rate:.4g}/s = {rate / (cus * 2.4e9):.3f} per CU-cycle at "
                    f"2.4 GHz")
        return

    if a.gr_only:
        ctx = pb.Context(0)
        reps = workloads.lj_replicas(256, 500, 20000, seed=5)
        sb = make_system(ctx, reps)
        log(f"g(r) only: {bench_gr(sb, reps)}")
        return

    if a.sph_only:
        ctx = pb.Context(0)
        log(f"sph only: {bench_sph(ctx, a)}")
        log(f"sph only (config physics): {bench_sph(ctx, a, config=True)}")
        return

    if a.copy_probe:
        import torch
        x = torch.empty(32 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        x.fill_(1.0)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copy probe: 5 x 256 MiB read + 256 MiB written")
        return

    ctx = pb.Context(0)
    name, cus = ctx.info()
    log(f"device {name} CUs={cus}")
    w = workloads.lj_fcc(16 if a.quick else 64, 0.8442, 0.72, seed=2)
    log(f"workload N={w.n} L={w.L:.4f}")

    # ---- mixed (chosen) form: tile shape first (the autotuner restores the
    # state it starts from), then the warm-up, then the timed steps, which
    # continue the warm-up's list and cell order
    s = make_system(ctx, w)
    tiles = choose_tiles(s, a, pb.PB_MIXED)
    log(f"tiles: {tiles}")
    timed_run(ctx, s, a.warmup, pb.PB_MIXED, a.skin, a.rebuild_every, 0, timing=False)
    el, e = timed_run(ctx, s, a.steps, pb.PB_MIXED, a.skin, a.rebuild_every, 10)
    tm = s.lj_run_timing()
    stats = s.lj_run_stats()
    value = w.n * a.steps / el
    kern_ms = tm["step_ms"] / max(tm["nstep"], 1)
    achieved = ALG_BYTES * w.n / (kern_ms * 1e-3) / 1e9
    if a.profile:
        return
    # energy drift of the same system: 100 more steps with E recorded every step
    _, e1 = timed_run(ctx, s, 100, pb.PB_MIXED, a.skin, a.rebuild_every, 1, timing=False)
    drift = drift_per_1e4(e1, w.n, 1)
    log(f"mixed: {value:.4g} atom-steps/s, {1e3 * el / a.steps:.3f} ms/step, "
        f"step kernel {kern_ms * 1e3:.1f} us, rebuild {tm['rebuild_ms'] / max(tm['nrebuild'], 1):.3f} ms,"
        f" {stats}")
    # force error of the mixed pair kernel against FP64 on the final state
    s.lj_force(pb.PB_FP64, energies=False)
    f64 = s.get("F")
    s.lj_force(pb.PB_MIXED, energies=False)
    f32 = s.get("F")
    frms = float(np.sqrt((f64 ** 2).sum(axis=1).mean()))
    ferr = float(np.abs(f32 - f64).max() / frms)
    # the same number of steps from host buffers to host buffers: import
    # (host -> HBM), the run (its list and cell order built from scratch) and
    # positions and velocities back to the host, all timed together
    ctx.sync()
    t0 = time.perf_counter()
    s.import_particles(w.r, w.m, v=w.v, type=w.type)
    ctx.sync()
    t_in = time.perf_counter() - t0
    s.lj_run(DT, a.steps, precision=pb.PB_MIXED, skin=a.skin, rebuild_every=a.rebuild_every,
             energy_every=0)
    ctx.sync()
    t_run = time.perf_counter() - t0 - t_in
    t1 = time.perf_counter()
    s.get("r")
    s.get("v")
    t_out = time.perf_counter() - t1
    host_io = {"value": w.n * a.steps / (t_in + t_run + t_out), "unit": "atom-steps/s",
               "import_ms": 1e3 * t_in, "run_ms": 1e3 * t_run, "get_r_v_ms": 1e3 * t_out,
               "steps": a.steps}
    log(f"host-buffer inclusive: {host_io}")
    del s

    # ---- plain FP64 form
    s = make_system(ctx, w)
    tiles64 = choose_tiles(s, a, pb.PB_FP64)
    timed_run(ctx, s, a.warmup, pb.PB_FP64, a.skin, a.rebuild_every, 0, timing=False)
    el64, e64 = timed_run(ctx, s, a.steps, pb.PB_FP64, a.skin, a.rebuild_every, 10)
    tm64 = s.lj_run_timing()
    value64 = w.n * a.steps / el64
    kern64 = tm64["step_ms"] / max(tm64["nstep"], 1)
    _, e64 = timed_run(ctx, s, 100, pb.PB_FP64, a.skin, a.rebuild_every, 1, timing=False)
    log(f"fp64: {value64:.4g} atom-steps/s, step kernel {kern64 * 1e3:.1f} us")
    del s

    traffic, traffic_src = load_traffic()
    out = {
        "metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair forces / f64 integration+reductions", "data": "synthetic",
        "config": {"workload": f"LJ fluid N={w.n} (FCC {round((w.n / 4) ** (1 / 3))}^3x4), "
                               f"rho 0.8442, T 0.72, rc 2.5, dt 0.005, Verlet skin {a.skin}, "
                               f"re-sort every {a.rebuild_every} steps",
                   "global_batch": w.n, "seq_len": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_lj_step (mixed, fused force+VV)",
                     "kernel_us": kern_ms * 1e3, "alg_bytes_per_atom": ALG_BYTES},
        "plain_fp64": {"value": value64, "ms_per_step": 1e3 * el64 / a.steps, "tiles": tiles64,
                       "kernel_us": kern64 * 1e3,
                       "roofline_frac": ALG_BYTES * w.n / (kern64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "energy_drift_per_1e4_steps": drift_per_1e4(e64, w.n, 1),
                       "drift_n4000_1e4_steps": None if a.no_drift else
                       drift_direct(ctx, pb.PB_FP64, a.skin, a.rebuild_every)},
        "mixed": {"force_err_max_over_frms": ferr, "energy_drift_per_1e4_steps": drift,
                  "energy_drift_run": "100 steps after the timed ones, E recorded every step, "
                                      "fitted slope of E/N",
                  "drift_n4000_1e4_steps": None if a.no_drift else
                  drift_direct(ctx, pb.PB_MIXED, a.skin, a.rebuild_every),
                  "tiles": tiles,
                  "max_nn": stats["max_nn"], "skin_violation": stats["skin_violation"],
                  "rebuild_ms_avg": tm["rebuild_ms"] / max(tm["nrebuild"], 1),
                  "host_buffers_inclusive": host_io},
    }

    # ---- batched replicas vs plain per-replica FP64
    if not a.no_batch:
        reps = workloads.lj_replicas(16 if a.quick else 256, 500, 20000, seed=5)
        ntot = int(reps.counts.sum())
        sb = make_system(ctx, reps)
        tiles_b = choose_tiles(sb, a, pb.PB_MIXED)
        timed_run(ctx, sb, 10, pb.PB_MIXED, a.skin, a.rebuild_every, 0, timing=False)
        elb, eb = timed_run(ctx, sb, a.batch_steps, pb.PB_MIXED, a.skin, a.rebuild_every, 0)
        vb = ntot * a.batch_steps / elb
        nreb = sb.lj_run_stats()["rebuilds"]
        log(f"batched (eager, host-polled rebuilds): {len(reps.systems)} replicas, {ntot} atoms, "
            f"{vb:.4g} atom-steps/s, {nreb} rebuilds")
        # the hipGraph-captured form: periods of R steps + the rebuild,
        # captured once and replayed.  The interval is fixed, so it is taken
        # from the eager run's rebuild rate with a 30 % margin (the hottest
        # replicas set it; a skin violation would raise)
        rg = max(2, min(a.rebuild_every, int(0.7 * a.batch_steps / max(nreb, 1))))
        ctx.sync()
        t0 = time.perf_counter()
        sb.lj_run(DT, a.batch_steps, precision=pb.PB_MIXED, skin=a.skin,
                  rebuild_every=rg, energy_every=0, use_graph=True)
        ctx.sync()
        elg = time.perf_counter() - t0
        vg = ntot * a.batch_steps / elg
        log(f"batched (hipGraph replay): {vg:.4g} atom-steps/s")
        # the same with the persistent step kernel (per-XCD work queues over
        # the replica x tile segments) in the captured periods
        ctx.sync()
        t0 = time.perf_counter()
        sb.lj_run(DT, a.batch_steps, precision=pb.PB_MIXED, skin=a.skin,
                  rebuild_every=rg, energy_every=0, use_graph=True, persistent=True)
        ctx.sync()
        elp = time.perf_counter() - t0
        vgp = ntot * a.batch_steps / elp
        log(f"batched (hipGraph replay, persistent step kernel): {vgp:.4g} atom-steps/s")
        # per-replica mixed force error against the FP64 kernel
        sb.lj_force(pb.PB_FP64, energies=False)
        fb64 = sb.get("F")
        sb.lj_force(pb.PB_MIXED, energies=False)
        fb32 = sb.get("F")
        off = 0
        berr = []
        for nk in reps.counts:
            f6, f3 = fb64[off:off + nk], fb32[off:off + nk]
            berr.append(float(np.abs(f3 - f6).max() / np.sqrt((f6 ** 2).sum(1).mean())))
            off += nk
        gr = bench_gr(sb, reps)
        log(f"g(r): {gr['pairs']} pairs in {gr['ms']:.2f} ms, {gr['pairs_per_s']:.4g} pairs/s")
        del sb
        plain_systems = [make_system(ctx, x) for x in reps.systems]
        for ps in plain_systems:
            ps.lj_run(DT, 2, precision=pb.PB_FP64, skin=a.skin, rebuild_every=a.rebuild_every,
                      energy_every=0)
        ctx.sync()
        t0 = time.perf_counter()
        for ps in plain_systems:
            ps.lj_run(DT, a.plain_steps, precision=pb.PB_FP64, skin=a.skin,
                      rebuild_every=a.rebuild_every, energy_every=0)
        ctx.sync()
        elp = time.perf_counter() - t0
        vp = ntot * a.plain_steps / elp
        del plain_systems
        log(f"plain sequential FP64: {vp:.4g} atom-steps/s")
        vbest = max(vb, vg, vgp)
        out["batched"] = {"replicas": len(reps.systems), "atoms": ntot, "value": vbest,
                          "eager": vb, "graph": vg, "graph_persistent": vgp, "tiles": tiles_b,
                          "eager_rebuilds": nreb, "graph_rebuild_every": rg,
                          "force_err_max_over_frms": max(berr),
                          "force_err_median_over_frms": float(np.median(berr)),
                          "force_err_reference": "GPU FP64 kernel, per replica F_rms",
                          "plain_sequential_fp64": vp, "speedup": vbest / vp,
                          "steps": a.batch_steps, "plain_steps": a.plain_steps, "gr": gr}

    if not a.no_sph:
        out["sph"] = bench_sph(ctx, a)
        out["sph_config"] = bench_sph(ctx, a, config=True)
    if not a.no_dem:
        out["dem"] = bench_dem(a)
        out["dem_hertz"] = bench_dem(a, law="hertz")

    out["host"] = host_info()
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(w, budget_s=5.0 if a.quick else 20.0)
        out["cpu_baseline"]["host"] = out["host"]["cpu_model"]
        c1 = cpu_baseline(w, budget_s=3.0 if a.quick else 8.0, cores=1)
        out["cpu_baseline"]["single_core"] = {"value": c1["value"], "sample": c1["sample"]}
        log(f"cpu: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
