"""Benchmark of the D3Q15 collide-stream step on one MI355X.

Default workload: 256^3 periodic domain decomposed into 64 boxes of 64^3,
FP64, tau 0.6, rho = 1 + 1e-3 N(0,1), u_a ~ U(-0.01, 0.01) (seed 2024).
A step is one lattice-Boltzmann time step over the whole domain.  The timed
region is one Iterate(K) call (K steps) with the state resident in HBM,
bracketed by stream synchronisations.  Before it, untimed: a clock warm-up
(steps of the same simulation for at least --clock-warmup-ms of wall time, so
that the GPU's clocks have left their idle state: after idling, the first
~25 launches of the fused kernel run up to 9 % slower; the FP32-shifted run
takes the same number of warm-up steps as the FP64 one it is compared
with), then the W warm-up steps.

Prints one JSON line.  `roofline` is computed for the fused stream-collide
kernel from its average duration, measured with HIP events on the library's
own stream; algorithmic bytes are 240 B per lattice update (15 x 8 B read +
15 x 8 B written) in FP64 and 120 B in FP32-shifted storage.
"""
import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from lambrex_amd import lbx, workloads  # noqa: E402

METRIC = "MLUPS & % HBM roofline (FP64 vs FP32-shifted); p50/p99 latency of batched runs"
HBM_PEAK_GBS = 8000.0
BYTES_PER_LU = {lbx.LBX_FP64: 240, lbx.LBX_FP32_SHIFTED: 120}
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"
COPY_BYTES = 2 << 30  # 2 GiB read + 2 GiB written per copy dispatch


CLOCK_WARMUP_MS = 200.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def clock_warmup(ctx, sim, ms, chunk=10, steps=None):
    """Untimed steps of `sim` until `ms` of wall time have passed (at least one
    chunk), so the GPU's clocks leave their idle state before anything is
    timed; or exactly `steps` steps, so that a run compared with an earlier
    one reaches the same simulated time.  Returns the steps taken."""
    if steps is not None:
        sim.iterate(steps)
        ctx.sync()
        return steps
    done = 0
    t0 = time.perf_counter()
    while True:
        sim.iterate(chunk)
        ctx.sync()
        done += chunk
        if (time.perf_counter() - t0) * 1e3 >= ms:
            return done


def measure(ctx, dims, mgs, tau, rho, u, precision, steps, warmup, keep_fields=False, periodic=(1, 1, 1),
            clock_steps=None):
    """Iterate(steps) after the clock warm-up and Iterate(warmup), timed; with
    keep_fields the final density and velocity come back as result["fields"].
    A 0 in periodic puts halfway bounce-back walls on both faces normal to
    that axis.  clock_steps fixes the clock warm-up's step count (to compare
    with an earlier run); result["clock_steps"] is the count taken."""
    boxes = workloads.chop(dims, mgs)
    fields = None
    sim = lbx.Sim(ctx, dims, precision=precision, rho0=1.0, periodic=periodic)
    try:
        sim.define_level(0, boxes)
        sim.set_tau(0, tau)
        sim.init_from_fields(rho, u)
        c0 = sim.conservation()
        clock_steps = clock_warmup(ctx, sim, CLOCK_WARMUP_MS, steps=clock_steps)
        if warmup:
            sim.iterate(warmup)
        ctx.sync()
        ctx.profile_enable(True)
        t0 = time.perf_counter()
        sim.iterate(steps)
        ctx.sync()
        t1 = time.perf_counter()
        kern_ms, launches = ctx.profile_read()
        ctx.profile_enable(False)
        c1 = sim.conservation()
        if keep_fields:
            fields = sim.get_hydro(0)
    finally:
        sim.close()
    cells = dims[0] * dims[1] * dims[2]
    wall = t1 - t0
    avg_ms = kern_ms / launches if launches else float("nan")
    achieved = BYTES_PER_LU[precision] * cells / (avg_ms * 1e-3) / 1e9
    return {
        "boxes": len(boxes),
        "mlups": cells * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_avg_ms": avg_ms,
        "kernel_launches": launches,
        "kernel_gbs": achieved,
        "mass_drift_rel": abs(c1[0] - c0[0]) / abs(c0[0]),
        "fields": fields,
        "clock_steps": clock_steps,
    }


def traffic_for(kernel, workload):
    """HBM bytes per dispatch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    try:
        data = json.loads(TRAFFIC_FILE.read_text())
        return data[kernel][workload]["bytes_per_dispatch"]
    except (OSError, KeyError, ValueError):
        return None


def step_traffic(key, workload):
    """Measured bytes per level-0 step of a hierarchy (every kernel of the
    step, FETCH_SIZE + WRITE_SIZE scaled) against its algorithmic 240 B per
    lattice update, from the committed counter summary, or None."""
    try:
        d = json.loads(TRAFFIC_FILE.read_text())[key][workload]
        return {"bytes_per_step": d["bytes_per_step"], "algorithmic_bytes_per_step": d["algorithmic_bytes"],
                "over_algorithmic": d["traffic_over_algorithmic"], "source": d["source"]}
    except (OSError, KeyError, ValueError):
        return None


def hierarchy_roofline(mlups, ms_per_step, precision, traffic):
    """Where a hierarchy step sits against the 8 TB/s HBM peak: its
    algorithmic bytes (240 B or 120 B per lattice update) per second, and,
    for FP64, the measured bytes per step (every kernel's FETCH_SIZE +
    WRITE_SIZE, which also count lines served by the MALL) per second."""
    bpl = 240 if precision == lbx.LBX_FP64 else 120
    alg = mlups * 1e6 * bpl / 1e9
    out = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_gbs": round(alg, 1),
           "algorithmic_frac": round(alg / HBM_PEAK_GBS, 4)}
    if traffic:
        fab = traffic["bytes_per_step"] / (ms_per_step * 1e-3) / 1e9
        out.update({"measured_gbs": round(fab, 1), "measured_frac": round(fab / HBM_PEAK_GBS, 4)})
    return out


def cpu_baseline(dims, mgs, tau, rho, u, steps):
    """The CPU oracle (the reference's algorithm restated in C, serial, same
    FillPatch / collide / FillBoundary / stream structure) on the same
    workload, timed for a bounded number of steps."""
    from oracle import oracle as orc
    boxes = workloads.chop(dims, mgs)
    s = orc.Sim(*dims, tau, boxes=[(list(a), list(b)) for a, b in boxes])
    s.init(np.ascontiguousarray(rho.transpose(2, 1, 0)).reshape(-1),
           np.ascontiguousarray(u.transpose(3, 2, 1, 0)).reshape(-1))
    t0 = time.perf_counter()
    s.iterate(steps)
    t1 = time.perf_counter()
    s.close()
    cells = dims[0] * dims[1] * dims[2]
    return cells * steps / (t1 - t0) / 1e6, t1 - t0


def serving_trace(n_requests, seed, rate_hz):
    """Requests of the serving workload: cubic edge from {32,48,64,96,128},
    tau ~ U(0.55, 1.5), plane-pulse amplitude log-uniform in [1e-4, 1e-2] on
    plane x = N/2, steps ~ U{100..2000}, Poisson arrivals at rate_hz."""
    rng = np.random.default_rng(seed)
    edges = rng.choice([32, 48, 64, 96, 128], size=n_requests)
    taus = rng.uniform(0.55, 1.5, size=n_requests)
    amps = 10.0 ** rng.uniform(-4.0, -2.0, size=n_requests)
    steps = rng.integers(100, 2001, size=n_requests)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, size=n_requests))
    arrivals -= arrivals[0]
    return [dict(edge=int(e), tau=float(t), amp=float(a), steps=int(s), arrival=float(at))
            for e, t, a, s, at in zip(edges, taus, amps, steps, arrivals)]


def serving_precision_sample(ctx, trace, n):
    """The first n requests of the trace through an FP64 and an FP32-shifted
    server: max relative density error, and max |du| relative to each run's
    pulse amplitude (the velocity scale the pulse starts; a run's own final
    max |u| can have decayed to round-off after up to 2000 steps)."""
    outs = {}
    for prec in (lbx.LBX_FP64, lbx.LBX_FP32_SHIFTED):
        srv = lbx.Server(ctx, precision=prec, arena_bytes=8 << 30, max_active=44)
        got = [srv.submit((r["edge"],) * 3, r["steps"], r["tau"], amplitude=r["amp"], axis=0, want_rho=True)[1]
               for r in trace[:n]]
        srv.drain()
        srv.poll()
        srv.close()
        outs[prec] = got
    e_rho = max(float(np.max(np.abs(b["rho"] - a["rho"]) / a["rho"]))
                for a, b in zip(outs[lbx.LBX_FP64], outs[lbx.LBX_FP32_SHIFTED]))
    e_u = max(float(np.max(np.abs(b["u"] - a["u"]))) / r["amp"]
              for a, b, r in zip(outs[lbx.LBX_FP64], outs[lbx.LBX_FP32_SHIFTED], trace[:n]))
    return e_rho, e_u


def serve(ctx, trace, max_active, policy=lbx.LBX_SERVE_FIFO, max_wait_s=0.0):
    """Replay the trace in real time against one server; returns aggregate
    MLUPS over the makespan and completion latencies (finish - arrival)."""
    srv = lbx.Server(ctx, precision=lbx.LBX_FP32_SHIFTED, arena_bytes=16 << 30, max_active=max_active,
                     policy=policy, max_wait_s=max_wait_s)
    results = []
    submitted = 0
    arrival_of = {}
    t0 = time.perf_counter()
    clock0 = None
    while len(results) < len(trace):
        now = time.perf_counter() - t0
        while submitted < len(trace) and trace[submitted]["arrival"] <= now:
            r = trace[submitted]
            rid, _ = srv.submit((r["edge"],) * 3, r["steps"], r["tau"], amplitude=r["amp"], axis=0)
            arrival_of[rid] = r["arrival"]
            submitted += 1
        srv.progress(2)
        for res in srv.poll():
            results.append(res)
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # completion latency against the scheduled arrival (host clock of the server)
    lat = []
    for res in results:
        if clock0 is None:
            clock0 = res["t_submit"] - arrival_of[res["id"]]
        lat.append(res["t_done"] - (clock0 + arrival_of[res["id"]]))
    lu = sum(r["cells"] * r["steps"] for r in results)
    return {"mlups": lu / makespan / 1e6, "makespan_s": makespan, "lattice_updates": lu,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "steps_launched": st["steps_launched"]}


def refined_mix(ctx, max_active):
    """A burst of 24 FP32-shifted requests at once: 16 single-level runs
    (edges 32-48, 2000-4000 steps) and 8 two-level runs (coarse 64^3 in
    boxes of 32^3, a fine slab over coarse x in [24, 40) and all y, z in
    boxes of 32^3, 300-600 coarse steps, plane pulse at coarse x = 32), served
    by one server; aggregate MLUPS over the makespan counts both levels'
    updates."""
    rng = np.random.default_rng(11)
    fine = [((48 + a[0], a[1], a[2]), (48 + b[0], b[1], b[2])) for a, b in workloads.chop((32, 128, 128), 32)]
    srv = lbx.Server(ctx, precision=lbx.LBX_FP32_SHIFTED, arena_bytes=8 << 30, max_active=max_active)
    t0 = time.perf_counter()
    for k in range(24):
        if k % 3 == 2:
            srv.submit((64, 64, 64), int(rng.integers(300, 601)), float(rng.uniform(0.6, 1.2)), amplitude=1e-3,
                       boxes=workloads.chop((64, 64, 64), 32), fine_boxes=fine, plane=32)
        else:
            e = int(rng.choice([32, 48]))
            srv.submit((e, e, e), int(rng.integers(2000, 4001)), float(rng.uniform(0.6, 1.2)), amplitude=1e-3)
    srv.drain()
    res = srv.poll()
    makespan = time.perf_counter() - t0
    lu = srv.stats()["lattice_updates"]
    srv.close()
    assert len(res) == 24 and all(r["status"] == lbx.LBX_OK for r in res)
    lat = [r["t_done"] - r["t_submit"] for r in res]
    return {"mlups": lu / makespan / 1e6, "makespan_s": makespan, "lattice_updates": lu,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3)}


def two_level_geometry():
    """Two-level configuration: coarse 128^3 periodic (boxes of 32^3), fine
    level over coarse x in [48, 80) and all y, z (fine 64x256x256 in boxes of
    32^3), tau_coarse 0.8 (fine tau 1.1, src/AmrSim.cpp:313-314), plane pulse
    rho = 1.01 at coarse x = 64.  The coarse cells under the fine level are
    masked (the fine-BoxArray mask; the reference's own MakeFineMask masks the
    coarse low octant instead, src/AmrSim.cpp:419-428, which zeroes coarse
    cells nowhere near the fine level every step).  A step is one coarse
    step: 128^3 + 2 x 64x256x256 lattice updates."""
    dims = (128, 128, 128)
    cboxes = workloads.chop(dims, 32)
    ext = (64, 256, 256)
    fboxes = [((96 + a[0], a[1], a[2]), (96 + b[0], b[1], b[2])) for a, b in workloads.chop(ext, 32)]
    mask = list(fboxes)
    rho, u = workloads.plane_pulse(dims, 0.01, axis=0, plane=64)
    lu_per_step = 128 ** 3 + 2 * 64 * 256 * 256
    return dims, cboxes, fboxes, mask, rho, u, lu_per_step


def two_level(ctx, steps, warmup, precision=lbx.LBX_FP64):
    """The two-level configuration on the GPU (one coarse step = one replay
    of the captured HIP graph).  Also returns the final density and velocity
    of both levels (for the FP32 vs FP64 comparison).  Its timed region
    (hundreds of coarse steps) is long enough that the GPU's clock ramp after
    idling does not matter, so there is no clock warm-up."""
    dims, cboxes, fboxes, mask, rho, u, lu_per_step = two_level_geometry()
    sim = lbx.Sim(ctx, dims, precision=precision, rho0=1.0)
    try:
        sim.define_level(0, cboxes)
        sim.set_tau(0, 0.8)
        sim.init_from_fields(rho, u)
        sim.define_level(1, fboxes)
        sim.set_tau(1, 2 * (0.8 - 0.5) + 0.5)
        sim.fill_from_coarse(1)
        sim.set_fine_mask(mask)
        sim.iterate(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        sim.iterate(steps)
        ctx.sync()
        t1 = time.perf_counter()
        fields = sim.get_hydro(0) + sim.get_hydro_bbox(1)[:2]
    finally:
        sim.close()
    lu = lu_per_step * steps
    name = "FP64" if precision == lbx.LBX_FP64 else "FP32-shifted"
    traffic = step_traffic("two_level_step", "128^3 + 64x256x256") if precision == lbx.LBX_FP64 else None
    return {"workload": "coarse 128^3 (64 boxes of 32^3) + fine 64x256x256 (128 boxes of 32^3), "
                        f"ratio 2, 2 fine substeps per coarse step, {name}",
            "coarse_steps": steps, "warmup_steps": warmup, "mlups": lu / (t1 - t0) / 1e6,
            "ms_per_coarse_step": (t1 - t0) * 1e3 / steps,
            "graph": "one coarse step captured as a HIP graph",
            "traffic_per_step": traffic,
            "roofline": hierarchy_roofline(lu / (t1 - t0) / 1e6, (t1 - t0) * 1e3 / steps, precision, traffic)}, fields


def three_level_geometry():
    """Three-level configuration: the two-level one with a level 2 nested
    inside level 1 as a slab too: level-1 x in [120, 136) and all y, z
    (level 2 32x512x512 in boxes of 32^3, tau 1.7), each level's cells under
    its child masked.  The interfaces are planes, as in the two-level case:
    the reference's cycle adds the average of each fine box grown by one
    coarse cell into unmasked coarse cells (SumFromFine, src/AmrSim.cpp:
    592-602), which a level-2 box enclosed on all six sides turns into a
    divergence within about 100 level-0 steps (in FP64 too).  A step is one
    level-0 step: 128^3 + 2 x 64x256x256 + 4 x 32x512x512 lattice updates
    (level l takes 2^l steps)."""
    dims, cboxes, fboxes, mask, rho, u, _ = two_level_geometry()
    ext = (32, 512, 512)
    l2 = [((240 + a[0], a[1], a[2]), (240 + b[0], b[1], b[2])) for a, b in workloads.chop(ext, 32)]
    levels = [cboxes, fboxes, l2]
    masks = [fboxes, l2]
    lu_per_step = 128 ** 3 + 2 * 64 * 256 * 256 + 4 * 32 * 512 * 512
    return dims, levels, masks, rho, u, lu_per_step


def three_level(ctx, steps, warmup, precision=lbx.LBX_FP64):
    """The three-level configuration on the GPU (one level-0 step = one
    replay of the captured HIP graph)."""
    dims, levels, masks, rho, u, lu_per_step = three_level_geometry()
    sim = lbx.Sim(ctx, dims, precision=precision, rho0=1.0)
    try:
        sim.define_level(0, levels[0])
        tau = 0.8
        sim.set_tau(0, tau)
        sim.init_from_fields(rho, u)
        for l in (1, 2):
            tau = 2 * (tau - 0.5) + 0.5
            sim.define_level(l, levels[l])
            sim.set_tau(l, tau)
            sim.fill_from_coarse(l)
            sim.set_level_mask(l - 1, masks[l - 1])
        sim.iterate(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        sim.iterate(steps)
        ctx.sync()
        t1 = time.perf_counter()
        rhos = [sim.get_hydro(0)[0]] + [sim.get_hydro_bbox(l)[0] for l in (1, 2)]
    finally:
        sim.close()
    lu = lu_per_step * steps
    name = "FP64" if precision == lbx.LBX_FP64 else "FP32-shifted"
    traffic = (step_traffic("three_level_step", "128^3 + 64x256x256 + 32x512x512")
               if precision == lbx.LBX_FP64 else None)
    return {"workload": "coarse 128^3 (64 boxes of 32^3) + level 1 64x256x256 (128 boxes of 32^3) + level 2 "
                        f"32x512x512 (256 boxes of 32^3), ratio 2, {name}",
            "coarse_steps": steps, "warmup_steps": warmup, "mlups": lu / (t1 - t0) / 1e6,
            "ms_per_coarse_step": (t1 - t0) * 1e3 / steps,
            "graph": "one level-0 step captured as a HIP graph",
            "traffic_per_step": traffic,
            "roofline": hierarchy_roofline(lu / (t1 - t0) / 1e6, (t1 - t0) * 1e3 / steps, precision, traffic),
            "final_rho_range_per_level": [[float(np.min(r)), float(np.max(r))] for r in rhos]}, rhos


def cpu_three_level(steps):
    """The oracle's recursive Rohde cycle on the three-level configuration."""
    from oracle import oracle as orc
    dims, levels, masks, rho, u, lu_per_step = three_level_geometry()
    s = orc.Sim(*dims, 0.8, boxes=[(list(a), list(b)) for a, b in levels[0]])
    s.init(np.ascontiguousarray(rho.transpose(2, 1, 0)).reshape(-1),
           np.ascontiguousarray(u.transpose(3, 2, 1, 0)).reshape(-1))
    for l in (1, 2):
        s.add_fine([(list(a), list(b)) for a, b in levels[l]], [(list(a), list(b)) for a, b in masks[l - 1]])
    t0 = time.perf_counter()
    s.iterate(steps)
    t1 = time.perf_counter()
    s.close()
    return lu_per_step * steps / (t1 - t0) / 1e6, t1 - t0


def precision_errors(ref, test):
    """Max relative density error and max |du| / max |u|, over both levels."""
    r0, u0, r1, u1 = ref
    s0, v0, s1, v1 = test
    e_rho = max(float(np.max(np.abs(s0 - r0) / r0)), float(np.max(np.abs(s1 - r1) / r1)))
    e_u = max(float(np.max(np.abs(v0 - u0)) / np.max(np.abs(u0))),
              float(np.max(np.abs(v1 - u1)) / np.max(np.abs(u1))))
    return e_rho, e_u


def cpu_two_level(steps):
    """The oracle's two-level Rohde cycle (serial C, per-FAB ghosted storage
    like the reference) on the same two-level configuration."""
    from oracle import oracle as orc
    dims, cboxes, fboxes, mask, rho, u, lu_per_step = two_level_geometry()
    s = orc.Sim(*dims, 0.8, boxes=[(list(a), list(b)) for a, b in cboxes])
    s.init(np.ascontiguousarray(rho.transpose(2, 1, 0)).reshape(-1),
           np.ascontiguousarray(u.transpose(3, 2, 1, 0)).reshape(-1))
    s.add_fine([(list(a), list(b)) for a, b in fboxes], [(list(a), list(b)) for a, b in mask])
    t0 = time.perf_counter()
    s.iterate(steps)
    t1 = time.perf_counter()
    s.close()
    return lu_per_step * steps / (t1 - t0) / 1e6, t1 - t0


def finite_or_none(x):
    """NaN and infinities (e.g. a kernel average when no fused pass ran)
    as JSON null, so the output line stays valid JSON."""
    if isinstance(x, float):
        return x if np.isfinite(x) else None
    if isinstance(x, dict):
        return {k: finite_or_none(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [finite_or_none(v) for v in x]
    return x


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    global CLOCK_WARMUP_MS
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="cubic domain edge")
    ap.add_argument("--max-grid-size", type=int, default=64)
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-shifted measurement")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-copy", action="store_true", help="skip the copy-bandwidth ceiling")
    ap.add_argument("--autotune", action="store_true",
                    help="time every workgroup shape first and use the fastest (default: the default shape)")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the copy kernel (counter calibration)")
    ap.add_argument("--clock-warmup-ms", type=float, default=CLOCK_WARMUP_MS,
                    help="untimed steps before the W warm-up steps of each single-level measurement, for at "
                         "least this much wall time (lets the GPU clocks leave their idle state)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-two-level-steps", type=int, default=1,
                    help="coarse steps of the two-level CPU baseline (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=512,
                    help="requests of the serving measurement (0 skips it)")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--two-level-steps", type=int, default=500,
                    help="coarse steps of the two-level measurement (0 skips it)")
    ap.add_argument("--three-level-steps", type=int, default=200,
                    help="level-0 steps of the three-level measurement (0 skips it)")
    ap.add_argument("--cpu-three-level-steps", type=int, default=1,
                    help="level-0 steps of the three-level CPU baseline (0 skips it)")
    ap.add_argument("--walls-steps", type=int, default=200,
                    help="steps of the FP64 measurement with bounce-back walls on both z faces (0 skips it)")
    args = ap.parse_args()
    CLOCK_WARMUP_MS = args.clock_warmup_ms

    if args.copy_only:
        with lbx.Context(0) as ctx:
            gbs = ctx.bench_copy(COPY_BYTES, reps=3)
        print(json.dumps({"copy_bytes_read_per_dispatch": COPY_BYTES,
                          "copy_bytes_written_per_dispatch": COPY_BYTES, "copy_gbs": round(gbs, 1)}))
        return

    if args.serve_only:
        trace = serving_trace(args.serve_requests, seed=7, rate_hz=200.0)
        with lbx.Context(0) as ctx:
            batched = serve(ctx, trace, max_active=44)
            seq = serve(ctx, trace, max_active=1)
        print(json.dumps({"serving": {"requests": len(trace), "batched": batched, "sequential": seq,
                                      "speedup_aggregate_mlups": round(batched["mlups"] / seq["mlups"], 3)}}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the domain would shard by boxes with one halo exchange per step).",
              file=sys.stderr)
        sys.exit(2)

    n = args.n
    dims = (n, n, n)
    tau = 0.6
    seed = 2024
    log(f"[bench] generating {n}^3 state (seed {seed})")
    rho, u = workloads.random_state(dims, seed=seed, rho_sigma=1e-3, u_amp=0.01)

    with lbx.Context(0) as ctx:
        ctx.set_autotune(args.autotune)
        copy_gbs = None
        if not args.no_copy:
            copy_gbs = ctx.bench_copy(COPY_BYTES, reps=10)
            log(f"[bench] copy ceiling: {copy_gbs:.0f} GB/s (2 GiB read + 2 GiB written per pass)")
        log(f"[bench] FP64 {n}^3 boxes of {args.max_grid_size}^3: warmup {args.warmup}, steps {args.steps}")
        r64 = measure(ctx, dims, args.max_grid_size, tau, rho, u, lbx.LBX_FP64, args.steps, args.warmup,
                      keep_fields=not args.no_fp32)
        log(f"[bench] FP64: {r64['mlups']:.0f} MLUPS, kernel {r64['kernel_avg_ms']:.4f} ms "
            f"({r64['kernel_gbs']:.0f} GB/s)")
        shape64 = ctx.tuned_shape(n, n, lbx.LBX_FP64)[0]
        walls = None
        if args.walls_steps:
            log(f"[bench] FP64 {n}^3 with bounce-back walls on both z faces: {args.walls_steps} steps")
            rw = measure(ctx, dims, args.max_grid_size, tau, rho, u, lbx.LBX_FP64, args.walls_steps, 5,
                         periodic=(1, 1, 0))
            rw.pop("fields")
            walls = {"workload": f"{n}^3, {rw['boxes']} boxes of {args.max_grid_size}^3, tau {tau}, halfway "
                                 "bounce-back walls on both z faces (x, y periodic)",
                     "periodic": [1, 1, 0], "steps": args.walls_steps, "mlups": rw["mlups"],
                     "ms_per_step": rw["ms_per_step"], "kernel": "k_step_f64<STEP_FUSED, walls>",
                     "kernel_avg_ms": rw["kernel_avg_ms"], "kernel_gbs": rw["kernel_gbs"],
                     "frac_of_hbm_peak": rw["kernel_gbs"] / HBM_PEAK_GBS,
                     "frac_of_copy_ceiling": rw["kernel_gbs"] / copy_gbs if copy_gbs else None,
                     "mass_drift_rel": rw["mass_drift_rel"]}
            log(f"[bench] walls: {rw['mlups']:.0f} MLUPS, kernel {rw['kernel_avg_ms']:.4f} ms "
                f"({rw['kernel_gbs']:.0f} GB/s), mass drift {rw['mass_drift_rel']:.1e}")
        twolev = None
        if args.two_level_steps:
            log(f"[bench] two-level: {args.two_level_steps} coarse steps")
            twolev, f64_fields = two_level(ctx, args.two_level_steps, 5)
            rf = twolev["roofline"]
            log(f"[bench] two-level: {twolev['mlups']:.0f} MLUPS, {twolev['ms_per_coarse_step']:.3f} ms per coarse step, "
                f"{rf['algorithmic_gbs']:.0f} GB/s algorithmic, {rf.get('measured_gbs', float('nan')):.0f} GB/s measured")
            if not args.no_fp32:
                tl32, f32_fields = two_level(ctx, args.two_level_steps, 5, lbx.LBX_FP32_SHIFTED)
                e_rho, e_u = precision_errors(f64_fields, f32_fields)
                tl32["max_rel_rho_err_vs_fp64"] = e_rho
                tl32["max_du_over_max_u_vs_fp64"] = e_u
                tl32["speedup_vs_fp64"] = round(tl32["mlups"] / twolev["mlups"], 3)
                twolev["fp32_shifted"] = tl32
                log(f"[bench] two-level FP32-shifted: {tl32['mlups']:.0f} MLUPS, max rel rho err {e_rho:.2e}, "
                    f"max |du|/max|u| {e_u:.2e}")
            del f64_fields
        threelev = None
        if args.three_level_steps:
            log(f"[bench] three-level: {args.three_level_steps} level-0 steps")
            threelev, r64s = three_level(ctx, args.three_level_steps, 5)
            log(f"[bench] three-level: {threelev['mlups']:.0f} MLUPS, "
                f"{threelev['ms_per_coarse_step']:.3f} ms per level-0 step, "
                f"rho ranges {threelev['final_rho_range_per_level']}")
            if not args.no_fp32:
                t32, r32s = three_level(ctx, args.three_level_steps, 5, lbx.LBX_FP32_SHIFTED)
                e_rho = max(float(np.max(np.abs(b - a) / a)) for a, b in zip(r64s, r32s))
                threelev["fp32_shifted"] = {"mlups": t32["mlups"], "ms_per_coarse_step": t32["ms_per_coarse_step"],
                                            "roofline": t32["roofline"],
                                            "speedup_vs_fp64": round(t32["mlups"] / threelev["mlups"], 3),
                                            "max_rel_rho_err_vs_fp64": e_rho}
                log(f"[bench] three-level FP32-shifted: {t32['mlups']:.0f} MLUPS, max rel rho err {e_rho:.2e}")
            del r64s
        serving = None
        if args.serve_requests:
            trace = serving_trace(args.serve_requests, seed=7, rate_hz=200.0)
            log(f"[bench] serving: {len(trace)} requests, Poisson 200/s, batched (<= 44 active)")
            batched = serve(ctx, trace, max_active=44)
            log(f"[bench] batched: {batched['mlups']:.0f} MLUPS, p50 {batched['p50_ms']:.0f} ms, "
                f"p99 {batched['p99_ms']:.0f} ms, makespan {batched['makespan_s']:.1f} s")
            log("[bench] serving: same trace, batched, shortest-first admission (queued > 2 s first)")
            sjf = serve(ctx, trace, max_active=44, policy=lbx.LBX_SERVE_SHORTEST_FIRST, max_wait_s=2.0)
            sjf["max_wait_s"] = 2.0
            log(f"[bench] shortest-first: {sjf['mlups']:.0f} MLUPS, p50 {sjf['p50_ms']:.0f} ms, "
                f"p99 {sjf['p99_ms']:.0f} ms, makespan {sjf['makespan_s']:.1f} s")
            log("[bench] serving: same trace, one run at a time")
            seq = serve(ctx, trace, max_active=1)
            log(f"[bench] sequential: {seq['mlups']:.0f} MLUPS, p50 {seq['p50_ms']:.0f} ms, "
                f"p99 {seq['p99_ms']:.0f} ms, makespan {seq['makespan_s']:.1f} s")
            log("[bench] serving: 24-request burst with 8 two-level runs, batched and one at a time")
            mix_b = refined_mix(ctx, 44)
            mix_s = refined_mix(ctx, 1)
            log(f"[bench] refined mix: batched {mix_b['mlups']:.0f} MLUPS (makespan {mix_b['makespan_s']:.2f} s), "
                f"one at a time {mix_s['mlups']:.0f} MLUPS ({mix_s['makespan_s']:.2f} s)")
            e_rho, e_u = serving_precision_sample(ctx, trace, 16)
            log(f"[bench] serving FP32 vs FP64 on the first 16 requests: max rel rho err {e_rho:.2e}, "
                f"max |du|/amplitude {e_u:.2e}")
            serving = {"requests": len(trace), "arrival": "Poisson 200/s (seed 7)", "precision": "FP32-shifted",
                       "batched": batched, "batched_shortest_first": sjf, "sequential": seq,
                       "speedup_aggregate_mlups": round(batched["mlups"] / seq["mlups"], 3),
                       "refined_mix": {"workload": "24 FP32-shifted requests at once: 16 single-level (32^3-48^3, "
                                                   "2000-4000 steps) + 8 two-level (64^3 + fine slab 32x128x128, "
                                                   "300-600 coarse steps)",
                                       "batched": mix_b, "one_at_a_time": mix_s,
                                       "speedup_aggregate_mlups": round(mix_b["mlups"] / mix_s["mlups"], 3)},
                       "fp32_vs_fp64_first_16": {"max_rel_rho_err": e_rho, "max_du_over_amplitude": e_u}}
        r32 = None
        if not args.no_fp32:
            log("[bench] FP32-shifted, same workload")
            r32 = measure(ctx, dims, args.max_grid_size, tau, rho, u, lbx.LBX_FP32_SHIFTED,
                          args.steps, args.warmup, keep_fields=True, clock_steps=r64["clock_steps"])
            (r0, u0), (s0, v0) = r64.pop("fields"), r32.pop("fields")
            r32["max_rel_rho_err_vs_fp64"] = float(np.max(np.abs(s0 - r0) / r0))
            r32["max_du_over_max_u_vs_fp64"] = float(np.max(np.abs(v0 - u0)) / np.max(np.abs(u0)))
            del r0, u0, s0, v0
            log(f"[bench] FP32-shifted vs FP64 after {args.steps} steps: max rel rho err "
                f"{r32['max_rel_rho_err_vs_fp64']:.2e}, max |du|/max|u| {r32['max_du_over_max_u_vs_fp64']:.2e}")
            shape32 = ctx.tuned_shape(n, n, lbx.LBX_FP32_SHIFTED)[0]
            log(f"[bench] FP32-shifted: {r32['mlups']:.0f} MLUPS, kernel {r32['kernel_avg_ms']:.4f} ms "
                f"({r32['kernel_gbs']:.0f} GB/s)")

    workload = f"{n}^3 periodic, {r64['boxes']} boxes of {args.max_grid_size}^3, tau {tau}"
    kernel = "k_step_f64<STEP_FUSED>"
    traffic = traffic_for(kernel, f"{n}^3")
    cells = n ** 3
    out = {
        "metric": METRIC,
        "value": round(r64["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "clock_warmup_ms": args.clock_warmup_ms,
        "clock_warmup_steps": r64["clock_steps"],
        "ms_per_step": round(r64["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 / Box-Muller initial state, equilibrium populations)",
        "config": {"workload": workload, "cells": cells, "boxes": r64["boxes"], "tau": tau,
                   "seed": seed, "precision": "FP64",
                   "workgroup": {"bx": shape64[0], "by": shape64[1], "max_workgroups_per_cu": shape64[2],
                                 "autotuned": bool(args.autotune)}},
        "roofline": {
            "bound": "hbm",
            "achieved": round(r64["kernel_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r64["kernel_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": kernel,
            "kernel_avg_ms": round(r64["kernel_avg_ms"], 5),
            "algorithmic_bytes_per_launch": BYTES_PER_LU[lbx.LBX_FP64] * cells,
            "copy_ceiling_gbs": round(copy_gbs, 1) if copy_gbs else None,
            "frac_of_copy_ceiling": round(r64["kernel_gbs"] / copy_gbs, 4) if copy_gbs else None,
        },
        "mass_drift_rel": r64["mass_drift_rel"],
    }
    if r32:
        out["fp32_shifted"] = {
            "value": round(r32["mlups"], 1),
            "unit": "MLUPS",
            "ms_per_step": round(r32["ms_per_step"], 5),
            "speedup_vs_fp64": round(r32["mlups"] / r64["mlups"], 3),
            "roofline": {"bound": "hbm", "achieved": round(r32["kernel_gbs"], 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(r32["kernel_gbs"] / HBM_PEAK_GBS, 4),
                         "traffic": traffic_for("k_step_f32s<STEP_FUSED>", f"{n}^3"),
                         "kernel": "k_step_f32s<STEP_FUSED>",
                         "kernel_avg_ms": round(r32["kernel_avg_ms"], 5),
                         "algorithmic_bytes_per_launch": BYTES_PER_LU[lbx.LBX_FP32_SHIFTED] * cells},
            "workgroup": {"bx": shape32[0], "by": shape32[1], "max_workgroups_per_cu": shape32[2]},
            "max_rel_rho_err_vs_fp64": r32["max_rel_rho_err_vs_fp64"],
            "max_du_over_max_u_vs_fp64": r32["max_du_over_max_u_vs_fp64"],
        }
    if walls:
        out["walls"] = walls
    if twolev:
        out["two_level"] = twolev
    if threelev:
        out["three_level"] = threelev
    if serving:
        out["serving"] = serving
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle, 1 thread, {n}^3 as {r64['boxes']} boxes, {args.cpu_steps} step(s)")
        mlups, secs = cpu_baseline(dims, args.max_grid_size, tau, rho, u, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": round(mlups, 3),
            "unit": "MLUPS",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_steps} step(s) of the same {workload} FP64 workload "
                      f"({secs:.1f} s), serial C restatement of the reference; host CPU {cpu_model()}, "
                      f"{os.cpu_count()} logical cores present",
        }
        if twolev and args.cpu_two_level_steps:
            log(f"[bench] CPU baseline, two-level: oracle, 1 thread, {args.cpu_two_level_steps} coarse step(s)")
            mlups2, secs2 = cpu_two_level(args.cpu_two_level_steps)
            twolev["cpu_baseline"] = {
                "value": round(mlups2, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
                "sample": f"{args.cpu_two_level_steps} coarse step(s) of the same two-level workload "
                          f"({secs2:.1f} s), serial C restatement of the reference's Rohde cycle",
            }
        if threelev and args.cpu_three_level_steps:
            log(f"[bench] CPU baseline, three-level: oracle, 1 thread, {args.cpu_three_level_steps} level-0 step(s)")
            mlups3, secs3 = cpu_three_level(args.cpu_three_level_steps)
            threelev["cpu_baseline"] = {
                "value": round(mlups3, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
                "sample": f"{args.cpu_three_level_steps} level-0 step(s) of the same three-level workload "
                          f"({secs3:.1f} s), serial C restatement of the reference's recursive Rohde cycle",
            }
    print(json.dumps(finite_or_none(out), allow_nan=False), flush=True)


if __name__ == "__main__":
    main()
