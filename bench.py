#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the Godunov step on one MI355X.

Workload: the 2D gas + pressureless dust disk, nr = 1024 x nphi = 3072, FP32
storage (St 0.01, dust/gas 0.01, q = 3e-4, alpha = 1e-3, wave killing, 1e-4
seeded noise on Sigma and v_r), synthetic initial state from the analytic
profiles.  A step is one full DS step (viscosity, half kick, x sweeps,
azimuthal sweeps + orbital advection, predictor/corrector kick, wave
killing) with the CFL dt chosen on the device, exactly as the reference
time loop does it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  `value` is interior cell updates per second over the
K timed steps (state resident in HBM before timing starts; the device loop
replays its captured step).  `roofline` describes the dominant kernel:
algorithmic bytes per launch over its average duration from HIP events
around every launch of further uncaptured steps, against 8 TB/s; its `traffic` is the
HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
passes (profiles/).  `cpu_baseline` times the FP64 host oracle (OpenMP) on a
bounded sample of the same workload.  `batched` times 64 uneven disks
advanced together (one launch per pass, captured step, autotuned schedule)
against looping the single-disk path over the same disks.  `adjoint` times
the backward sweep of the gap-profile loss on the 512x1024 disk against the
plain forward run.  `sph` times the 3D adiabatic disk (384x1536x64).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# the CPU baseline's OpenMP threads stay on neighbouring cores (read when the
# oracle library loads)
os.environ.setdefault("OMP_PROC_BIND", "close")

PEAK_HBM = 8.0e12  # B/s, MI355X HBM3E

# algorithmic HBM bytes per interior cell per launch, in units of the storage
# word: (fields read + fields written) that the kernel must move at minimum
WORDS_PER_CELL = {
    "sweep_x": 17,               # gas and dust x sweeps in one launch (8 + 9 below)
    "ring": 12,                  # gas and dust azimuthal passes in one launch (6 + 6)
    "sweep_x_gas": 8,            # rho,u,vp in and out; + the viscous forces of the fused first half kick
    "ring_gas": 6,               # azimuthal sweep + orbital shift + residual advection
    "sweep_x_dust": 9,           # dust in and out; + gas rho,u,vp for the fused kick's drag
    "ring_dust": 6,
    "viscosity": 5,              # rho,u,vp in; two forces out (once per run)
    "predictor_viscosity": 7,    # gas 3 + forces 2 in; forces of the predicted state 2 out
    "corrector_viscosity": 14,   # gas 3 + forces 2 + dust u,vp in; gas 3 (wave killing) + dust 2 + forces 2 out
    "sources_corrector": 12,     # (without viscosity) gas 3 + dust 3 in; gas and dust velocities out
    "cfl": 4,                    # gas u, vp and dust u, vp in (c_s from the per-radius table)
    "prologue": 0,
}


def note(msg: str) -> None:
    """One progress line on stderr as each leg of the bench ends."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", type=int, default=4, choices=(4, 8))
    ap.add_argument("--nr", type=int, default=1024)
    ap.add_argument("--nphi", type=int, default=3072)
    ap.add_argument("--cpu-steps", type=int, default=20, help="oracle steps of the CPU baseline sample")
    ap.add_argument("--cpu-c1-steps", type=int, default=100, help="oracle steps of the 256x512 disk (CPU baseline)")
    ap.add_argument("--fp64-steps", type=int, default=20, help="steps of the FP64 leg (0 skips it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the uncaptured loop (one launch per pass) instead of the replayed hipGraph step")
    ap.add_argument("--kernel-steps", type=int, default=50,
                    help="uncaptured steps with HIP events around every launch, for the per-kernel durations")
    ap.add_argument("--batch-disks", type=int, default=64, help="disks of the batched leg (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the grid-barrier probe of the batched leg (its cooperative launch is not traced: "
                         "rocprofv3's kernel tracing ends the process in a host SIGSEGV there)")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="steps of the adjoint leg (0 skips it)")
    ap.add_argument("--sph-steps", type=int, default=20, help="steps of the 3D disk leg (0 skips it)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates the HBM byte counters)")
    return ap.parse_args()


def adjoint_leg(steps: int, precision: int) -> dict:
    """Gap-profile loss on the 512x1024 gas disk (q 1e-4, alpha 1e-3):
    forward `steps` with checkpoints every 20, backward through the adjoint;
    cost = backward wall time (recompute included) / plain forward wall time
    (pg_run with its captured step) over the same steps, both in the steady
    state of a loop that calls them repeatedly: the forward on a handle whose
    step graphs a warm-up run captured, the backward on its second call (the
    first also captures the segment graphs; its time is reported apart).
    The tape_* keys: the same with every step's intermediates kept in HBM
    (ckpt_every 0, no recompute)."""
    import numpy as np
    from penguin_amd import Disk, config_c5
    from penguin_amd.adjoint import gap_profile

    p = config_c5(precision=precision)
    target = gap_profile(p)
    d = Disk(p)
    d.run(3, use_graph=True)                     # warm up kernels and capture the step's graphs
    t0 = time.perf_counter()
    d.run(steps, use_graph=True)
    t_fwd = time.perf_counter() - t0
    d.close()

    def sweep(ckpt, dL=None):
        d = Disk(p)
        dev0 = d.device_bytes()
        t0 = time.perf_counter()
        d.adjoint_forward(steps, ckpt)
        t_f = time.perf_counter() - t0
        work = d.device_bytes() - dev0
        g, _ = d.get_state()
        prof = g[0].mean(axis=0)
        if dL is None:
            dL = np.zeros((5, d.ny, d.nx))
            dL[0] = np.broadcast_to(2.0 * (prof - target) / d.ny, (d.ny, d.nx))
        out = np.zeros((5, d.ny, d.nx))          # the caller's gradient buffer, reused across calls
        t0 = time.perf_counter()
        d.adjoint_backward(dL, out=out)
        t_b1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        dq, dal, d0 = d.adjoint_backward(dL, out=out)
        t_b2 = time.perf_counter() - t0
        d.close()
        return t_f, t_b1, t_b2, work, dL, float(np.sum((prof - target) ** 2)), dq, dal, d0

    t_afwd, t_bwd1, t_bwd, _, dL, loss, dq, dal, d0 = sweep(20)
    t_tfwd, t_tbwd1, t_tbwd, tape_bytes, _, _, tq, _, _ = sweep(0, dL)
    return {"workload": f"512x1024 gas disk, q 1e-4, alpha 1e-3, {steps} steps, checkpoints every 20, "
                        "loss = L2 of <Sigma>_phi(r) to a Gaussian-gap target",
            "forward_s": round(t_fwd, 4), "forward_with_checkpoints_s": round(t_afwd, 4),
            "backward_s": round(t_bwd, 4), "cost_ratio": round(t_bwd / t_fwd, 3),
            "backward_first_call_s": round(t_bwd1, 4), "cost_ratio_first_call": round(t_bwd1 / t_fwd, 3),
            "tape_forward_s": round(t_tfwd, 4), "tape_backward_s": round(t_tbwd, 4),
            "tape_cost_ratio": round(t_tbwd / t_fwd, 3), "tape_backward_first_call_s": round(t_tbwd1, 4),
            "tape_bytes": int(tape_bytes), "tape_dL_dq_rel_diff": abs(tq - dq) / max(abs(dq), 1e-300),
            "loss": loss, "dL_dq": dq, "dL_dalpha": dal, "dL_dsigma0_norm": float(np.linalg.norm(d0[0]))}


# words (reals) each 3D kernel reads and writes per interior cell: the sweeps
# and the orbital advection read and write the five fields (the radial sweep,
# which applies the first half kick as it loads, also reads the two viscous
# force components); the viscosity
# reads rho, u, vp and writes two components; the fused predictor-viscosity
# reads rho, u, vp, w and two viscous components and writes two; the
# sources read five fields and two viscous components and write three
# velocities (in place) or five
SPH_WORDS = {"sweep_x_gas": 12, "sweep_y_gas": 10, "sweep_z_gas": 10, "orbital_advection": 10,
             "viscosity": 5, "predictor_viscosity": 8, "sources": 10, "cfl": 5}


def sph_leg(steps: int, precision: int, cpu: bool) -> dict:
    """The 3D adiabatic disk (384 x 1536 x 64, PEM4, viscosity, planet,
    orbital advection): `steps` of the host-driven loop (CFL dt read back
    every step, as the reference's global_dt does), then a short profiled run
    for the per-kernel times."""
    import numpy as np
    from penguin_amd import Disk, config_c3

    p = config_c3(precision=precision)
    cells = p.xres * p.yres * p.zres
    d = Disk(p)
    d.run(2)
    t0 = time.perf_counter()
    dts = d.run(steps)
    el = time.perf_counter() - t0
    d.profile(True)
    d.run(3)
    kernels = {}
    for name, words in SPH_WORDS.items():
        avg_us, cnt = d.kernel_time_us(name)
        if cnt:
            kernels[name] = {"avg_us": round(avg_us, 2), "launches_per_step": cnt // 3,
                             "GBps": round(words * precision * cells / (avg_us * 1e-6) / 1e9, 1)}
    d.profile(False)
    d.close()
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches_per_step"])
    out = {"workload": "3D adiabatic disk 384x1536x64 (r, phi, theta): PEM4 + HLLC with the internal-energy "
                       "equation, polar sweep with reflecting walls, orbital advection, viscosity, planet "
                       "(q 1e-4, spline softening), host CFL dt",
           "value": round(cells * steps / el, 1), "unit": "cell-updates/s", "steps": steps,
           "ms_per_step": round(1e3 * el / steps, 3), "mean_dt": float(np.mean(dts)),
           "kernels": kernels,
           "roofline": {"bound": "hbm", "kernel": dom,
                        "achieved": kernels[dom]["GBps"], "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                        "frac": round(kernels[dom]["GBps"] * 1e9 / PEAK_HBM, 4),
                        "bytes_per_launch": SPH_WORDS[dom] * precision * cells}}
    if cpu:
        import oracle
        oracle.build()
        q = config_c3(precision=8, xres=96, yres=384, zres=64)
        o = oracle.Oracle(q)
        c0 = time.perf_counter()
        o.run(8)
        cdt = time.perf_counter() - c0
        out["cpu_baseline"] = {"value": round(96 * 384 * 64 * 8 / cdt, 1), "unit": "cell-updates/s",
                               "cores": oracle.num_threads(), "kind": "port",
                               "sample": f"8 steps of a 96x384x64 3D disk (FP64 oracle, {cdt:.1f} s)"}
    return out


TRAFFIC_FILE = os.path.join(HERE, "profiles", "c2_gas_dust_1024x3072_hbm_counters.md")


def measured_traffic(kernel: str, nr: int, nphi: int, precision: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE passes (copy-calibrated), for the workload they
    were taken on (1024x3072 FP32); None for any other workload."""
    if (nr, nphi, precision) != (1024, 3072, 4) or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return int(f[2])
    return None


def copy_probe(mib: int):
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))


def persistent_probe(npass: int = 200) -> dict:
    """What one persistent launch of the step would pay per pass boundary:
    a grid-wide barrier of a cooperative launch, against the gap between two
    kernel nodes of a captured graph (the batch's step has five boundaries:
    prologue, x sweep, azimuthal pass, predictor, corrector)."""
    from penguin_amd.disk import probe_grid_barrier

    out = {"passes": npass, "threads_per_block": 256, "us_per_grid_barrier": {}, "us_per_graph_node": {}}
    for blocks in (256, 1024, 2048):
        ub, un = probe_grid_barrier(blocks, npass)
        out["us_per_grid_barrier"][str(blocks)] = round(ub, 3)
        out["us_per_graph_node"][str(blocks)] = round(un, 3)
    return out


def kernel_times(d, cells: int, word: int, steps: int) -> dict:
    """Per-kernel durations from HIP events around every launch of `steps`
    uncaptured steps of disk d."""
    d.profile(True)
    d.run(steps, use_graph=False)
    kernels = {}
    for name, words in WORDS_PER_CELL.items():
        avg_us, cnt = d.kernel_time_us(name)
        if cnt:
            kernels[name] = {"avg_us": round(avg_us, 2), "launches": cnt,
                             "GBps": round(words * word * cells / (avg_us * 1e-6) / 1e9, 1) if words else None}
    d.profile(False)
    return kernels


def dominant(kernels: dict) -> str:
    return max((k for k in kernels if WORDS_PER_CELL[k]), key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])


def fp64_leg(nr: int, nphi: int, steps: int) -> dict:
    """The same disk at the reference's own precision (FP64 storage and
    arithmetic): the timed device loop and its dominant kernel."""
    from penguin_amd import Disk, config_c2

    p = config_c2(precision=8, xres=nr, yres=nphi)
    cells = nr * nphi
    d = Disk(p)
    d.run(3, use_graph=True)
    t0 = time.perf_counter()
    d.run(steps, use_graph=True)
    el = time.perf_counter() - t0
    kernels = kernel_times(d, cells, 8, 10)
    d.close()
    dom = dominant(kernels)
    achieved = WORDS_PER_CELL[dom] * 8 * cells / (kernels[dom]["avg_us"] * 1e-6) / 1e9
    return {"steps": steps, "ms_per_step": round(1e3 * el / steps, 4), "value": round(cells * steps / el, 1),
            "unit": "cell-updates/s", "dtype": "fp64",
            "step_roofline_frac": round(cells * steps / el / (PEAK_HBM / (3 * 2 * 6 * 8)), 4),
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM / 1e9,
                         "unit": "GB/s", "frac": round(achieved * 1e9 / PEAK_HBM, 4),
                         "bytes_per_launch": WORDS_PER_CELL[dom] * 8 * cells},
            "kernels": kernels}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nr: int, nphi: int, steps: int, c1_steps: int) -> dict:
    """The FP64 oracle (OpenMP over rows) on this host's cores: `steps` of
    the benchmarked disk from its initial state (the per-step rate
    extrapolates to any run length: every step does the same work), and
    c1_steps of the 256x512 gas disk in full."""
    import oracle
    from penguin_amd import config_c1, config_c2
    oracle.build()
    o = oracle.Oracle(config_c2(precision=8, xres=nr, yres=nphi))
    c0 = time.perf_counter()
    o.run(steps)
    cdt = time.perf_counter() - c0
    o.close()
    out = {"value": round(nr * nphi * steps / cdt, 1), "unit": "cell-updates/s",
           "cores": oracle.num_threads(), "kind": "port", "cpu_model": cpu_model(),
           "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
           "sample": f"{steps} steps of the {nr}x{nphi} gas+dust disk from its initial state (FP64 oracle, "
                     f"OpenMP over rows, {cdt:.1f} s, {1e3 * cdt / steps:.1f} ms per step; a run of K steps "
                     f"takes K times that)"}
    if c1_steps > 0:
        q = config_c1(precision=8)
        o = oracle.Oracle(q)
        c0 = time.perf_counter()
        o.run(c1_steps)
        c1t = time.perf_counter() - c0
        o.close()
        out["c1"] = {"value": round(q.xres * q.yres * c1_steps / c1t, 1), "unit": "cell-updates/s",
                     "sample": f"{c1_steps} steps of the 256x512 gas disk (FP64 oracle, {c1t:.2f} s)"}
    return out


def batched_leg(ndisk: int, steps: int, precision: int) -> dict:
    """The batched workload (ndisk uneven 2D gas disks, own CFL dt each,
    captured step, autotuned schedule) against looping the single-disk path
    (pg_run with its own captured step) over the same disks, same steps."""
    from penguin_amd import Batch, Disk, batch_configs

    ps = batch_configs(ndisk, seed=1, precision=precision)
    cells = sum(int(p.xres) * int(p.yres) for p in ps)
    b = Batch(ps)
    tune_ms = b.autotune(steps=3)
    b.run(5)
    t0 = time.perf_counter()
    b.run(steps)
    tb = time.perf_counter() - t0
    sched = b.schedule()
    # per-pass device time of the batch, from HIP events around each launch
    # of a short uncaptured run after the timed region
    b.profile(True)
    b.run(10, use_graph=False)
    passes = {}
    for name in WORDS_PER_CELL:
        us, cnt = b.kernel_time_us(name)
        if cnt:
            passes[name] = round(us, 2)
    b.close()
    tl = 0.0
    for p in ps:
        d = Disk(p)
        d.run(5, use_graph=True)
        t0 = time.perf_counter()
        d.run(steps, use_graph=True)
        tl += time.perf_counter() - t0
        d.close()
    vb = cells * steps / tb
    vl = cells * steps / tl
    return {"disks": ndisk, "steps": steps, "cells": cells, "value": round(vb, 1), "unit": "cell-updates/s",
            "ms_per_step": round(1e3 * tb / steps, 4), "looped_value": round(vl, 1),
            "looped_ms_per_step": round(1e3 * tl / steps, 4), "speedup": round(vb / vl, 3),
            "schedule": {"tile_x": sched[0], "ring_threads": sched[1], "ring_order": sched[2], "pt_rows": sched[3],
                         "autotune_ms_per_step": round(tune_ms, 4)},
            "pass_us": passes,
            "workload": f"{ndisk} 2D gas disks, nr in {{128,256,384,512}}, nphi in {{256,512,1024,1536}}, "
                        "q logU[1e-5,1e-3], alpha logU[1e-4,1e-2], h U[0.03,0.08], seed 1"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the whole disk lives on one MI355X; --gpus {a.gpus} is not "
              "supported (multi-GPU radial slabs are not built)", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return

    from penguin_amd import Disk, config_c2

    p = config_c2(precision=a.precision, xres=a.nr, yres=a.nphi)
    word = a.precision
    cells = p.xres * p.yres
    d = Disk(p)
    graph = not a.no_graph
    d.run(a.warmup, use_graph=graph)

    # timed region: K steps of the device loop (dt on the device), replaying
    # the captured step; nothing else runs on the stream
    t0 = time.perf_counter()
    dts = d.run(a.steps, use_graph=graph)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = cells * a.steps / elapsed
    # per-kernel durations: HIP events on the launch stream around every
    # launch of further uncaptured steps (the events cost ~40 us per step
    # in launch gaps, so they stay out of the throughput region)
    k0 = time.perf_counter()
    kernels = kernel_times(d, cells, word, a.kernel_steps)
    kel = time.perf_counter() - k0
    dom = dominant(kernels)
    bytes_launch = WORDS_PER_CELL[dom] * word * cells
    achieved = bytes_launch / (kernels[dom]["avg_us"] * 1e-6) / 1e9
    # full-step roofline: 3 read+write passes of 6 fields (x, phi+OA, sources)
    step_bytes = 3 * 2 * 6 * word
    # the boundary takes and returns host FP64 state: one upload + download
    # of the whole disk (all fields, as pg_set_state / pg_get_state move it)
    hg, hd = d.get_state()
    h0 = time.perf_counter()
    d.set_state(hg, hd)
    d.get_state()
    host_rt = time.perf_counter() - h0

    note(f"single disk: {1e3 * elapsed / a.steps:.4f} ms per step, {value:.4g} cell-updates/s")

    batched = batched_leg(a.batch_disks, a.batch_steps, a.precision) if a.batch_disks > 0 else None
    if batched:
        if not a.no_probe:
            batched["persistent_probe"] = persistent_probe()
        note(f"batched: speedup {batched['speedup']}")
    adjoint = adjoint_leg(a.adjoint_steps, a.precision) if a.adjoint_steps > 0 else None
    if adjoint:
        note(f"adjoint: cost ratio {adjoint['cost_ratio']}")
    sph = sph_leg(a.sph_steps, a.precision, not a.no_cpu) if a.sph_steps > 0 else None
    if sph:
        note(f"3D disk: {sph['ms_per_step']} ms per step")

    fp64 = fp64_leg(a.nr, a.nphi, a.fp64_steps) if a.fp64_steps > 0 and a.precision == 4 else None
    if fp64:
        note(f"fp64: {fp64['ms_per_step']} ms per step")
    cpu = cpu_baseline(a.nr, a.nphi, a.cpu_steps, a.cpu_c1_steps) if not a.no_cpu and a.cpu_steps > 0 else None

    out = {
        # the headline metric as a whole: `value` is the cell-updates/s, the
        # roofline object the %HBM, the batched and adjoint objects the rest
        "metric": "cell-updates/s on 1 GPU + %HBM roofline; batched vs looped; adjoint/forward cost",
        "value": round(value, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 4),
        "timing": ("captured step replayed (hipGraph)" if graph else "uncaptured launches") +
                  f"; kernel durations from HIP events over {a.kernel_steps} further uncaptured steps "
                  f"({1e3 * kel / a.kernel_steps:.4f} ms per step with the events)",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if a.precision == 4 else "fp64",
        "data": "synthetic",
        "config": {"workload": f"2D gas+dust disk {a.nr}x{a.nphi} (PPM4+HLLC, orbital advection, "
                               f"viscosity, planet, wave killing), device-side CFL dt",
                   "nr": a.nr, "nphi": a.nphi, "mean_dt": float(sum(dts) / len(dts))},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": round(achieved * 1e9 / PEAK_HBM, 4),
                     "traffic": measured_traffic(dom, a.nr, a.nphi, a.precision),
                     "bytes_per_launch": bytes_launch},
        "step_roofline": {"bytes_per_cell_update": step_bytes,
                          "roofline_cell_updates_per_s": PEAK_HBM / step_bytes,
                          "frac": round(value / (PEAK_HBM / step_bytes), 4)},
        "kernels": kernels,
        "host_state_roundtrip_ms": round(1e3 * host_rt, 3),
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        out["fp64"] = fp64
    if batched is not None:
        out["batched"] = batched
    if adjoint is not None:
        out["adjoint"] = adjoint
    if sph is not None:
        out["sph"] = sph
    print(json.dumps(out))


if __name__ == "__main__":
    main()
