#!/usr/bin/env python3
"""Born forward + exact adjoint throughput on one MI355X.

Workload (the 256^3 Born fwd+adj pair at 10 m): 304^3 padded domain, 3 shots,
128x128 receivers, 1500 steps at 1 ms.  One benchmark step runs, for every
shot, the Born forward (data from the model) and the exact adjoint (image
from those data), all inputs resident in HBM.  Work unit: a Born cell-step =
one interior cell advanced one time step through the forward (2 wavefield
updates) and the adjoint (2 wavefield updates: background and adjoint
state); Mcell-updates/s = 4 x Born cell-steps/s.  The unit is fixed per Born
cell-step: with the background history kept in HBM the adjoint advances only
the adjoint state, and the work counted is still 4 updates per cell-step.

Prints one JSON line (see the keys below).  --gpus N with N > 1 is refused:
the path shards by shot and the single-GPU number is the one measured.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "sc17-born_amd")

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per interior cell and step, FP32 / FP16 / 24-bit
# levels (4 / 2 / 3 bytes per cell; vel, model, image and traces stay FP32):
#   forward: s, r read cur + old, vel, model; s, r written (FP32: s into
#     the kept history with scatter 0; FP16 and 24-bit: s and r stored
#     together per cell and the history written apart)         32 / 22 / 29
#   adjoint state: nu read cur + old, vel, sigma, image; nu and image
#     written                                                   28 / 20 / 24
#   background recompute (steps of checkpoint segments), scatter 0: s read
#     cur + old, vel; s written (into the segment history)      16 / 10 / 13
#   background sweep of an adjoint whose forward kept neither the history
#     nor its checkpoint plan (the same kernel over every step, storing
#     checkpoints)                                              16 / 10 / 13
FWD_BYTES_PER_CELL = 32
BYTES = {"f32": {"fwd": 32, "lam": 28, "srec": 16}, "f16": {"fwd": 22, "lam": 20, "srec": 10},
         "f24": {"fwd": 29, "lam": 24, "srec": 13}}


def born_bytes(storage, nt, recomputed, swept=0):
    """Algorithmic bytes per Born cell-step of one forward + exact adjoint,
    for the checkpoint plan that ran (recomputed background steps, and the
    background steps swept first when the forward kept neither the history
    nor the checkpoints)."""
    b = BYTES[storage]
    return b["fwd"] + b["lam"] + b["srec"] * (recomputed + swept) / nt
METRIC = "Born fwd+adj Mcell-updates/s + HBM GB/s vs peak; FP16-store batch vs FP32 plain"
# HBM bytes per launch of the forward step kernel on the 304^3 workload, from
# rocprofv3 FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0) runs of this script
# (profiles/born_fwd_adj_304cube_hbm_counters.md); reported as roofline.traffic
# only when the measured workload, kernel and z chunk are the profiled ones.
PROFILED_TRAFFIC = {("c2", "FwdOp<0, float, true>, 2, 4, 1, 1, 64", 74, "f32"): 961.7e6}


def load(name):
    spec = importlib.util.spec_from_file_location("sc17born_" + name,
                                                  os.path.join(PKG, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, nthreads, target_s=10.0):
    """The oracle (C restatement of the reference's cpu_prop.cc) on this host:
    rtmForward + sourceProp + rtmAdjoint on the same domain, a bounded number
    of steps, timed with all threads given.  The source is built here for
    this host's CPU, -O3 -march=native and -Ofast -march=native (the
    reference's CMakeLists.txt:10 flags); the faster build is the value."""
    sys.path.insert(0, ROOT)
    import subprocess

    import oracle

    variants = []
    try:
        subprocess.run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle"), "native"],
                       check=True, timeout=300, stdout=subprocess.DEVNULL)
        variants = [("native", "-O3 -march=native"), ("ofast", "-Ofast -march=native")]
    except (OSError, subprocess.SubprocessError) as e:
        log("native oracle build failed (%s); timing the checker build" % e)
        variants = [(None, "-O3 -ffp-contract=off")]

    def run(variant, nt):
        P = oracle.Prop(nthreads, variant=variant)
        P.create_space(p.d, p.d, p.d, p.nx, p.ny, p.nz)
        P.transfer_sinc_table_s(p.tab)
        P.transfer_sinc_table_d(p.tab)
        locs, vals = p.shots[0]
        P.transfer_source_func(locs, vals)
        P.transfer_vel_func1(p.V)
        P.transfer_vel_func2(p.V)
        rec = np.zeros((p.nrec, p.nt_big), np.float32)
        P.transfer_receiver_func(p.rec_locs, rec)
        a0 = np.zeros(p.nx * p.ny * p.nz, np.float32)
        a1 = np.zeros_like(a0)
        img = np.zeros_like(a0)
        t0 = time.perf_counter()
        P.rtm_forward(1, p.model, nt)
        P.source_prop(a0, a1, 1, nt)
        P.rtm_adjoint(a1, a0, img, nt)
        return time.perf_counter() - t0

    results = {}
    for variant, flags in variants:
        t4 = run(variant, 4)
        t12 = run(variant, 12)
        per_step = max((t12 - t4) / 8, 1e-4)
        nt = int(max(4, min(1500, target_s / per_step)))
        t = run(variant, nt)
        results[flags] = (4 * p.interior_cells * nt / t / 1e6, nt, t)
        log("cpu baseline %s: %.1f Mcell-updates/s (%d steps, %.1f s)"
            % (flags, results[flags][0], nt, t))
    best = max(results, key=lambda k: results[k][0])
    v, nt, t = results[best]
    return {
        "value": round(v, 2),
        "unit": "Mcell-updates/s",
        "cores": nthreads,
        "kind": "port",
        "cpu": cpu_model(),
        "host_cpus": len(os.sched_getaffinity(0)),
        "flags": best,
        "by_flags": {k: round(r[0], 2) for k, r in results.items()},
        "sample": "oracle rtmForward + sourceProp + rtmAdjoint (reference CPU semantics, "
                  "OpenMP over z like the reference's TBB loop), %dx%dx%d domain, 1 shot, "
                  "%d steps (%.1f s), built %s on this host"
                  % (p.nx, p.ny, p.nz, nt, t, best),
    }


def batch_leg(args, pyborn, configs, torch):
    """32 shots on windowed subgrids (BATCHES["c4"]), Born forward: the FP16
    batch (one launch per step for all shots, time loop replayed from a HIP
    graph) against the same shots run one by one in FP32, each on its own
    context; the FP32 batch separates batching from storage."""
    cfg = dict(configs.BATCHES["c4"])
    if args.batch_nt:
        cfg["nt"] = args.batch_nt
    if args.batch_shots:
        cfg["nshots"] = args.batch_shots
    t0 = time.perf_counter()
    shots, tab = configs.make_batch(**cfg)
    nt = cfg["nt"]
    log("batch problem built in %.1f s: %d shots" % (time.perf_counter() - t0, len(shots)))
    dev = torch.device("cuda:0")
    cells = sum(s.interior_cells for s in shots)
    # FP32 one shot at a time
    singles = []
    for s in shots:
        P = pyborn.Propagator(0)
        P.create_space(10.0, 10.0, 10.0, s.nx, s.ny, s.nz)
        P.transfer_sinc_table_s(tab)
        P.transfer_sinc_table_d(tab)
        P.transfer_vel_func1(s.V)
        P.transfer_vel_func2(s.V)
        P.transfer_receiver_func(s.rec_nx, s.rec_ny, s.rec_locs,
                                 np.zeros((s.nrec, s.nt_big), np.float32))
        P.transfer_source_func(s.src_locs, s.src_vals)
        m = torch.from_numpy(s.model).to(dev)
        tr = torch.zeros((s.nt_big, s.nrec), dtype=torch.float32, device=dev)
        singles.append((P, m, tr))

    def run_singles():
        for P, m, tr in singles:
            tr.zero_()
            P.dev_forward(1, nt, m, tr, scatter=0, keep_background=False)
        for P, _, _ in singles:
            P.synchronize()

    run_singles()  # autotune
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_singles()
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    log("batch leg: FP32 one shot at a time %.2f s" % t_single)
    ref0 = singles[0][2].cpu().numpy().T.copy()
    for P, m, tr in singles:
        P.close()
    del singles
    torch.cuda.empty_cache()
    # the batch
    P = pyborn.Propagator(0)
    P.transfer_sinc_table_s(tab)
    P.transfer_sinc_table_d(tab)
    B = pyborn.Batch(P, [{"nx": s.nx, "ny": s.ny, "nz": s.nz, "vel": s.V, "model": s.model,
                          "src_locs": s.src_locs, "src_vals": s.src_vals, "rec_nx": s.rec_nx,
                          "rec_ny": s.rec_ny, "rec_locs": s.rec_locs} for s in shots],
                     (10.0, 10.0, 10.0))
    out, err, tile = {}, {}, {}
    for storage in ("f32", "f16", "f24"):
        P.set_storage(storage)
        B.forward(1, nt, 0)  # autotune and graph capture
        P.synchronize()
        t0 = time.perf_counter()
        B.forward(1, nt, 0)
        P.synchronize()
        out[storage] = time.perf_counter() - t0
        log("batch leg: %s batch %.2f s" % (storage, out[storage]))
        d = B.traces(0)
        err[storage] = float(np.linalg.norm(d - ref0) / np.linalg.norm(ref0))
        tile[storage] = B.tile(0)
    B.close()
    P.close()
    mcu = lambda t: round(2 * cells * nt / t / 1e6, 1)
    pair = None if args.no_batch_pair else pair_leg(args, pyborn, shots, tab, nt, torch)
    return {
        "workload": "Born forward, %d shots on windows of [%d, %d]^2 x %d cells at 10 m "
                    "(padded per shot), 96x96 receivers each, %d steps"
                    % (len(shots), cfg["nxy"][0], cfg["nxy"][1], cfg["nz"], nt),
        "interior_cells": cells,
        "unit": "Mcell-updates/s (2 wavefield updates per cell-step)",
        "f32_one_shot_at_a_time": mcu(t_single),
        "f32_batch": mcu(out["f32"]),
        "f16_batch": mcu(out["f16"]),
        "f16_batch_speedup_vs_f32_plain": round(t_single / out["f16"], 3),
        "f16_batch_tile": tile["f16"],
        "f16_vs_f32_data_rel_l2_shot0": round(err["f16"], 6),
        "f24_batch": mcu(out["f24"]),
        "f24_batch_speedup_vs_f32_plain": round(t_single / out["f24"], 3),
        "f24_batch_tile": tile["f24"],
        "f24_vs_f32_data_rel_l2_shot0": round(err["f24"], 7),
        "pair": pair,
    }


def pair_leg(args, pyborn, shots, tab, nt, torch):
    """The Born pair (forward, then the exact adjoint of its traces) on the
    first args.pair_shots windowed shots: FP32 one shot at a time on its own
    context (the background history kept by the forward when it fits), against
    one FP16 and one 24-bit batch running born_batch_forward then
    born_batch_adjoint_exact (background rebuilt from checkpoints of the whole
    batch).  Per-shot image errors against the FP32 images."""
    sel = shots[:args.pair_shots]
    cells = sum(s.interior_cells for s in sel)
    dev = torch.device("cuda:0")

    def load(P, s):
        P.create_space(10.0, 10.0, 10.0, s.nx, s.ny, s.nz)
        P.transfer_sinc_table_s(tab)
        P.transfer_sinc_table_d(tab)
        P.transfer_vel_func1(s.V)
        P.transfer_vel_func2(s.V)
        P.transfer_receiver_func(s.rec_nx, s.rec_ny, s.rec_locs,
                                 np.zeros((s.nrec, s.nt_big), np.float32))
        P.transfer_source_func(s.src_locs, s.src_vals)

    # FP32 single-shot pairs on one context (its history buffer is allocated
    # once, by an untimed pair on the largest shot); the launch shapes tuned
    # on that pair are fixed for every shot (all have the same nz), so no
    # tuning or allocation of the history is timed
    P = pyborn.Propagator(0)
    big = max(sel, key=lambda s: s.nx * s.ny * s.nz)
    load(P, big)
    m = torch.from_numpy(big.model).to(dev)
    tr = torch.zeros((big.nt_big, big.nrec), dtype=torch.float32, device=dev)
    img = torch.zeros(big.nx * big.ny * big.nz, dtype=torch.float32, device=dev)
    P.dev_forward(1, nt, m, tr, scatter=0, keep_background=True)
    P.dev_adjoint_exact(1, nt, tr, img, scatter=0)
    P.synchronize()
    tuned = {k2: v for k2, v in P.tuning().items()
             if k2 in ("fwd", "background", "adjoint_state") and v["ms"] >= 0}
    ref, t32 = [], 0.0
    for s in sel:
        load(P, s)
        for name, t in tuned.items():
            P.set_tuning(name, t["rows_per_thread"], t["waves"], t["cells_per_lane"],
                         t["lookahead"], t["zchunk"])
        m = torch.from_numpy(s.model).to(dev)
        tr = torch.zeros((s.nt_big, s.nrec), dtype=torch.float32, device=dev)
        img = torch.zeros(s.nx * s.ny * s.nz, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.dev_forward(1, nt, m, tr, scatter=0, keep_background=True)
        P.synchronize()
        t1 = time.perf_counter()
        P.dev_adjoint_exact(1, nt, tr, img, scatter=0)
        P.synchronize()
        t2 = time.perf_counter()
        t32 += t2 - t0
        log("pair leg: FP32 shot %dx%dx%d forward %.2f s, adjoint %.2f s, plan %s, swept %d"
            % (s.nx, s.ny, s.nz, t1 - t0, t2 - t1, P.last_checkpoint_plan(),
               P.last_background_sweep()))
        ref.append(img.cpu().numpy())
    P.close()
    del m, tr, img
    torch.cuda.empty_cache()
    log("pair leg: FP32 one shot at a time %.2f s" % t32)
    out = {"workload": "Born forward + exact adjoint of the first %d of the windowed shots, "
                       "%d steps" % (len(sel), nt),
           "interior_cells": cells,
           "unit": "Mcell-updates/s (4 wavefield updates per Born cell-step)",
           "f32_one_shot_at_a_time": round(4 * cells * nt / t32 / 1e6, 1)}
    P = pyborn.Propagator(0)
    P.transfer_sinc_table_s(tab)
    P.transfer_sinc_table_d(tab)
    B = pyborn.Batch(P, [{"nx": s.nx, "ny": s.ny, "nz": s.nz, "vel": s.V, "model": s.model,
                          "src_locs": s.src_locs, "src_vals": s.src_vals, "rec_nx": s.rec_nx,
                          "rec_ny": s.rec_ny, "rec_locs": s.rec_locs} for s in sel],
                     (10.0, 10.0, 10.0))
    for storage in ("f16", "f24"):
        P.set_storage(storage)
        B.forward(1, nt, 0)  # autotune and graph capture
        B.adjoint_exact(1, nt, 0)
        P.synchronize()
        t0 = time.perf_counter()
        B.forward(1, nt, 0)
        B.adjoint_exact(1, nt, 0)
        P.synchronize()
        t = time.perf_counter() - t0
        segs, rec = B.last_checkpoint_plan()
        errs = [float(np.linalg.norm(B.image(k) - r) / np.linalg.norm(r)) for k, r in enumerate(ref)]
        log("pair leg: %s batch %.2f s, %d segments" % (storage, t, segs))
        out[storage + "_batch"] = round(4 * cells * nt / t / 1e6, 1)
        out[storage + "_batch_speedup_vs_f32_plain"] = round(t32 / t, 3)
        out[storage + "_checkpoint_segments"] = segs
        out[storage + "_recomputed_steps"] = rec
        out[storage + "_image_rel_l2_vs_f32_max"] = round(max(errs), 7)
        out[storage + "_image_rel_l2_vs_f32_per_shot"] = [round(e, 7) for e in errs]
    B.close()
    P.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--budget-gb", type=float, default=0.0,
                    help="HBM budget for the background history (0: free memory)")
    ap.add_argument("--nt", type=int, default=0, help="time steps (profiling runs)")
    ap.add_argument("--shots", type=int, default=0, help="shots (profiling runs)")
    ap.add_argument("--no-f16", action="store_true",
                    help="skip the FP16- and 24-bit-storage legs")
    ap.add_argument("--no-checkpointed", action="store_true",
                    help="skip the leg whose HBM budget forces a checkpointed background")
    ap.add_argument("--leg-steps", type=int, default=3,
                    help="timed steps of the legs after the headline one (at most --steps)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-shots leg")
    ap.add_argument("--batch-nt", type=int, default=0, help="time steps of the batch leg")
    ap.add_argument("--batch-shots", type=int, default=0, help="shots of the batch leg")
    ap.add_argument("--no-batch-pair", action="store_true",
                    help="skip the Born pair on the first shots of the batch")
    ap.add_argument("--pair-shots", type=int, default=8,
                    help="shots of the batch's Born pair sub-leg")
    ap.add_argument("--fwd-shape", default="",
                    help="fix the FP32 forward launch shape: rows,waves,cells_per_lane"
                         "[,lookahead] (64-cell rows; profiling runs), e.g. 1,8,1 or 2,8,1,2")
    ap.add_argument("--tuning", default="",
                    help="fix the launch shapes of the single-shot kernels: a JSON file or "
                         "string in the form of config.tuning of an earlier line (profiling "
                         "runs that must launch the shapes a full run chose)")
    ap.add_argument("--placement", default="timed", choices=["timed", "off", "cycle"],
                    help="where the work pair of the FP32 / narrow history sits (profiling "
                         "runs compare \"off\": no gaps, the work buffers)")
    ap.add_argument("--step-pairs", type=int, default=0,
                    help="FP32 forward in two-step launches: 0 off, 1 when timed faster, 2 on")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device (a kernel of known traffic "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs one shot batch per GPU and "
              "has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    pyborn = load("pyborn")
    configs = load("configs")
    cfg = dict(configs.CONFIGS[args.config])
    if args.nt:
        cfg["nt"] = args.nt
    if args.shots:
        cfg["nshots"] = args.shots
    t_setup = time.perf_counter()
    p = configs.make_problem(**cfg)
    log("problem built in %.1f s: %dx%dx%d, %d shots, nt=%d"
        % (time.perf_counter() - t_setup, p.nx, p.ny, p.nz, len(p.shots), p.nt))

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        a = torch.zeros(256 * 2 ** 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("copy calibration: 3 x (1 GiB read + 1 GiB write)")
    P = pyborn.Propagator(0)
    P.create_space(p.d, p.d, p.d, p.nx, p.ny, p.nz)
    P.transfer_sinc_table_s(p.tab)
    P.transfer_sinc_table_d(p.tab)
    P.transfer_vel_func1(p.V)
    P.transfer_vel_func2(p.V)
    P.transfer_receiver_func(p.rec_nx, p.rec_ny, p.rec_locs,
                             np.zeros((p.nrec, p.nt_big), np.float32))
    if args.zchunk:
        P.set_zchunk(args.zchunk)
    P.set_step_pairs(args.step_pairs)
    P.set_placement(args.placement)
    if args.tuning:
        text = open(args.tuning).read() if os.path.exists(args.tuning) else args.tuning
        fixed = json.loads(text)
        fixed = fixed.get("config", fixed).get("tuning", fixed)
        for name, t in fixed.items():
            P.set_tuning(name, t["rows_per_thread"], t["waves"], t["cells_per_lane"],
                         t["lookahead"], t["zchunk"])
    if args.fwd_shape:
        sh = [int(v) for v in args.fwd_shape.split(",")]
        ry, wy, vx, la = sh + [1] * (4 - len(sh))
        P.set_tuning("fwd", ry, wy, vx, la, args.zchunk or 64)
    if args.budget_gb:
        P.set_checkpoint_budget(int(args.budget_gb * 2 ** 30))
    model = torch.from_numpy(p.model).to(dev)
    traces = torch.zeros((p.nt_big, p.nrec), dtype=torch.float32, device=dev)
    img = torch.zeros(p.nx * p.ny * p.nz, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(P.stream)
    nt = p.nt

    def run_leg(storage, steps, warmup, budget=0):
        """warmup + timed steps with the given wavefield storage and history
        budget (0: free memory); returns (elapsed s, forward events,
        adjoint events)"""
        P.set_storage(storage)
        P.set_checkpoint_budget(budget)
        fwd_events, adj_events = [], []

        def step(record):
            img.zero_()
            torch.cuda.synchronize()
            for locs, vals in p.shots:
                P.transfer_source_func(locs, vals)
                traces.zero_()
                torch.cuda.synchronize()
                e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
                e[0].record(stream)
                P.dev_forward(1, nt, model, traces, scatter=0, keep_background=True)
                e[1].record(stream)
                P.dev_adjoint_exact(1, nt, traces, img, scatter=0)
                e[2].record(stream)
                if record:
                    fwd_events.append((e[0], e[1]))
                    adj_events.append((e[1], e[2]))
            P.synchronize()

        for i in range(warmup):
            step(False)
            log("%s warmup step %d done" % (storage, i))
        torch.cuda.synchronize()
        P.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            step(True)
            log("%s timed step %d done" % (storage, i))
        torch.cuda.synchronize()
        P.synchronize()
        return time.perf_counter() - t0, fwd_events, adj_events

    budget = int(args.budget_gb * 2 ** 30)
    elapsed, fwd_events, adj_events = run_leg("f32", args.steps, args.warmup, budget)
    segs, recomputed = P.last_checkpoint_plan()
    moves = P.last_placement()
    # background steps the adjoint swept before its segments (0: the forward
    # kept the history or its checkpoint plan)
    swept = P.last_background_sweep()
    tuning = P.tuning()
    leg_steps = max(1, min(args.steps, args.leg_steps))
    leg_warm = min(args.warmup, 1)
    legs = {}
    if not args.no_f16:
        for storage in ("f16", "f24"):
            el, fe, ae = run_leg(storage, leg_steps, leg_warm)
            legs[storage] = (el, fe, ae, P.last_checkpoint_plan())
        tuning = P.tuning()
    ckpt = None
    if not args.no_checkpointed and not args.budget_gb:
        # a budget of half the FP32 history: two segments, the second half of
        # the background recomputed once from a checkpoint
        half_gb = nt * p.nx * p.ny * p.nz * 4 / 2 ** 30 / 2 + 2
        elc, fec, aec = run_leg("f32", leg_steps, leg_warm, int(half_gb * 2 ** 30))
        ckpt = (elc, fec, aec, P.last_checkpoint_plan(), half_gb, P.last_background_sweep())
        P.set_checkpoint_budget(0)

    nshots = len(p.shots)
    cell_steps = p.interior_cells * nt * nshots * args.steps
    value = 4 * cell_steps / elapsed / 1e6
    fwd_ms = np.mean([a.elapsed_time(b) for a, b in fwd_events])
    adj_ms = np.mean([a.elapsed_time(b) for a, b in adj_events])
    # dominant kernel: the fused forward step (one launch per time step)
    fwd_launch_s = fwd_ms / 1e3 / nt
    fwd_alg_bytes = FWD_BYTES_PER_CELL * p.interior_cells
    achieved = fwd_alg_bytes / fwd_launch_s / 1e9
    alg_gbs = born_bytes("f32", nt, recomputed, swept) * cell_steps / elapsed / 1e9
    ft = tuning["fwd"]
    fwd_kernel = "FwdOp<0, float, true>, %d, %d, %d, %d, %d" % (
        ft["rows_per_thread"], ft["waves"], ft["cells_per_lane"], ft["lookahead"],
        64 // ft["cells_per_lane"])
    traffic = PROFILED_TRAFFIC.get((args.config, fwd_kernel, ft["zchunk"], "f32"))
    if args.nt or args.shots:
        traffic = None
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Born fwd+adj (exact adjoint), 256^3 model at 10 m -> %dx%dx%d "
                        "domain, %d shots, %dx%d receivers, %d steps at 1 ms"
                        % (p.nx, p.ny, p.nz, nshots, p.rec_nx, p.rec_ny, nt),
            "global_batch": nshots,
            "parallelism": "single GPU",
            "born_cell_steps_per_s": round(cell_steps / elapsed, 1),
            "bytes_per_born_cell_step": round(born_bytes("f32", nt, recomputed, swept), 2),
            # the unit counts 4 wavefield updates per Born cell-step (s and r
            # in the forward, background and adjoint state in the adjoint);
            # with the background history kept by the forward the adjoint
            # executes only the adjoint-state update, 3 of the 4 counted
            "updates_executed_per_born_cell_step": 3 if swept == 0 and recomputed == 0 else
                                                   round(3 + (recomputed + swept) / nt, 3),
            "algorithmic_GBps": round(alg_gbs, 1),
            "algorithmic_frac_of_hbm": round(alg_gbs / HBM_PEAK_GBS, 4),
            "forward_ms_per_shot": round(float(fwd_ms), 2),
            "adjoint_ms_per_shot": round(float(adj_ms), 2),
            "checkpoint_segments": segs,
            "recomputed_steps": recomputed,
            # moves of the forward's r pair and of the adjoint state's pair
            # between the gaps of the history, per shot
            "placement": {"mode": args.placement, "forward_moves": moves[0],
                          "adjoint_moves": moves[1]},
            "tuning": {k: v for k, v in tuning.items() if v["ms"] >= 0},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "march_kernel<%s> (fused Born forward step)" % fwd_kernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_cell": FWD_BYTES_PER_CELL,
            "avg_launch_us": round(fwd_launch_s * 1e6, 2),
        },
    }
    leg_cells = p.interior_cells * nt * nshots * leg_steps
    names = {"f16": "f16_storage", "f24": "f24_storage"}
    for storage, (el, fe, ae, (sg, rc)) in legs.items():
        v = 4 * leg_cells / el / 1e6
        out["config"][names[storage]] = {
            "value": round(v, 1),
            "unit": "Mcell-updates/s",
            "timed_steps": leg_steps,
            "ms_per_step": round(el / leg_steps * 1e3, 2),
            "speedup_vs_f32": round(v / value, 3),
            "forward_ms_per_shot": round(float(np.mean([a.elapsed_time(b) for a, b in fe])), 2),
            "adjoint_ms_per_shot": round(float(np.mean([a.elapsed_time(b) for a, b in ae])), 2),
            "checkpoint_segments": sg,
            "recomputed_steps": rc,
            "bytes_per_born_cell_step": round(born_bytes(storage, nt, rc), 2),
            "algorithmic_GBps": round(born_bytes(storage, nt, rc) * leg_cells / el / 1e9, 1),
        }
    if ckpt is not None:
        elc, fec, aec, (sgc, rcc), gb, swc = ckpt
        vc = 4 * leg_cells / elc / 1e6
        bc = born_bytes("f32", nt, rcc, swc)
        bg = tuning.get("background", {})
        out["config"]["checkpointed"] = {
            "workload": "the same pair with a %.0f GiB history budget: the forward keeps the "
                        "checkpoint plan (checkpoints at segment starts and the last segment's "
                        "levels), the adjoint recomputes the earlier segments from their "
                        "checkpoints" % gb,
            "value": round(vc, 1),
            "unit": "Mcell-updates/s",
            "timed_steps": leg_steps,
            "ms_per_step": round(elc / leg_steps * 1e3, 2),
            "forward_ms_per_shot": round(float(np.mean([a.elapsed_time(b) for a, b in fec])), 2),
            "adjoint_ms_per_shot": round(float(np.mean([a.elapsed_time(b) for a, b in aec])), 2),
            "checkpoint_segments": sgc,
            "recomputed_steps": rcc,
            "background_steps_swept_per_adjoint": swc,
            "background_steps_per_adjoint": swc + rcc,
            "updates_executed_per_born_cell_step": round(3 + (swc + rcc) / nt, 3),
            "bytes_per_born_cell_step": round(bc, 2),
            "algorithmic_GBps": round(bc * leg_cells / elc / 1e9, 1),
            "background_kernel_tuned_us": round(bg.get("ms", -1) * 1e3, 1),
        }
    P.close()
    del model, traces, img
    torch.cuda.empty_cache()
    if not args.no_batch:
        out["config"]["batch"] = batch_leg(args, pyborn, configs, torch)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(p, nthreads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
