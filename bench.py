"""Benchmark: periodic linked-cell k-nearest search on one MI355X.

Default workload: the clustered slab box 120x120x360 with N = 4,194,304
FP64 points and k = 16 (hint 3.0), inputs resident in HBM. One step is
one full search (fold, count, scan, scatter, cell sort, walk, packed
outputs) through the lc_gpu_* C waist. Prints one JSON line.

The library's autotuner first picks the walk launch for the frame's size
bucket (the row walk launched per tile group or persistent; --no-autotune
keeps the default row walk). The plain walk is timed first; then the
learned scheduler is trained on 32
further frames of the same mixture (labels from the plain walk, training on
the device), and the learned walk is timed, checked bitwise against the
plain walk's outputs. `value` is the learned walk's throughput; the plain
walk's sits under "plain" (--schedule plain skips the learned leg).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]

--config c4 runs the uneven 256-frame batch through the captured-graph
path instead (its own JSON line).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("source-queries/sec at N=4M,k=16 FP64, learned vs plain schedule; "
          "% HBM peak")
HBM_PEAK_GBS = 8000.0
BYTES_PER_CANDIDATE = 28  # 3 FP64 SoA coordinates + int32 index


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from linkcell_amd import workloads as wl
    if name == "c1":
        return wl.uniform_c1()
    if name == "c2":
        return wl.triclinic_c2()
    if name == "c3":
        return wl.clustered_c3(0)
    if name == "c5ii":
        return wl.fcc_c5ii()
    raise SystemExit(f"unknown config {name}")


def fold_cells(ws, xyz, c12, hint):
    """Home cell ids of every point, used only to count algorithmic bytes:
    the library's own plan of the cell (widths, orthorhombic flag, mesh,
    lc_gpu_dcell_plan) and a numpy fold and bin with it."""
    import torch
    dc = torch.tensor(c12, dtype=torch.float64, device="cuda:0")
    plan = ws.dcell_plan(dc.data_ptr(), 12, hint)
    if plan["status"] != 0:
        raise SystemExit(f"bench: the device could not plan the cell (status {plan['status']})")
    H = np.asarray(plan["h"]).reshape(3, 3)
    o = np.asarray(plan["origin"])
    if plan["ortho"]:
        s = (xyz - o) / np.asarray(plan["widths"])
    else:
        s = (xyz - o) @ np.linalg.inv(H)
    s = s - np.floor(s)
    s = np.where(s >= 1.0, 0.0, s)
    n = np.array([plan["nx"], plan["ny"], plan["nz"]])
    b = np.clip((s * n).astype(np.int64), 0, n - 1)
    cid = (b[:, 2] * n[1] + b[:, 1]) * n[0] + b[:, 0]
    return cid, plan


def periodic_window_sum(a, axis, r):
    """sum_{d=-r..r} a[(i + d) mod n] along axis (each wrap counts)."""
    out = np.zeros_like(a)
    for d in range(-r, r + 1):
        out += np.roll(a, -d, axis=axis)
    return out


def walk_bytes(cid, plan, reach, k):
    """B_walk = N_act (24 + 4 + 12 k) + sum over home cells c with an
    active source of sum_{|o|inf <= R_c} 28 occ(c + o), R_c the largest
    final reach among c's sources."""
    nx, ny, nz = plan["nx"], plan["ny"], plan["nz"]
    act = reach > 0
    n_act = int(act.sum())
    occ = np.bincount(cid[act], minlength=nx * ny * nz).astype(np.float64)
    rc = np.zeros(nx * ny * nz, dtype=np.int64)
    np.maximum.at(rc, cid[act], reach[act])
    occ3 = occ.reshape(nz, ny, nx)
    total = 0.0
    for r in np.unique(rc[rc > 0]):
        w = occ3
        for ax in range(3):
            w = periodic_window_sum(w, ax, int(r))
        total += float(w.ravel()[rc == r].sum())
    return n_act * (24 + 4 + 12 * k) + BYTES_PER_CANDIDATE * total, n_act


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo), as lscpu prints it."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(xyz, cell, k, hint, mask, target_s=12.0):
    """Oracle (C restatement of the reference CPU walk) on this host's
    cores: full fold/bin/list preparation, the walk on a random sample of
    sources, extrapolated to all sources."""
    from oracle import lcoracle as orc
    n = len(xyz)
    threads = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(7)
    m = min(n, 16384)
    src = np.sort(rng.choice(n, m, replace=False))
    _, _, _, _, prep, walk = orc.knearest_sample(xyz, cell, k, src, mask=mask,
                                                 cell_hint=hint, nthreads=threads)
    per = walk / m
    m2 = int(min(n, max(m, (target_s - prep) / max(per, 1e-9))))
    if m2 > m:
        src = np.sort(rng.choice(n, m2, replace=False))
        _, _, _, _, prep, walk = orc.knearest_sample(xyz, cell, k, src, mask=mask,
                                                     cell_hint=hint, nthreads=threads)
        m = m2
    t_full = prep + walk * (n / m)
    return {"value": n / t_full, "unit": "source-queries/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": (f"C restatement of the reference CPU walk, {threads} threads: fold/bin/"
                       f"lists over all {n} points ({prep:.2f} s) + walk of {m} random "
                       f"sources ({walk:.2f} s), extrapolated to {n} sources")}


PMC_PROFILES = {"plain": ROOT / "profiles" / "walk_rows_clustered_4M_k16_pmc.txt",
                "learned": ROOT / "profiles" / "walk_rows_learned_clustered_4M_k16_pmc.txt"}


def pmc_traffic(workload_name, schedule):
    """HBM bytes per walk of the given schedule from the committed rocprofv3
    --pmc summary of the same workload (FETCH_SIZE + WRITE_SIZE of the walk
    kernels, each scaled by the copy-kernel calibration recorded there), or
    None."""
    prof = PMC_PROFILES.get(schedule)
    if prof is None or not prof.exists():
        return None, None
    vals = {}
    for line in prof.read_text().splitlines():
        if line.startswith("#") or ":" not in line:
            continue
        key, _, val = line.partition(":")
        vals[key.strip()] = val.strip()
    if vals.get("workload") != workload_name or "traffic_bytes_per_launch" not in vals:
        return None, None
    return float(vals["traffic_bytes_per_launch"]), prof.name


def copy_calibration(mib):
    """One elementwise kernel of known traffic (reads and writes mib MiB),
    for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = (mib << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    log(f"bench: copy calibration kernel, {mib} MiB read + {mib} MiB written")


def bench_batch(args):
    """Uneven batch of 256 frames (n from 1e3 to 2e5, orthorhombic and
    triclinic, uniform and clustered, k in {4, 12, 32}) in one pipeline,
    captured once into a hipGraph and replayed: the plain schedule, then
    the learned one (trained on four other 256-frame batches of the same
    generator), checked bitwise against the plain outputs."""
    import torch
    import linkcell_amd as lc
    from linkcell_amd import workloads as wl
    dev = torch.device("cuda:0")

    def device_batch(frames):
        keep, recs = [], []
        for fr in frames:
            n, k = len(fr["xyz"]), fr["k"]
            x = torch.from_numpy(np.ascontiguousarray(fr["xyz"])).to(dev)
            nn = torch.empty((n, k), dtype=torch.int32, device=dev)
            d2 = torch.empty((n, k), dtype=torch.float64, device=dev)
            keep += [x, nn, d2]
            c = np.asarray(fr["cell"], dtype=np.float64)
            recs.append(dict(xyz=x.data_ptr(), n=n, cell=list(c.reshape(9)) + [0.0, 0.0, 0.0],
                             k=k, cell_hint=fr["hint"], out_nn=nn.data_ptr(),
                             out_d2=d2.data_ptr()))
        torch.cuda.synchronize()
        return keep, recs

    frames = wl.batched_c4(256)
    ntot = sum(len(fr["xyz"]) for fr in frames)
    keep, recs = device_batch(frames)
    ws = lc.Workspace()
    best, tune_ms = ws.autotune(recs)
    log(f"bench: batch of {len(frames)} frames, {ntot} points; tuned split {best} "
        f"({tune_ms:.3f} ms)")

    def timed_graph(tag):
        ws.graph_capture(recs)
        for _ in range(args.warmup):
            ws.graph_launch(wait=False)
        ws.wait()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ws.graph_launch(wait=False)
        ws.wait()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        ws.graph_reset()
        log(f"bench: {tag}: {ms:.3f} ms per replay")
        return ms

    ws.set_schedule(0)
    ms_plain = timed_graph("plain graph")
    ref = [t.clone() for t in keep[1::3]] + [t.clone() for t in keep[2::3]]
    learned = None
    sched = "plain"
    ms_step = ms_plain
    if args.schedule == "learned" and best == -1:
        t0 = time.perf_counter()
        ws.sched_clear()
        for b in range(4):
            tk, tr = device_batch(wl.batched_c4(256, seed0=10000 + 256 * b))
            ws.knearest_batch(tr)
            ws.sched_collect()
            del tk
        loss, margin = ws.sched_train(epochs=args.epochs, batch=4096, lr=args.lr, seed=1,
                                      quantile=args.quantile)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        def same_as_plain():
            got = keep[1::3] + keep[2::3]
            return all(bool(torch.equal(g.view(torch.int64) if g.dtype == torch.float64 else g,
                                        r.view(torch.int64) if r.dtype == torch.float64 else r))
                       for g, r in zip(got, ref))

        ws.set_schedule(3)
        ms_d = timed_graph("learned depth graph")
        if not same_as_plain():
            raise SystemExit("bench: the depth schedule differs from the plain walk")
        ws.set_schedule(1)
        ms_l = timed_graph("learned graph")
        same = same_as_plain()
        if not same:
            raise SystemExit("bench: the learned walk differs from the plain walk")
        learned = {"ms_per_step": ms_l, "train_batches": 4, "samples": ws.sched_samples(),
                   "loss": loss, "margin_log2": margin, "train_s": t1 - t0,
                   "bitwise_equal_to_plain": same,
                   "depth_schedule": {"ms_per_step": ms_d, "over_plain": ms_plain / ms_d}}
        ms_step, sched = ms_l, "learned"
    out = {
        "metric": "source-queries/sec, 256 uneven frames in one captured pipeline",
        "value": ntot / (ms_step * 1e-3),
        "unit": "source-queries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, see linkcell_amd/workloads.py)",
        "config": {"workload": "uneven-batch-256-frames", "points": ntot,
                   "k": [4, 12, 32], "split": best, "graph": True, "schedule": sched},
        "plain": {"value": ntot / (ms_plain * 1e-3), "ms_per_step": ms_plain},
    }
    if learned is not None:
        out["learned"] = learned
        out["learned_over_plain"] = ms_plain / ms_step
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the library's autotuner (which picks the walk launch "
                         "for the frame's size bucket) and use the default row walk")
    ap.add_argument("--mode", type=int, default=0)
    ap.add_argument("--dense-min", type=int, default=None,
                    help="walk split: -1 row walk, else the home-cell occupancy at "
                         "which the wave-per-cell walk takes over")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MiB",
                    help="also run one elementwise copy kernel of this size "
                         "(for scaling HBM counters)")
    ap.add_argument("--schedule", choices=("learned", "plain"), default="learned",
                    help="headline schedule (learned: train the scheduler on "
                         "--train-frames frames of the same mixture, then time it "
                         "beside the plain walk)")
    ap.add_argument("--train-frames", type=int, default=None,
                    help="training frames of the learned leg (default 32 for c3, 8 for c5ii)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=4e-3, help="Adam learning rate of the scheduler")
    ap.add_argument("--quantile", type=float, default=0.999,
                    help="margin quantile of the training residuals (a higher one re-walks fewer rows)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the search runs on one GPU; sharding it across GPUs is "
              "designed but not built, so --gpus must be 1", file=sys.stderr)
        raise SystemExit(2)

    if args.config == "c4":
        return bench_batch(args)
    import torch
    import linkcell_amd as lc

    cfg = workload(args.config)
    xyz, cell, k, hint, mask = cfg["xyz"], cfg["cell"], cfg["k"], cfg["hint"], cfg["mask"]
    n = len(xyz)
    log(f"bench: config {cfg['name']} n={n} k={k}")
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(xyz).to(dev)
    dm = torch.from_numpy(mask).to(dev) if mask is not None else None
    nn = torch.empty((n, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((n, k), dtype=torch.float64, device=dev)
    rr = torch.empty((n,), dtype=torch.int32, device=dev)
    c12 = list(np.asarray(cell, dtype=np.float64).reshape(-1))
    if len(c12) == 3:
        c12 = [c12[0], 0, 0, 0, c12[1], 0, 0, 0, c12[2], 0, 0, 0]
    elif len(c12) == 9:
        c12 = c12 + [0.0, 0.0, 0.0]
    ws = lc.Workspace()
    ws.set_mode(args.mode)
    if args.dense_min is not None:
        ws.set_dense_min(args.dense_min)
    mptr = dm.data_ptr() if dm is not None else 0
    torch.cuda.synchronize()
    tuned = None
    if args.autotune and args.dense_min is None and args.mode == 0:
        # the library's autotuner picks the walk launch for this frame's
        # size bucket (lc_gpu_autotune); later searches use its choice
        rec = dict(xyz=dx.data_ptr(), n=n, cell=c12, k=k, cell_hint=hint, mask=mptr,
                   out_nn=nn.data_ptr(), out_d2=d2.data_ptr())
        tuned, tune_ms = ws.autotune([rec])
        log(f"bench: tuned split {tuned} ({tune_ms:.3f} ms)")

    def step(wait=False, reach=False):
        ws.knearest_ptr(dx.data_ptr(), n, 1, c12, k, nn.data_ptr(), d2.data_ptr(),
                        mptr, hint, rr.data_ptr() if reach else 0, wait)

    def leg(tag):
        """warmup, then K timed steps with hipEvents on the workspace stream
        around every stage of every step (medians over the K steps), then
        one step with the work counters on"""
        for _ in range(args.warmup):
            step()
        ws.wait()
        torch.cuda.synchronize()
        ws.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ws.wait()
        torch.cuda.synchronize()
        ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
        hist = ws.timing_history(args.steps)
        ws.set_timing(False)
        log(f"bench: {tag}: timed {args.steps} steps, {ms_step:.3f} ms/step")
        kern = {key: float(np.median([t[key] for t in hist])) for key in hist[0]}
        kern["walk_exact"] = float(np.median([t["walk"] + t["exact"] for t in hist]))
        kern["prep"] = float(np.median([t["fold"] + t["scan"] + t["scatter"] + t["sort"]
                                        for t in hist]))
        kern["steps_timed"] = len(hist)
        ws.set_timing(True)
        step(wait=True)
        last = ws.last_timing()
        ws.set_timing(False)
        kern["deferred"], kern["retried"] = last["deferred"], last["retried"]
        ws.set_stats(True)
        step(wait=True)
        work = ws.last_stats()
        ws.set_stats(False)
        return ms_step, kern, work

    # plain schedule (also records every source's final reach for B_walk)
    ws.set_schedule(0)
    step(wait=True, reach=True)
    reach = rr.cpu().numpy()
    ms_plain, kern_plain, work_plain = leg("plain")
    n_act = int((reach > 0).sum()) if mask is None else int(mask.sum())
    cid, plan = fold_cells(ws, xyz, c12, hint)
    bwalk, _ = walk_bytes(cid, plan, reach, k)
    split = args.dense_min if args.dense_min is not None else (tuned if tuned is not None else -1)
    row_split = split in (-1, -2)
    kc = 4 if k <= 4 else 8 if k <= 8 else 16 if k <= 16 else 32 if k <= 32 else 0

    def roof(kern):
        t_walk = kern["walk_exact"] * 1e-3
        achieved = bwalk / t_walk / 1e9
        return t_walk, achieved

    # preparation stage (fold, count, scan, scatter, cell sort, boxes):
    # streaming, B_prep = N (124 + 4 with a mask) + 12 nC bytes
    ncell = plan["nx"] * plan["ny"] * plan["nz"]
    bprep = n * (124 + (4 if mask is not None else 0)) + 12 * ncell

    def prep_of(kern):
        t = kern["prep"] * 1e-3
        return {"bytes": bprep, "ms": kern["prep"], "achieved_GBs": bprep / t / 1e9,
                "frac": bprep / t / 1e9 / HBM_PEAK_GBS}

    t_walk_plain, ach_plain = roof(kern_plain)
    plain = {"value": n_act / (ms_plain * 1e-3), "ms_per_step": ms_plain,
             "walk_ms": t_walk_plain * 1e3, "walk_qps": n_act / t_walk_plain,
             "roofline_frac": ach_plain / HBM_PEAK_GBS, "kernels_ms": kern_plain,
             "prep": prep_of(kern_plain), "walk_work": work_plain}

    learned = None
    learned_ok = (args.schedule == "learned" and args.config in TRAIN_FRAMES and kc != 0
                  and row_split and args.mode == 0)
    train_frames = (args.train_frames if args.train_frames is not None
                    else TRAIN_FRAMES.get(args.config, 0))
    if learned_ok:
        ref_nn, ref_d2 = nn.clone(), d2.clone()
        tn = torch.empty((n, k), dtype=torch.int32, device=dev)
        td = torch.empty((n, k), dtype=torch.float64, device=dev)
        t0 = time.perf_counter()
        ws.sched_clear()
        for f in range(1, train_frames + 1):
            tx = wl_device_frame(args.config, f, dev)
            # drawn on torch's stream; the workspace searches on its own
            torch.cuda.synchronize()
            ws.knearest_ptr(tx.data_ptr(), n, 1, c12, k, tn.data_ptr(), td.data_ptr(),
                            0, hint, 0, True)
            added = ws.sched_collect()
            del tx
            if f % 8 == 0 or f == 1:
                log(f"bench: training frame {f}: {added} tiles, "
                    f"{time.perf_counter() - t0:.2f} s")
        samples = ws.sched_samples()
        t1 = time.perf_counter()
        loss, margin = ws.sched_train(epochs=args.epochs, batch=4096, lr=args.lr, seed=1,
                                      quantile=args.quantile)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        del tn, td
        log(f"bench: learned scheduler: {train_frames} frames, {samples} tiles "
            f"collected in {t1 - t0:.2f} s, trained in {t2 - t1:.2f} s "
            f"(loss {loss:.4f}, margin {margin:.3f})")
        ws.set_schedule(1)
        step(wait=True)
        same = bool(torch.equal(nn, ref_nn)) and bool(
            torch.equal(d2.view(torch.int64), ref_d2.view(torch.int64)))
        if not same:
            raise SystemExit("bench: the learned walk differs from the plain walk")
        ms_l, kern_l, work_l = leg("learned")
        ws.set_timing(True)
        step(wait=True)
        retried = ws.last_retried()
        ws.set_timing(False)
        ntiles = ws.tile_count()
        # the depth schedule on the same network: shells below each tile's
        # predicted depth listed in one pass, tiles packed by depth
        ws.set_schedule(3)
        step(wait=True)
        same_d = bool(torch.equal(nn, ref_nn)) and bool(
            torch.equal(d2.view(torch.int64), ref_d2.view(torch.int64)))
        if not same_d:
            raise SystemExit("bench: the depth schedule differs from the plain walk")
        ms_d, kern_d, work_d = leg("learned depth")
        ws.set_schedule(1)
        depth = {"ms_per_step": ms_d, "walk_ms": kern_d["walk_exact"], "kernels_ms": kern_d,
                 "over_plain": ms_plain / ms_d, "bitwise_equal_to_plain": same_d,
                 "row_rounds": work_d["row_tiles"], "row_useful": work_d["row_useful"],
                 "merges": work_d["row_flushes"], "accepts": work_d["row_accepts"]}
        learned = {"ms_per_step": ms_l, "kernels_ms": kern_l, "walk_work": work_l,
                   "train_frames": train_frames, "samples": samples,
                   "epochs": args.epochs, "loss": loss, "margin_log2": margin,
                   "quantile": args.quantile, "collect_s": t1 - t0, "train_s": t2 - t1,
                   "retried_rows": retried, "rows": ntiles,
                   "bitwise_equal_to_plain": same, "depth_schedule": depth}

    if learned is not None:
        ms_step, kern, work, sched = learned["ms_per_step"], learned["kernels_ms"], \
            learned["walk_work"], "learned"
        walk_kernel = (f"k_walk_rows<{kc}> capped pass + retry pass "
                       "(+ k_mlp_predict, k_walk_exact)")
    else:
        ms_step, kern, work, sched = ms_plain, kern_plain, work_plain, "plain"
        if args.mode == 1 or kc == 0:
            walk_kernel = "k_walk_exact"
        elif row_split:
            walk_kernel = (f"k_walk_rows<{kc}>{' (persistent)' if split == -2 else ''} "
                           "(+ k_walk_exact for deferred sources)")
        else:
            walk_kernel = (f"k_walk_dense<{kc}> + k_walk_sparse<{kc}> "
                           "(+ k_walk_exact for deferred sources)")
    t_walk, achieved = roof(kern)
    traffic, traffic_src = (pmc_traffic(cfg["name"], sched)
                            if row_split and args.mode == 0 else (None, None))
    value = n_act / (ms_step * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "source-queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, see linkcell_amd/workloads.py)",
        "config": {"workload": cfg["name"], "n": n, "k": k, "cell_hint": hint,
                   "mesh": [plan["nx"], plan["ny"], plan["nz"]], "schedule": sched,
                   "mode": "auto" if args.mode == 0 else "exact-only",
                   "split": split, "autotuned": tuned is not None},
        "walk_ms": t_walk * 1e3,
        "walk_qps": n_act / t_walk,
        "kernels_ms": kern,
        "walk_work": work,
        "reach_hist": np.bincount(reach).tolist(),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": bwalk,
                     "kernel": walk_kernel,
                     "timing": f"hipEvent median over the {kern['steps_timed']} timed steps"},
        "prep": prep_of(kern),
        "plain": plain,
    }
    if learned is not None:
        out["learned"] = learned
        out["learned_over_plain"] = ms_plain / ms_step
    if traffic_src:
        out["roofline"]["traffic_source"] = traffic_src
    if args.copy_calib:
        copy_calibration(args.copy_calib)
    if not args.no_cpu:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(xyz, cell, k, hint, mask)
    print(json.dumps(out), flush=True)


# training frames of the learned leg per workload (drawn on the device)
TRAIN_FRAMES = {"c3": 32, "c5ii": 8}


def wl_device_frame(config, f, dev):
    from linkcell_amd import workloads as wl
    if config == "c5ii":
        return wl.fcc_c5ii_device(5500 + f, dev)
    return wl.clustered_c3_device(f, dev)


if __name__ == "__main__":
    main()
