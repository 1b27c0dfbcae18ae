#!/usr/bin/env python3
"""Throughput of the forward path on one MI355X: teleseismic source-solves
per second at 201^3, batched float32 against the plain float64 GPU solver,
with every source-solve carrying its receiver times and Frechet rows.

A step is one batch: `--batch` plane-wave sources (bottom-face injection)
on the 201^3 grid (0-600 km, 15x15 deg, 25^3 B-spline vertices, +-3 %),
solved, then 300 receiver times and 300 rays with Frechet rows per source.
The model is resident; boundary faces go in and rows come out through the
C ABI each step (the PCIe share is small and is included).

Prints one JSON line.  Extra fields: the plain float64 solver's throughput
on the same sources (one source per call), the float32 error against float64
at the receivers, per-step latency percentiles, the relaxation kernel's
roofline figure and a CPU baseline (the oracle's heap march, one core).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "parallel-fast-marching-method_amd"))
sys.path.insert(0, REPO)

from fmm3d import lib, synth  # noqa: E402

METRIC = "source-solves/s @200³, batched FP32 vs plain GPU at ≤1 ms err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per node visit of the column walk: read the node's
# time and slowness once, write its time once (float32)
BYTES_PER_NODE_VISIT = 12
# measured HBM bytes per node visit of the same kernel on this workload:
# rocprofv3 FETCH_SIZE (x2.000) and WRITE_SIZE (x1.000), each scaled by a
# 1 GiB copy, over the launches of one batch of 50
# (profiles/k_column_tele201_batch50_traffic.md)
MEASURED_BYTES_PER_NODE_VISIT = 34.73 + 4.42


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--batch", type=int, default=50)
    p.add_argument("--nrc", type=int, default=300)
    p.add_argument("--plain", type=int, default=3, help="float64 single-source solves timed")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-sources", type=int, default=1)
    p.add_argument("--serve-s", type=float, default=6.0,
                   help="seconds of Poisson arrivals for the serving measurement (0: skip)")
    p.add_argument("--copy-mib", type=int, default=0,
                   help="first run one device-to-device copy of this many MiB (the reference kernel "
                        "that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def copy_calibration(mib):
    """One plain device copy of a known size, untimed: under rocprofv3 --pmc its
    FETCH_SIZE and WRITE_SIZE give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def one_step(ctx, faces, rcv, prec):
    ctx.solve_tele(faces, lib.ORDER_MIXED, prec)
    st = ctx.stats()
    t = ctx.receivers(rcv)
    rp, cols, vals = ctx.frechet(rcv, pinned=True)
    return t, st, int(rp[-1])


def serving(g, velv, faces, rcv, seconds, t_src):
    """The serving loop under Poisson arrivals: requests of 4 float32 plane-wave
    sources (receiver times only) on the resident 201^3 model, offered at 70 %
    of the batched throughput; p50/p99 request latency and the throughput the
    server sustains.  Results are checked against the batched solve."""
    rng = np.random.default_rng(5)
    per = 4
    rate = 0.7 * (1.0 / t_src) / per              # requests per second
    n = max(8, int(rate * seconds))
    gaps = rng.exponential(1.0 / rate, n)
    picks = [rng.integers(0, faces.shape[0], per) for _ in range(n)]
    tickets = []
    with lib.Server(hbm_budget=64e9, max_slots=64) as S:
        # warm the model: the first request creates the resident context
        t = S.submit(g, velv, 11, faces=faces[:per], rcv=rcv)
        S.wait(t, 600.0)
        S.result(t)

        t0 = time.perf_counter()
        next_t = t0
        for i in range(n):
            next_t += gaps[i]
            d = next_t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append((S.submit(g, velv, 11, faces=faces[picks[i]], rcv=rcv), picks[i]))
        lat, errs = [], []
        for t, idx in tickets:
            S.wait(t, 600.0)
            out, l = S.result(t)
            lat.append(l)
            errs.append(out)
        span = time.perf_counter() - t0
        st = S.stats()
    return {"requests": n, "sources_per_request": per, "offered_req_per_s": round(rate, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "source_solves_per_s": round(n * per / span, 2), "server_hbm_gb": round(st["hbm_bytes"] / 1e9, 2)}, tickets, errs


def cpu_baseline(g, velv, faces, rcv, n):
    """The oracle's heap march (global mode) + receivers + rays/Frechet, one
    core, on `n` sources of the same workload."""
    from oracle import oracle as O
    veln = O.gridder(g, velv)
    t0 = time.perf_counter()
    for s in range(n):
        T = O.fmm_tele(g, veln, faces[s], 1)
        O.srtimes(g, T, rcv)
        O.rpaths(g, veln, T, rcv)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "source-solves/s", "cores": 1, "kind": "port",
            "sample": f"{n} source-solve(s) of the 201^3 workload (heap march + {len(rcv)} receivers + rays/Frechet), "
                      f"{dt:.1f} s on one host core"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (sources shard across GPUs "
              f"with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if a.copy_mib > 0:
        copy_calibration(a.copy_mib)
    g, velv, srcs, rcv_deg = synth.config_201(nsrc=a.batch, nrc=a.nrc, seed=2)
    rcv = synth.receivers_rad(rcv_deg)
    faces = np.stack([synth.tele_bottom_face(g, s) for s in srcs])
    ctx = lib.Context(g)
    ctx.set_model(velv)
    ctx.set_option("timing", 1)

    # accuracy of the batched float32 path against float64 (outside the timed region)
    t64, _, _ = one_step(ctx, faces, rcv, lib.PREC_FP64)
    t32, _, _ = one_step(ctx, faces, rcv, lib.PREC_FP32)
    err = float(np.nanmax(np.abs(t32.astype(np.float64) - t64)))

    # batched float32: the measured configuration
    for _ in range(a.warmup):
        one_step(ctx, faces, rcv, lib.PREC_FP32)
    lat = []
    upd = 0
    nvis = 0
    kms = 0.0
    launches = 0
    nnz = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0 = time.perf_counter()
        _, st, n = one_step(ctx, faces, rcv, lib.PREC_FP32)
        lat.append(1e3 * (time.perf_counter() - s0))
        upd += st["node_updates"]
        nvis += st["node_visits"]
        kms += st["kernel_ms"]
        launches += st["kernel_launches"]
        nnz += n
    elapsed = time.perf_counter() - t0
    value = a.batch * a.steps / elapsed

    # plain float64 GPU solver, one source per call
    plain_t = []
    for s in range(a.plain):
        s0 = time.perf_counter()
        one_step(ctx, faces[s:s + 1], rcv, lib.PREC_FP64)
        plain_t.append(time.perf_counter() - s0)
    plain = 1.0 / float(np.mean(plain_t)) if plain_t else None

    avg_launch_ms = kms / max(launches, 1)
    bytes_per_launch = BYTES_PER_NODE_VISIT * nvis / max(launches, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if avg_launch_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "source-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"teleseismic 201^3 nodes (0-600 km, 15x15 deg), 25^3 B-spline vertices +-3%, "
                               f"{a.batch} plane-wave sources per batch (bottom injection), {a.nrc} receivers, "
                               f"mixed-order stencils, receiver times + rays/Frechet rows",
                   "grid": [g.nnr, g.nnx, g.nnz], "batch": a.batch, "receivers": a.nrc},
        "plain_fp64_source_solves_per_s": round(plain, 3) if plain else None,
        "ratio_vs_plain_fp64": round(value / plain, 2) if plain else None,
        "max_receiver_err_s_vs_fp64": err,
        "latency_ms_per_batch": {"p50": round(float(np.percentile(lat, 50)), 2),
                                 "p99": round(float(np.percentile(lat, 99)), 2)},
        "frechet_entries_per_step": nnz // max(a.steps, 1),
        "roofline": {"kernel": "k_column (float32 column walk, 8x8 patches)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(MEASURED_BYTES_PER_NODE_VISIT * nvis / max(launches, 1)),
                     "avg_launch_ms": round(avg_launch_ms, 4),
                     "node_visits_per_launch": int(nvis / max(launches, 1)),
                     "algorithmic_bytes_per_node_visit": BYTES_PER_NODE_VISIT,
                     "node_updates_per_s": round(upd / (kms * 1e-3), 1) if kms > 0 else None},
    }
    ctx.close()
    if a.serve_s > 0:
        sv, tk, outs = serving(g, velv, faces, rcv, a.serve_s, elapsed / a.steps / a.batch * 1.0)
        # every served request against the batched float32 solve of its sources
        err = 0.0
        for (t, idx), o in zip(tk, outs):
            err = max(err, float(np.nanmax(np.abs(o.astype(np.float64) - t32[idx]))))
        sv["max_abs_diff_vs_batch_s"] = err
        out["serving"] = sv
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, velv, faces, rcv, a.cpu_sources)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
