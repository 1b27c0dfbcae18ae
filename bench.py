#!/usr/bin/env python3
"""Throughput of the forward path on one MI355X: teleseismic source-solves
per second at 201^3, batched float32 against the plain float64 GPU solver,
with every source-solve carrying its receiver times and Frechet rows.

A step is one batch: `--batch` plane-wave sources on the 201^3 grid (0-600
km, 15x15 deg, 25^3 B-spline vertices, +-3 %), each injected on the bottom
face and on the two side faces it enters through with 1-D (tau-p) boundary
times, solved, then 300 receiver times and 300 rays with Frechet rows per
source.  The model is resident; boundary faces go in and rows come out
through the C ABI each step (the PCIe share is small and is included).

Prints one JSON line.  Extra fields: the plain float64 solver's throughput
on the same sources (one source per call), the float32 error against float64
at the receivers, per-step latency percentiles, the relaxation kernel's
roofline figure, the same figures with bottom-face injection only (the
reference's own boundary condition), the serving loop, and CPU baselines:
the reference program compiled from its sources (oracle/_ref) under
mpirun -np 9 on 6 sources (when built), and the oracle's heap march from
the same fixed nodes on one core.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "parallel-fast-marching-method_amd"))
sys.path.insert(0, REPO)

from fmm3d import lib, synth  # noqa: E402

METRIC = "source-solves/s @200³, batched FP32 vs plain GPU at ≤1 ms err; p50/p99 latency"
# the reference program compiled from its own sources (make -C oracle ref)
REF_EXE = os.path.join(REPO, "oracle", "_ref", "fm3dt_parallel")
MPIRUN = "/opt/conda/bin/mpirun"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per node visit of the column walk: read the node's
# time and slowness once, write its time once (float32)
BYTES_PER_NODE_VISIT = 12
# measured HBM bytes per node visit of the same kernel on this workload:
# rocprofv3 FETCH_SIZE (x2.000) and WRITE_SIZE (x1.000), each scaled by a
# 1 GiB copy, over the launches of one batch of 50
# (profiles/k_column_tele201_batch50_traffic.md)
MEASURED_BYTES_PER_NODE_VISIT = 44.94 + 8.06
# wave64 VALU instructions per 64 float32 node updates of the column walk,
# from the rocprofv3 SQ_INSTS_VALU pass over one batch of 50, and the
# device's wave64 VALU issue rate at the clock that pass measured
# (GRBM_GUI_ACTIVE per XCD over the launch time, 2.28 GHz): 256 CUs x 4
# SIMDs, one wave64 instruction per 4 cycles of a 16-lane SIMD
# (profiles/k_column_tele201_batch50_valu.md)
VALU_PER_64_UPDATES = 338.8
VALU_WAVE_ISSUE_PER_S = 256 * 4 * 2.281e9 / 4


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
    p.add_argument("--cpu-port-only", action="store_true",
                   help="skip the reference program (oracle/_ref under mpirun -np 9 on 6 sources, about 30 s on "
                        "16 host cores) and report the oracle port on one core as the CPU baseline")
    p.add_argument("--serve-s", type=float, default=6.0,
                   help="seconds of Poisson arrivals for the serving measurement (0: skip)")
    p.add_argument("--workload", choices=("tele201", "mixed256", "serving-mix"), default="tele201",
                   help="tele201: the measured configuration (default); mixed256: 192 plane waves + 64 local "
                        "sources in one batch at 201^3; serving-mix: Poisson requests over 65-257^3 grids "
                        "with teleseismic and local sources")
    p.add_argument("--serve-rate", type=float, default=1.0, help="serving-mix: requests per second offered")
    p.add_argument("--no-mixed", action="store_true",
                   help="skip the 256-source mixed batch (192 teleseismic + 64 local) reported beside the headline")
    p.add_argument("--only-batch", action="store_true",
                   help="profiling: one warm-up float32 batch of the measured workload, then --steps timed ones "
                        "(hipGraph replays), nothing else; prints their times and the solver's counts")
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


class Workload:
    """Boundary data of the batch: bottom faces, and with sides the theta and
    phi side faces the waves enter through (fmm3d_solve_tele_sides)."""

    def __init__(self, g, srcs, sides):
        parts = [synth.tele_faces(g, s) for s in srcs]
        self.faces = np.stack([q[0] for q in parts])
        self.sides = sides
        if sides:
            self.xf = np.stack([q[1] for q in parts])
            self.zf = np.stack([q[2] for q in parts])
            self.side = np.stack([q[3] for q in parts])

    def take(self, lo, hi):
        w = Workload.__new__(Workload)
        w.faces, w.sides = self.faces[lo:hi], self.sides
        if self.sides:
            w.xf, w.zf, w.side = self.xf[lo:hi], self.zf[lo:hi], self.side[lo:hi]
        return w

    def solve(self, ctx, prec):
        if self.sides:
            ctx.solve_tele_sides(self.faces, self.xf, self.zf, self.side, lib.ORDER_MIXED, prec)
        else:
            ctx.solve_tele(self.faces, lib.ORDER_MIXED, prec)

    def stage(self, ctx):
        """The boundary faces into device memory, ahead of the timed steps."""
        if self.sides:
            ctx.stage_tele(self.faces, self.xf, self.zf, self.side)
        else:
            ctx.stage_tele(self.faces)


def one_step(ctx, work, rcv, prec, staged=False):
    """One batch: the solves (from the staged faces, or from the host
    arrays), receiver times and Frechet rows into pinned host memory."""
    if staged:
        ctx.solve_tele_staged(lib.ORDER_MIXED, prec)
    else:
        work.solve(ctx, prec)
    st = ctx.stats()
    t = ctx.receivers(rcv)
    rp, cols, vals = ctx.frechet(rcv, pinned=True)
    return t, st, int(rp[-1])


def row_errors(ctx, work, rcv, nv):
    """Frechet rows traced on the float32 field against those on the float64
    field of the same batch: per row the relative 2-norm of the difference,
    ||r32 - r64|| / ||r64||; median, p99, max and the share of rows above
    1e-4."""
    from scipy.sparse import csr_matrix

    def rows(prec):
        work.solve(ctx, prec)
        rp, cols, vals = ctx.frechet(rcv)
        return csr_matrix((vals.astype(np.float64), cols.astype(np.int64) - 1, rp), shape=(len(rp) - 1, nv))
    a, b = rows(lib.PREC_FP64), rows(lib.PREC_FP32)
    d = np.sqrt(np.asarray((b - a).multiply(b - a).sum(axis=1)).ravel())
    n = np.sqrt(np.asarray(a.multiply(a).sum(axis=1)).ravel())
    ok = n > 0
    rel = d[ok] / n[ok]
    return {"rows": int(ok.sum()), "median": float(np.median(rel)), "p99": float(np.percentile(rel, 99)),
            "max": float(rel.max()), "frac_above_1e-4": round(float((rel > 1e-4).mean()), 4),
            "definition": "per row ||r32 - r64||_2 / ||r64||_2, rows traced on the float32 and the float64 "
                          "fields of the timed batch"}


def serving(g, velv, faces, rcv, seconds):
    """The serving loop under Poisson arrivals: requests of 4 float32 plane-wave
    sources (bottom faces) on the resident 201^3 model, each with receiver
    times and Frechet rows (the source-solve of the headline), offered at 70 %
    of the server's own capacity (a burst of 32 requests in flight together);
    p50/p99 request latency and the throughput the server sustains.  Results
    are checked against the batched solve of the same sources."""
    rng = np.random.default_rng(5)
    per = 4
    tickets = []
    with lib.Server(hbm_budget=64e9, max_slots=64) as S:
        # warm the model: the first request creates the resident context
        t = S.submit(g, velv, 11, faces=faces[:per], rcv=rcv, rows=True)
        S.wait(t, 600.0)
        S.rows(t)
        S.result(t)
        # capacity: a burst of 32 requests submitted at once
        b0 = time.perf_counter()
        burst = [S.submit(g, velv, 11, faces=faces[rng.integers(0, faces.shape[0], per)], rcv=rcv, rows=True)
                 for _ in range(32)]
        for t in burst:
            S.wait(t, 600.0)
            S.rows(t)
            S.result(t)
        cap = 32 * per / (time.perf_counter() - b0)   # source-solves per second
        rate = 0.7 * cap / per                         # requests per second
        n = max(8, int(rate * seconds))
        gaps = rng.exponential(1.0 / rate, n)
        picks = [rng.integers(0, faces.shape[0], per) for _ in range(n)]

        t0 = time.perf_counter()
        next_t = t0
        for i in range(n):
            next_t += gaps[i]
            d = next_t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append((S.submit(g, velv, 11, faces=faces[picks[i]], rcv=rcv, rows=True), picks[i]))
        lat, errs, nnz = [], [], 0
        for t, idx in tickets:
            S.wait(t, 600.0)
            nnz += int(S.rows(t)[0][-1])
            out, l = S.result(t)
            lat.append(l)
            errs.append(out)
        span = time.perf_counter() - t0
        st = S.stats()
    return {"requests": n, "sources_per_request": per, "capacity_source_solves_per_s": round(cap, 2),
            "offered_req_per_s": round(rate, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "source_solves_per_s": round(n * per / span, 2), "server_hbm_gb": round(st["hbm_bytes"] / 1e9, 2),
            "frechet_entries": nnz, "includes_rays": True}, tickets, errs


def cpu_baseline_reference(g, velv, srcs, rcv_deg):
    """The reference program itself under `mpirun -np 9` (its hard-wired
    2x2x2 layout plus the writer rank) on 6 sources of the workload -- the
    fewest one receiver array takes -- with bottom faces, its own boundary
    condition, the same receivers and Frechet rows; the whole run is timed.
    None when the build or mpirun is missing or the run fails."""
    if not (os.path.exists(REF_EXE) and os.path.exists(MPIRUN)):
        return None
    d = tempfile.mkdtemp(prefix="fmm3d_ref_")
    try:
        synth.write_case(d, g, velv, srcs[:6], rcv_deg, fom=1, cfd=1)
        t0 = time.perf_counter()
        r = subprocess.run([MPIRUN, "-np", "9", REF_EXE], cwd=d, capture_output=True, text=True, timeout=300)
        dt = time.perf_counter() - t0
        ok = r.returncode == 0 and os.path.getsize(os.path.join(d, "rtravel.out")) > 0
    except (OSError, subprocess.SubprocessError):
        ok = False
    finally:
        shutil.rmtree(d, ignore_errors=True)
    if not ok:
        return None
    return {"value": 6 / dt, "unit": "source-solves/s", "cores": 9, "kind": "reference",
            "sample": f"6 plane-wave sources of the 201^3 workload with bottom faces (the reference's boundary "
                      f"condition), {len(rcv_deg)} receivers, rtravel.out + frechet.out: fm3dt_parallel compiled "
                      f"from its sources, mpirun -np 9 (2x2x2 ranks + writer) on {os.cpu_count()} visible host "
                      f"cores, whole run {dt:.1f} s"}


def cpu_baseline(g, velv, srcs, rcv, n):
    """The oracle's heap march (global mode) from the same fixed nodes
    (bottom and entry side faces) + receivers + rays/Frechet, one core, on
    `n` sources of the same workload."""
    from oracle import oracle as O
    veln = O.gridder(g, velv)
    sets = []
    for s in srcs[:n]:
        node, tv = synth.tele_boundary_nodes(g, s)
        fixed = np.zeros(g.nnr * g.nnx * g.nnz, np.uint8)
        tfix = np.zeros(fixed.shape)
        fixed[node] = 1
        tfix[node] = tv
        sets.append((fixed.reshape(g.nnr, g.nnx, g.nnz), tfix.reshape(g.nnr, g.nnx, g.nnz)))
    t0 = time.perf_counter()
    for fixed, tfix in sets:
        T = O.fmm_fixed(g, veln, fixed, tfix, 1)
        O.srtimes(g, T, rcv)
        O.rpaths(g, veln, T, rcv)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "source-solves/s", "cores": 1, "kind": "port",
            "sample": f"{n} source-solve(s) of the 201^3 workload (heap march from bottom + side faces + {len(rcv)} receivers + rays/Frechet), "
                      f"{dt:.1f} s on one host core"}


def mixed256(a):
    """256 sources in one batch on the 201^3 model: 192 plane waves (bottom
    and entry side faces, seed 3) and 64 local sources (depth 5-80 km,
    refinement factor 5 over radius 5-20 nodes, seed 3), solved by one
    fmm3d_solve_mixed call, then receiver times and Frechet rows.  A step
    is that whole batch in float32; the float64 batch gives the receiver
    error.  The same sources split by kind are timed beside it."""
    g, velv, _, rcv_deg = synth.config_201(nsrc=1, nrc=a.nrc, seed=2)
    rcv = synth.receivers_rad(rcv_deg)
    tele = synth.tele_sources(g, 192, 3)
    loc = synth.local_sources(g, 64, 3)
    parts = [synth.tele_faces(g, s) for s in tele]
    F, X, Z, S = (np.stack([p[k] for p in parts]) for k in range(4))
    src = np.stack([synth.point_source_rad(g, s) for s in loc])
    rad = np.array([s.radius_nodes for s in loc], np.int32)
    with lib.Context(g) as c:
        c.set_model(velv)

        def step(prec, tele_on=True, point_on=True):
            s0 = time.perf_counter()
            c.solve_mixed(F if tele_on else None, (X, Z, S) if tele_on else None,
                          src if point_on else None, rad if point_on else None, 5, lib.ORDER_MIXED, prec)
            s1 = time.perf_counter()
            st = c.stats()
            t = c.receivers(rcv)
            rp, _, _ = c.frechet(rcv, pinned=True)
            print(f"mixed256: step ({'fp64' if prec == lib.PREC_FP64 else 'fp32'}) {time.perf_counter() - s0:.2f} s",
                  file=sys.stderr, flush=True)
            return {"t": t, "solve_wall_s": s1 - s0, "device_solve_ms": st["solve_ms"], "sweeps": st["sweeps"],
                    "step_s": time.perf_counter() - s0, "entries": int(rp[-1])}
        r64 = step(lib.PREC_FP64)
        for _ in range(a.warmup):
            step(lib.PREC_FP32)
        runs = [step(lib.PREC_FP32) for _ in range(a.steps)]
        tele_only = step(lib.PREC_FP32, point_on=False)
        point_only = step(lib.PREC_FP32, tele_on=False)
        # the plain float64 GPU solver, one source per call (same pipeline:
        # receivers and rows), on the first teleseismic and the first local
        # source, weighted 192 : 64
        def plain(f):
            f()
            s0 = time.perf_counter()
            f()
            c.receivers(rcv)
            c.frechet(rcv, pinned=True)
            return time.perf_counter() - s0
        t_tele = plain(lambda: c.solve_tele_sides(F[:1], X[:1], Z[:1], S[:1], lib.ORDER_MIXED, lib.PREC_FP64))
        t_point = plain(lambda: c.solve_point(src[:1], 5, int(rad[0]), lib.ORDER_MIXED, lib.PREC_FP64))
    plain_rate = 256.0 / (192 * t_tele + 64 * t_point)
    el = sum(r["step_s"] for r in runs)
    err = float(np.nanmax(np.abs(runs[-1]["t"].astype(np.float64) - r64["t"])))
    med = lambda k: round(float(np.median([r[k] for r in runs])), 3)
    return {"metric": "source-solves/s, 192 teleseismic + 64 local sources in one batch at 201^3",
            "value": round(256 * a.steps / el, 2), "unit": "source-solves/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(1e3 * el / a.steps, 1), "higher_is_better": True,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "201^3 (0-600 km, 15x15 deg), 25^3 vertices +-3 %; 192 plane waves on bottom + entry "
                                   "side faces, 64 local sources (5-80 km deep, refinement 5 over 5-20 nodes); "
                                   f"{a.nrc} receivers; receiver times + Frechet rows"},
            "solve_wall_s": med("solve_wall_s"), "device_solve_ms": med("device_solve_ms"),
            "sweeps": runs[-1]["sweeps"], "frechet_entries": runs[-1]["entries"],
            "max_receiver_err_s_vs_fp64": err, "fp64_step_s": round(r64["step_s"], 3),
            "plain_fp64_source_solves_per_s": round(plain_rate, 3),
            "ratio_vs_plain_fp64": round(256 * a.steps / el / plain_rate, 2),
            "plain_fp64_s_per_source": {"teleseismic": round(t_tele, 4), "local": round(t_point, 4)},
            "tele_only": {k: round(tele_only[k], 3) for k in ("solve_wall_s", "device_solve_ms", "step_s")} | {"sweeps": tele_only["sweeps"]},
            "point_only": {k: round(point_only[k], 3) for k in ("solve_wall_s", "device_solve_ms", "step_s")} | {"sweeps": point_only["sweeps"]}}


def serving_mix(a):
    """The serving loop under a Poisson stream of uneven requests: grid size
    uniform over {65, 129, 201, 257}^3 (0-600 km, 15x15 deg, 25^3 vertices),
    two +-5 % velocity models per size that requests reuse (as successive
    inversion iterations would), 8-128 sources per request, three requests
    in four teleseismic (bottom faces) and one local (5-80 km deep,
    refinement 5 over 5 nodes), 100 receivers per grid, float32.  Reports
    p50/p99 request latency and the sustained source-solves/s; a sample of
    requests is checked against float64 solves of the same sources (1 ms)."""
    rng = np.random.default_rng(5)
    sizes = (65, 129, 201, 257)
    cases = {}
    for n in sizes:
        g = synth.make_grid((n, n, n), 600.0, (15.0, 15.0), nv=(25, 25, 25))
        tele = synth.tele_sources(g, 32, 5)
        cases[n] = {"g": g, "velv": [synth.make_velv(g, 0.05, 5 + k) for k in range(2)],
                    "faces": np.stack([synth.tele_bottom_face(g, s) for s in tele]),
                    "src": np.stack([synth.point_source_rad(g, s) for s in synth.local_sources(g, 32, 5)]),
                    "rcv": synth.receivers_rad(synth.receivers(g, 100, 5))}
    n_req = max(8, int(a.serve_rate * a.serve_s))
    gaps = rng.exponential(1.0 / a.serve_rate, n_req)
    reqs = []
    for i in range(n_req):
        n = sizes[rng.integers(0, 4)]
        m = int(rng.integers(0, 2))
        k = int(rng.integers(8, 129))
        pick = rng.integers(0, 32, k)
        reqs.append((n, m, "tele" if rng.random() < 0.75 else "point", pick))
    tickets = []
    with lib.Server(hbm_budget=200e9, max_slots=128) as srv:
        t0 = time.perf_counter()
        nxt = t0
        for i, (n, m, kind, pick) in enumerate(reqs):
            nxt += gaps[i]
            d = nxt - time.perf_counter()
            if d > 0:
                time.sleep(d)
            c = cases[n]
            kw = {"faces": c["faces"][pick]} if kind == "tele" else {"src": c["src"][pick], "refine": 5, "radius": 5}
            tickets.append(srv.submit(c["g"], c["velv"][m], 100 * n + m, rcv=c["rcv"], precision=lib.PREC_FP32, **kw))
            if i % 10 == 9:
                print(f"serving-mix: {i + 1} of {n_req} requests submitted", file=sys.stderr, flush=True)
        lat, outs = [], []
        for i, t in enumerate(tickets):
            srv.wait(t, 900.0)
            o, l = srv.result(t)
            lat.append(l)
            outs.append(o)
            if i % 10 == 9:
                print(f"serving-mix: {i + 1} of {n_req} requests done", file=sys.stderr, flush=True)
        span = time.perf_counter() - t0
        st = srv.stats()
    # accuracy: the first request of each (size, kind) against float64 solves of its sources
    err, seen = 0.0, set()
    for (n, m, kind, pick), o in zip(reqs, outs):
        if (n, kind) in seen:
            continue
        seen.add((n, kind))
        c = cases[n]
        with lib.Context(c["g"]) as ctx:
            ctx.set_model(c["velv"][m])
            if kind == "tele":
                ctx.solve_tele(c["faces"][pick], lib.ORDER_MIXED, lib.PREC_FP64)
            else:
                ctx.solve_point(c["src"][pick], 5, 5, lib.ORDER_MIXED, lib.PREC_FP64)
            ref = ctx.receivers(c["rcv"])
        err = max(err, float(np.nanmax(np.abs(o.astype(np.float64) - ref))))
    nsrc = int(sum(len(r[3]) for r in reqs))
    return {"metric": "serving: source-solves/s and request latency, uneven grids and source kinds",
            "value": round(nsrc / span, 2), "unit": "source-solves/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "Poisson requests over 65/129/201/257^3 grids (2 models each), 8-128 sources, "
                                   "3/4 teleseismic (bottom faces) and 1/4 local (refinement 5 over 5 nodes), "
                                   "100 receivers, float32",
                       "requests": n_req, "offered_req_per_s": a.serve_rate, "sources": nsrc},
            "p50_ms": round(float(np.percentile(lat, 50)), 1), "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "p50_ms_by_kind": {kd: round(float(np.percentile([l for l, r in zip(lat, reqs) if r[2] == kd], 50)), 1)
                               for kd in ("tele", "point") if any(r[2] == kd for r in reqs)},
            "span_s": round(span, 2), "models_resident_at_end": st["models_resident"],
            "server_hbm_gb": round(st["hbm_bytes"] / 1e9, 2),
            "max_receiver_err_s_vs_fp64": err}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: this build runs on one GPU "
              f"(sources would shard across GPUs with no exchange between them)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "mixed256":
        print(json.dumps(mixed256(a)), flush=True)
        return
    if a.workload == "serving-mix":
        print(json.dumps(serving_mix(a)), flush=True)
        return
    g, velv, srcs, rcv_deg = synth.config_201(nsrc=a.batch, nrc=a.nrc, seed=2)
    # the reference on the host cores, first: a child process started
    # before this process touches the GPU
    ref_cpu = None if (a.no_cpu or a.cpu_port_only) else cpu_baseline_reference(g, velv, srcs, rcv_deg)
    if a.copy_mib > 0:
        copy_calibration(a.copy_mib)
    rcv = synth.receivers_rad(rcv_deg)
    ctx = lib.Context(g)
    ctx.set_model(velv)
    if a.only_batch:
        work = Workload(g, srcs, True)
        one_step(ctx, work, rcv, lib.PREC_FP32)
        out = []
        for _ in range(a.steps):
            s0 = time.perf_counter()
            _, st, n = one_step(ctx, work, rcv, lib.PREC_FP32)
            out.append({"ms": round(1e3 * (time.perf_counter() - s0), 2), "solve_ms": round(st["solve_ms"], 2),
                        "sweeps": st["sweeps"], "node_updates": st["node_updates"], "node_visits": st["node_visits"],
                        "frechet_entries": n})
        print(json.dumps({"only_batch": out}), flush=True)
        return

    def measure(work, steps, warmup, plain_n):
        """float32 error against float64, the timed float32 batches (sweeps
        replayed from captured hipGraphs), the relaxation kernel's launches
        timed with HIP events on the solver's stream over as many batches
        again (direct launches), and the plain float64 solver (one source
        per call) on the same sources"""
        t64, _, _ = one_step(ctx, work, rcv, lib.PREC_FP64)
        t32, _, _ = one_step(ctx, work, rcv, lib.PREC_FP32)
        r = {"t32": t32, "err": float(np.nanmax(np.abs(t32.astype(np.float64) - t64))), "lat": [],
             "upd": 0, "nvis": 0, "kms": 0.0, "launches": 0, "nnz": 0}
        for _ in range(warmup):
            one_step(ctx, work, rcv, lib.PREC_FP32)
        # the faces coming in from pageable host memory inside every step
        t0 = time.perf_counter()
        for _ in range(steps):
            one_step(ctx, work, rcv, lib.PREC_FP32)
        r["pcie_elapsed"] = time.perf_counter() - t0
        # the measured steps: the faces staged in HBM beforehand
        work.stage(ctx)
        for _ in range(warmup):
            one_step(ctx, work, rcv, lib.PREC_FP32, staged=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            _, st, n = one_step(ctx, work, rcv, lib.PREC_FP32, staged=True)
            r["lat"].append(1e3 * (time.perf_counter() - s0))
            r["nnz"] += n
        r["elapsed"] = time.perf_counter() - t0
        r["value"] = a.batch * steps / r["elapsed"]
        # per-launch kernel time: every relaxation launch between HIP events
        ctx.set_option("timing", 1)
        d0 = time.perf_counter()
        for _ in range(steps):
            _, st, n = one_step(ctx, work, rcv, lib.PREC_FP32, staged=True)
            r["upd"] += st["node_updates"]
            r["nvis"] += st["node_visits"]
            r["kms"] += st["kernel_ms"]
            r["launches"] += st["kernel_launches"]
        r["direct_elapsed"] = time.perf_counter() - d0
        ctx.set_option("timing", 0)
        plain_t = []
        for s in range(plain_n):
            s0 = time.perf_counter()
            one_step(ctx, work.take(s, s + 1), rcv, lib.PREC_FP64)
            plain_t.append(time.perf_counter() - s0)
        r["plain"] = 1.0 / float(np.mean(plain_t)) if plain_t else None
        return r

    # the measured configuration: bottom + entry side faces
    side = measure(Workload(g, srcs, True), a.steps, a.warmup, a.plain)
    # the reference's own boundary condition: bottom faces only
    bottom_work = Workload(g, srcs, False)
    bottom = measure(bottom_work, a.steps, a.warmup, a.plain)
    value, elapsed, err, lat, plain = side["value"], side["elapsed"], side["err"], side["lat"], side["plain"]
    upd, nvis, kms, launches, nnz = side["upd"], side["nvis"], side["kms"], side["launches"], side["nnz"]
    faces = bottom_work.faces

    rows_err = row_errors(ctx, Workload(g, srcs, True), rcv, (g.nvr + 2) * (g.nvx + 2) * (g.nvz + 2))
    # the reference's own result on the GPU: the 2x2x2-partitioned float64
    # solve (bottom faces) with receivers and rows, on the first sources
    npart = min(8, a.batch)
    part_w = bottom_work.take(0, npart)
    one_step_part = lambda: (ctx.solve_tele(part_w.faces, lib.ORDER_MIXED, lib.PREC_FP64, partition=True),
                             ctx.receivers(rcv), ctx.frechet(rcv, pinned=True))
    one_step_part()
    p0 = time.perf_counter()
    one_step_part()
    part_s = time.perf_counter() - p0
    part_passes = ctx.passes()
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
                               f"{a.batch} plane-wave sources per batch injected on the bottom and entry side faces "
                               f"(1-D tau-p times), {a.nrc} receivers, "
                               f"mixed-order stencils, receiver times + rays/Frechet rows",
                   "grid": [g.nnr, g.nnx, g.nnz], "batch": a.batch, "receivers": a.nrc},
        "plain_fp64_source_solves_per_s": round(plain, 3) if plain else None,
        "ratio_vs_plain_fp64": round(value / plain, 2) if plain else None,
        "max_receiver_err_s_vs_fp64": err,
        "frechet_row_err_vs_fp64": rows_err,
        "graph_timed": True,
        "inputs": "boundary faces staged in HBM before the timed steps (fmm3d_stage_tele_sides, then "
                  "fmm3d_solve_tele_staged per step); receiver times and Frechet rows come back to host "
                  "memory inside every step",
        "pcie_inclusive": {"value": round(a.batch * a.steps / side["pcie_elapsed"], 3),
                           "ms_per_step": round(1e3 * side["pcie_elapsed"] / a.steps, 2),
                           "what": "the same batches with the faces uploaded from pageable host memory inside "
                                   "every step (fmm3d_solve_tele_sides)"},
        "ms_per_step_direct_launches": round(1e3 * side["direct_elapsed"] / a.steps, 2),
        "reference_partition_fp64": {
            "source_solves_per_s": round(npart / part_s, 3), "sources": npart,
            "passes": part_passes.tolist(),
            "what": "the reference's own 2x2x2-partitioned float64 solve (FMM3D_REFERENCE_PARTITION, bottom "
                    "faces), receivers and rows; reproduces the reference program's rtravel.out"},
        "latency_ms_per_batch": {"p50": round(float(np.percentile(lat, 50)), 2),
                                 "p99": round(float(np.percentile(lat, 99)), 2)},
        "frechet_entries_per_step": nnz // max(a.steps, 1),
        "roofline": {"kernel": "k_column (float32 column walk, 8x8 patches)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(MEASURED_BYTES_PER_NODE_VISIT * nvis / max(launches, 1)),
                     "traffic_source": "extrapolated: HBM bytes per node visit from the rocprofv3 FETCH_SIZE/"
                                       "WRITE_SIZE pass in profiles/ times this run's node visits per launch",
                     "launch_timing": "HIP events around each k_column launch on the solver's stream, over "
                                      "as many batches again as the timed ones (direct launches)",
                     "avg_launch_ms": round(avg_launch_ms, 4),
                     "node_visits_per_launch": int(nvis / max(launches, 1)),
                     "algorithmic_bytes_per_node_visit": BYTES_PER_NODE_VISIT,
                     # one source-solve's algorithmic bytes (every node once) at the measured rate
                     "useful_work_gbs": round(g.nnr * g.nnx * g.nnz * BYTES_PER_NODE_VISIT * value / 1e9, 1),
                     "node_updates_per_s": round(upd / (kms * 1e-3), 1) if kms > 0 else None,
                     "updates_per_node": round(upd / (a.steps * a.batch * g.nnr * g.nnx * g.nnz), 2),
                     # the kernel's actual bound: VALU issue (this run's updates/s times the counter
                     # pass's VALU instructions per update; not the contract's roofline)
                     "valu_issue_frac": round(upd / (kms * 1e-3) / 64 * VALU_PER_64_UPDATES
                                              / VALU_WAVE_ISSUE_PER_S, 3) if kms > 0 else None,
                     "valu_source": f"rocprofv3 SQ_INSTS_VALU pass in profiles/: {VALU_PER_64_UPDATES} wave64 VALU "
                                    f"instructions per 64 node updates; issue rate 1024 SIMDs x "
                                    f"{VALU_WAVE_ISSUE_PER_S * 4 / 1024 / 1e9:.3f} GHz / 4"},
        "bottom_faces_only": {"value": round(bottom["value"], 3),
                              "ms_per_step": round(1e3 * bottom["elapsed"] / a.steps, 2),
                              "plain_fp64_source_solves_per_s": round(bottom["plain"], 3) if bottom["plain"] else None,
                              "ratio_vs_plain_fp64": round(bottom["value"] / bottom["plain"], 2) if bottom["plain"] else None,
                              "max_receiver_err_s_vs_fp64": bottom["err"],
                              "pcie_inclusive_value": round(a.batch * a.steps / bottom["pcie_elapsed"], 3)},
    }
    ctx.close()
    if not a.no_mixed:
        m = mixed256(argparse.Namespace(nrc=a.nrc, steps=1, warmup=1))
        out["mixed256"] = {k: m[k] for k in ("value", "unit", "ms_per_step", "plain_fp64_source_solves_per_s",
                                             "ratio_vs_plain_fp64", "max_receiver_err_s_vs_fp64", "sweeps")}
        out["mixed256"]["workload"] = m["config"]["workload"] + " (float32; 1 timed step after 1 warm-up)"
    if a.serve_s > 0:
        sv, tk, outs = serving(g, velv, faces, rcv, a.serve_s)
        # every served request against the batched float32 solve of its sources
        err = 0.0
        for (t, idx), o in zip(tk, outs):
            err = max(err, float(np.nanmax(np.abs(o.astype(np.float64) - bottom["t32"][idx]))))
        sv["max_abs_diff_vs_batch_s"] = err
        out["serving"] = sv
    if not a.no_cpu:
        port = cpu_baseline(g, velv, srcs, rcv, a.cpu_sources)
        if ref_cpu is not None:
            out["cpu_baseline"] = ref_cpu
            out["cpu_baseline_port"] = port
        else:
            out["cpu_baseline"] = port
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
