"""Headline benchmark: Born forward + adjoint (LSRTM pair), 80 shots, 500 x 1200.

One step = one Born forward (model -> data) plus one Born adjoint
(data -> image) over all shots, with model, data and image resident in HBM
(``mem=DEVICE``).  Work unit: one interior point-update of one fine time step;
the pair runs 4 propagations per shot (source + scattered, source + receiver),
so a step is 4 * 3998 * 797,696 * 80 = 1.0205e12 point-updates.

Prints one JSON line with the contract fields plus ``roofline`` (dominant step
kernel, HIP-event timed over the timed region) and ``cpu_baseline`` (the CPU
oracle on the box's host cores, one shot of the same workload).

The dominant (pair) kernel is timed over windows of --profile-every
consecutive launches between two HIP events on the engine's stream; the
other step kernels carry events on every --profile-every-th launch, which
over-reads their duration by a few microseconds (``kernels``: upper bounds).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per interior point and fine step of each step-kernel
# variant: 16 B for the stencil (read c, o, v2dt2; write n) plus the
# coarse-rate traffic of that pass spread over its `sub` fine steps:
# source pass writes W (4 B), scattered pass reads W and m~ (8 B), receiver
# pass reads W and updates the image (12 B).
# The lockstep forward advances both forward fields (two fine steps each per
# launch): its point-updates count all of them, with (4 + 8) / 2 B of coarse
# traffic each.
CO_BYTES = {"K_SRC": 4.0, "K_BORN_SCAT": 8.0, "K_BORN_REC": 12.0, "K_BORN_PAIR": 6.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=80)
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0 = by HBM budget)")
    ap.add_argument("--tile", type=str, default="", help="tz,tx for every call (default: engine's per-direction tiles)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-shots", type=int, default=8, help="CPU baseline: at most this many shots")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline: stop after the shot that passes this")
    ap.add_argument("--opt", action="append", default=[], help="engine option key=value (aiso_ctx_set_option)")
    ap.add_argument("--workload", default="born",
                    choices=["born", "bornfs", "bornext21", "borntime21", "tomo21", "wemva21", "tomotime21",
                             "wemvatime21", "fwi32", "serve", "windows128"],
                    help="born: the headline Born fwd+adj; bornfs: the same with a free surface; bornext21: subsurface-offset extended Born "
                         "fwd+adj with 21 offsets on the same grid and shots; borntime21: the same with 21 "
                         "time lags; tomo21: the offset-extended tomographic operator fwd+adj about a "
                         "21-offset reflectivity; wemva21: the offset-extended WEMVA operator fwd+adj "
                         "(21 offsets); tomotime21, wemvatime21: the same with 21 time lags; fwi32: L2 misfit and "
                         "its adjoint-state velocity gradient through torch.autograd, 32 shots on 400x1000 "
                         "(use --shots 32); serve: a Poisson stream of one-shot requests (nonlinear, Born "
                         "forward/adjoint, 21-offset extended Born adjoint) on 4 resident models; windows128: "
                         "the batched form, 128 nonlinear-forward shots on windows of width 300-1500 at nz 400 "
                         "(nts 3000), tuned and graph-captured, beside the same shots one per launch")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests in the stream")
    ap.add_argument("--rate", type=float, default=50.0, help="serve: mean arrivals per second")
    ap.add_argument("--slots", type=int, default=4, help="serve: engine calls running at once")
    ap.add_argument("--profile-every", type=int, default=16,
                    help="profiling in the timed region: the pair kernel timed over windows of N consecutive "
                         "launches between two events, other step kernels on every N-th launch (1: events on "
                         "every launch, which costs about 9 %% of the Born benchmark's throughput)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy 1 GiB device-to-device 3 times (HBM counter calibration)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_fwd_pair_born_lsrtm_80x500x1200_hbm_traffic.md")
KERNEL_SRC = os.path.join(ROOT, "acoustic_isotropic_operators_amd", "csrc", "kernels.hip")


def _sha16(path):
    import hashlib
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def measured_traffic():
    """Fabric bytes per launch of the lockstep Born forward kernel from the
    committed FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), with the
    summary's file name; the figure is used only when the summary was taken
    on the kernels.hip in this tree (its KERNELS_SHA16 line), else None."""
    vals = {}
    try:
        for line in open(TRAFFIC_FILE):
            if "=" in line and line.split("=", 1)[0].isupper():
                k, v = line.strip().split("=", 1)
                vals[k] = v
    except OSError:
        return None, "no counter summary"
    src = os.path.relpath(TRAFFIC_FILE, ROOT)
    if vals.get("KERNELS_SHA16") != _sha16(KERNEL_SRC):
        return None, f"{src} was taken on another kernels.hip"
    try:
        return float(vals["TRAFFIC_BYTES_PER_LAUNCH"]), src
    except (KeyError, ValueError):
        return None, f"{src} holds no TRAFFIC_BYTES_PER_LAUNCH"


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, seconds=20.0, max_shots=8):
    """Oracle (float, OpenMP over ix, the vectorized timing build of
    oracle/Makefile picked for this CPU) on the host: Born fwd + adj shot by
    shot over the same workload until `seconds` have passed or `max_shots`
    ran (at least one); Gpts/s over the shots run, per-shot p50/p99, and the
    80-shot time extrapolated linearly from them.  Threads: OMP_NUM_THREADS
    (the host-core share of one GPU), else every core of the affinity mask."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import oracle as O
    build = O.use_timing_build()
    p = w.params
    g = O.Grid(nz=p["nz"], nx=p["nx"], nts=p["nts"], sub=p["sub"], minPad=min(p["zPadMinus"], p["zPadPlus"], p["xPadMinus"], p["xPadPlus"]),
               dz=p["dz"], dx=p["dx"], dts=p["dts"], alphaCos=p["alphaCos"])
    rec = [O.Device.grid(*w.rec, p["nz"], p["nx"], p["nts"])]
    per = []
    t_all = time.perf_counter()
    for iz, ix in w.src[:max_shots]:
        src = [O.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"])]
        t0 = time.perf_counter()
        d = O.born_forward(g, w.vel, src, [w.wavelet], rec, w.model, "f")
        O.born_adjoint(g, w.vel, src, [w.wavelet], rec, d, "f")
        per.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all >= seconds:
            break
    t = sum(per)
    pts = 4.0 * w.fine_steps * w.interior * len(per)
    return {"value": pts / t / 1e9, "unit": "Gpts/s", "cores": threads, "kind": "port",
            "sample": f"{len(per)} of {w.nShot} shots run (full grid and nts, Born fwd+adj, {t:.1f} s); "
                      f"80-shot time extrapolated linearly: {t / len(per) * w.nShot:.0f} s",
            "shot_latency_ms": {"p50": 1e3 * float(np.percentile(per, 50)), "p99": 1e3 * float(np.percentile(per, 99))},
            "cpu": _cpu_model(), "build": build}


def shot_latencies(bms, bsh, nShot, calls):
    """Per-shot latency of a job of `calls` operator calls per step, each
    call carrying all nShot shots in consecutive batches (engine batch
    events in issue order, shots in order within a call): a shot's latency
    is the sum over the calls of the duration of the batch that carried it."""
    lat, i = [], 0
    while i < len(bms):
        per_shot = np.zeros(nShot)
        for _ in range(calls):
            done = 0
            while done < nShot:
                if i >= len(bms):
                    raise RuntimeError("batch events do not cover whole calls")
                n = int(bsh[i])
                per_shot[done:done + n] += float(bms[i])
                done += n
                i += 1
        lat.extend(per_shot.tolist())
    return lat


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(shots shard across GPUs with one image all-reduce, not implemented here)", file=sys.stderr)
        sys.exit(2)
    import torch
    from acoustic_isotropic_operators_amd import _capi as A
    from acoustic_isotropic_operators_amd.workloads import born_lsrtm

    if a.workload == "borntime21":
        return bench_ext(a, torch, A, born_lsrtm, time_lags=True)
    if a.workload == "bornext21":
        return bench_ext(a, torch, A, born_lsrtm)
    if a.workload == "fwi32":
        return bench_fwi(a, torch, A)
    if a.workload == "serve":
        return bench_serve(a)
    if a.workload == "windows128":
        return bench_windows(a, torch, A)
    if a.workload in ("tomo21", "wemva21", "tomotime21", "wemvatime21"):
        return bench_tomo(a, torch, A, born_lsrtm, wemva=a.workload.startswith("wemva"),
                          time_lags="time" in a.workload)
    fs = a.workload == "bornfs"
    w = born_lsrtm(nShot=a.shots, free_surface=fs)
    p = w.params
    ctx = A.Context(p)
    ctx.set_velocity(w.vel)
    configure(ctx, a)
    srcs = [A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"]) for iz, ix in w.src]
    rec = [A.Device.grid(*w.rec, p["nz"], p["nx"], p["nts"])]
    nrec = rec[0].nDev
    dev = torch.device("cuda:0")
    model = torch.from_numpy(w.model).to(dev)
    data = torch.empty((w.nShot, nrec, p["nts"]), dtype=torch.float32, device=dev)
    image = torch.empty_like(model)
    torch.cuda.synchronize()

    def step():
        ctx.born_forward(model, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
        ctx.born_adjoint(image, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    ok = bool(torch.isfinite(image).all().item()) and bool(torch.isfinite(data).all().item())
    if a.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    kinds = {name: getattr(A, name) for name in CO_BYTES}
    stats = {name: ctx.profile_read(k) for name, k in kinds.items()}
    kern = {name: {"launches": v["launches"], "ms_per_step": v["ms"] / a.steps,
                   "avg_us": 1e3 * v["ms"] / max(v["launches"], 1)} for name, v in stats.items() if v["launches"]}
    dom = max(stats, key=lambda n: stats[n]["ms"])
    st = stats[dom]
    bpp = 16.0 + CO_BYTES[dom] / p["sub"]
    avg_ms = st["ms"] / max(st["launches"], 1)
    per_launch_bytes = bpp * st["point_updates"] / max(st["launches"], 1)
    achieved = per_launch_bytes / (avg_ms / 1e3) / 1e9
    lanes = ctx.get_option("lanes")
    bms, bsh = ctx.batch_read()
    info = ctx.info()
    lat = shot_latencies(bms, bsh, w.nShot, 2)           # the forward call, then the adjoint call
    traffic, traffic_src = measured_traffic() if dom == "K_BORN_PAIR" else (None, "not the pair kernel")
    pts = 4.0 * w.fine_steps * w.interior * w.nShot
    job_bytes = w.nShot * w.interior * (4 * w.fine_steps * 16.0 + p["nts"] * 28.0)
    out = {
        "metric": f"Gpts/s, Born fwd+adj {w.nShot} shots 500x1200" + (", free surface" if fs else ""),
        "value": pts / t / 1e9,
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "shots": w.nShot, "grid": "500x1200 (padded %dx%d)" % (p["nz"], p["nx"]),
                   "nts": p["nts"], "sub": p["sub"], "shots_per_launch_adj": info["batch"], "tile": a.tile or "32x16 forward, 32x32 adjoint", "lanes": lanes,
                   "inputs": "model/data/image resident in HBM"},
        "job_roofline_frac": job_bytes / t / 1e9 / HBM_PEAK_GBS,
        "shot_latency_ms": {"p50": float(np.percentile(lat, 50)) if lat else None,
                            "p99": float(np.percentile(lat, 99)) if lat else None,
                            "shots": len(lat),
                            "definition": "forward batch + adjoint batch carrying the shot"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "traffic_counts": "L2-to-fabric bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated), "
                                       "HBM plus Infinity Cache hits", "avg_launch_ms": avg_ms,
                     "launches": st["launches"], "bytes_per_point": bpp,
                     "point_updates_per_launch": st["point_updates"] / max(st["launches"], 1),
                     "lanes": lanes,
                     "timing": ("HIP events on the engine's stream around each forward pass, before its "
                                f"{lanes} lanes fork and after they join; a launch is one fine step of every "
                                "lane's shots (the lanes' kernels run concurrently)") if lanes > 1 else
                               ("HIP events on the engine's stream around windows of consecutive launches")},
        "kernels": kern,
        "finite": ok,
    }
    if not a.no_cpu and not fs:   # the CPU baseline is the headline workload's
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds, a.cpu_shots)
    print(json.dumps(out), flush=True)


def configure(ctx, a):
    """Tile, batch and engine options from the command line."""
    if a.tile:
        tz, tx = (int(v) for v in a.tile.split(","))
        ctx.set_tile(tz, tx)
    if a.batch:
        ctx.set_batch(a.batch)
    for kv in a.opt:
        k, v = kv.split("=")
        ctx.set_option(k, int(v))
    ctx.set_option("profileEvery", a.profile_every)


def bench_ext(a, torch, A, born_lsrtm, nExt=21, time_lags=False):
    """Extended Born (nExt subsurface offsets, or time lags) forward + adjoint
    on the headline grid and shots: one JSON line, same fields as the
    headline where they apply.  Point-updates count the four propagations per
    shot."""
    w = born_lsrtm(nShot=a.shots)
    p = dict(w.params, nExt=nExt, extension=A.EXT_TIME if time_lags else A.EXT_OFFSET)
    ctx = A.Context(p)
    ctx.set_velocity(w.vel)
    configure(ctx, a)
    srcs = [A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"]) for iz, ix in w.src]
    rec = [A.Device.grid(*w.rec, p["nz"], p["nx"], p["nts"])]
    dev = torch.device("cuda:0")
    mext = np.zeros((nExt,) + w.model.shape, np.float32)
    mext[nExt // 2] = w.model
    model = torch.from_numpy(mext).to(dev)
    data = torch.empty((w.nShot, rec[0].nDev, p["nts"]), dtype=torch.float32, device=dev)
    image = torch.empty_like(model)
    torch.cuda.synchronize()

    def step():
        ctx.bornext_forward(model, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
        ctx.bornext_adjoint(image, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    stats = {n: ctx.profile_read(getattr(A, n)) for n in ("K_SRC", "K_EXT_SCAT", "K_EXT_REC")}
    pts = 4.0 * w.fine_steps * w.interior * w.nShot
    kind = "time-lag" if time_lags else "extended"
    out = {"metric": f"Gpts/s, {kind} Born fwd+adj nExt={nExt}, {w.nShot} shots 500x1200", "value": pts / t / 1e9,
           "unit": "Gpts/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"born{'time' if time_lags else 'ext'}{nExt}_{w.nShot}x500x1200", "nExt": nExt,
                      "shots": w.nShot,
                      "shots_per_launch_adj": ctx.info()["batch"]},
           "lanes": ctx.get_option("lanes"),
           "step_kernels_us_per_fine_step": {n: 1e3 * v["ms"] / v["launches"] for n, v in stats.items()
                                             if v["launches"]},
           "step_kernel_timing": ("pass-timed: events around each pass, before its lanes fork and after they join, "
                                  "per fine step of all the batch's shots" if ctx.get_option("lanes") > 1 else
                                  "events on sampled dispatches (upper bounds)"),
           "finite": bool(torch.isfinite(image).all().item())}
    print(json.dumps(out), flush=True)


def bench_tomo(a, torch, A, born_lsrtm, nExt=21, wemva=False, time_lags=False):
    """Offset-extended tomographic (or WEMVA) operator, both legs, forward +
    adjoint on the headline grid: one JSON line.  Point-updates count the
    propagations per shot: tomographic source + 4 forward (two per leg),
    source + receiver + 2 adjoint; WEMVA source + receiver + 2 each way."""
    w = born_lsrtm(nShot=a.shots)
    p = dict(w.params, nExt=nExt, extension=A.EXT_TIME if time_lags else A.EXT_OFFSET)
    ctx = A.Context(p)
    ctx.set_velocity(w.vel)
    configure(ctx, a)
    srcs = [A.Device.grid(1, iz, 1, 1, ix, 1, p["nz"], p["nx"], p["nts"]) for iz, ix in w.src]
    rec = [A.Device.grid(*w.rec, p["nz"], p["nx"], p["nts"])]
    dev = torch.device("cuda:0")
    refl = np.zeros((nExt,) + w.model.shape, np.float32)
    refl[nExt // 2] = w.model
    ctx.set_ext_reflectivity(refl)
    dv = torch.from_numpy((w.vel * 0.01).astype(np.float32)).to(dev)
    data = torch.empty((w.nShot, rec[0].nDev, p["nts"]), dtype=torch.float32, device=dev)
    grad = torch.empty_like(dv)
    if wemva:   # observed data: the Born data of the reflectivity
        ctx.born_forward(torch.from_numpy(w.model).to(dev), data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
        image = torch.empty((nExt,) + tuple(dv.shape), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def step():
        if wemva:
            ctx.wemva_forward(dv, image, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
            ctx.wemva_adjoint(grad, image, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
        else:
            ctx.tomo_forward(dv, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)
            ctx.tomo_adjoint(grad, data, srcs, w.wavelet, rec, mem=A.MEM_DEVICE)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    stats = {n: ctx.profile_read(getattr(A, n)) for n in ("K_SRC", "K_TOMO")}
    pts = (8.0 if wemva else 9.0) * w.fine_steps * w.interior * w.nShot
    name = ("WEMVA" if wemva else "tomographic") + (" time-lag" if time_lags else "")
    out = {"metric": f"Gpts/s, {name} fwd+adj nExt={nExt}, {w.nShot} shots 500x1200", "value": pts / t / 1e9,
           "unit": "Gpts/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{'wemva' if wemva else 'tomo'}{'time' if time_lags else ''}{nExt}_{w.nShot}x500x1200",
                      "nExt": nExt,
                      "shots": w.nShot,
                      "shots_per_launch": ctx.info()["batch"]},
           "step_kernels_us_per_launch_sampled": {n: 1e3 * v["ms"] / v["launches"] for n, v in stats.items()
                                                  if v["launches"]},
           "finite": bool(torch.isfinite(grad).all().item()) and bool(torch.isfinite(data).all().item())}
    print(json.dumps(out), flush=True)


def bench_fwi(a, torch, A):
    """FWI step on the layered model with a salt blob: nonlinear data of the
    starting model, L2 misfit against the true model's data, and the
    adjoint-state velocity gradient (checkpointed recomputation every
    sqrt(fine steps)), through torch.autograd with every tensor on the GPU.
    Point-updates count 4 propagations per shot: the forward, and the
    gradient's checkpointing forward, recomputation and adjoint."""
    from acoustic_isotropic_operators_amd.fwi import NonlinearVelocityOperator
    from acoustic_isotropic_operators_amd.workloads import fwi_layered
    w = fwi_layered(nShot=a.shots)
    p = w.params
    op = NonlinearVelocityOperator(p, [("grid", (1, iz, 1, 1, ix, 1)) for iz, ix in w.src], [("grid", w.rec)],
                                   w.wavelet[None])
    for kv in a.opt:
        k, v = kv.split("=")
        op.context.set_option(k, int(v))
    dev = torch.device("cuda:0")
    with torch.no_grad():
        observed = op(torch.from_numpy(w.model).to(dev))
    v0 = torch.from_numpy(w.vel).to(dev)
    torch.cuda.synchronize()

    def step():
        v = v0.clone().requires_grad_(True)
        r = op(v) - observed
        loss = 0.5 * (r * r).sum()
        loss.backward()
        return loss, v.grad

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        loss, grad = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    pts = 4.0 * w.fine_steps * w.interior * w.nShot
    out = {"metric": f"Gpts/s, FWI misfit + velocity gradient, {w.nShot} shots 400x1000", "value": pts / t / 1e9,
           "unit": "Gpts/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": t * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": w.name, "shots": w.nShot, "nts": p["nts"], "sub": p["sub"],
                      "grid": "400x1000 (padded %dx%d)" % (p["nz"], p["nx"])},
           "gradients_per_s": 1.0 / t, "misfit": float(loss.item()),
           "finite": bool(torch.isfinite(grad).all().item())}
    print(json.dumps(out), flush=True)


def bench_windows(a, torch, A):
    """Batched form over windows of uneven aperture: 128 nonlinear-forward
    shots, each on its own window (width 300-1500, nz 400, nts 3000 at 2 ms,
    sub 2), advanced together by k_step_win, one launch per fine step over
    every window of a launch group, the time loop captured in a HIP graph;
    tile, windows per launch and workgroups per CU picked by the autotuner.
    One step = one call over all 128 shots with wavelet and data resident in
    HBM.  Beside it, the plain form: the same 128 shots through
    aiso_nl_forward one shot per call and per launch, on per-window contexts
    made before the clock starts.  Work unit: interior point-updates of one
    fine step; the step kernel's algorithmic bytes are 16 per point-update."""
    from acoustic_isotropic_operators_amd.workloads import batched_windows
    w = batched_windows()
    p = w.params
    dev = torch.device("cuda:0")
    model = torch.from_numpy(w.wavelet.astype(np.float32)).to(dev)
    b = A.Batch(p)
    devs = []
    for k in range(w.nShot):
        NX, NZ = w.vels[k].shape
        iz, ix = w.src[k]
        devs.append((A.Device.grid(1, iz, 1, 1, ix, 1, NZ, NX, p["nts"]), A.Device.grid(*w.rec[k], NZ, NX, p["nts"])))
    b.set_windows(w.vels, [s for s, _ in devs], [r for _, r in devs])
    t0 = time.perf_counter()
    tuned = b.tune(steps=12)
    tune_s = time.perf_counter() - t0
    if a.tile:
        c = b.config()
        tz, tx = (int(v) for v in a.tile.split(","))
        b.set_config(tz, tx, c["windows_per_launch"], c["blocks_per_cu"], True, c["lanes"])
    out_d = [torch.zeros((r.nDev, p["nts"]), dtype=torch.float32, device=dev) for _, r in devs]
    torch.cuda.synchronize()
    for _ in range(max(1, a.warmup)):     # the first call captures the graphs
        b.nl_forward([model], out_d, mem=A.MEM_DEVICE)
    capture_ms = b.info()["capture_ms"]
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        b.nl_forward([model], out_d, mem=A.MEM_DEVICE)
        dev_ms.append(b.info()["last_ms"])
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    info = b.info()
    pts = float(info["point_updates"])
    steps_per_call = w.fine_steps * info["launch_groups"]
    avg_launch_ms = float(np.mean(dev_ms)) / steps_per_call
    achieved = 16.0 * pts / (float(np.mean(dev_ms)) / 1e3) / 1e9
    # the plain form: one shot per call and per launch
    ctxs = []
    for k in range(w.nShot):
        c = A.Context(w.window_params(k))
        c.set_batch(1)
        c.set_velocity(w.vels[k])
        ctxs.append(c)
    plain_d = [torch.zeros((1, r.nDev, p["nts"]), dtype=torch.float32, device=dev) for _, r in devs]
    m3 = model[None]
    ctxs[0].nl_forward(m3, plain_d[0], [devs[0][0]], [devs[0][1]], mem=A.MEM_DEVICE)   # warm
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for k in range(w.nShot):
        ctxs[k].nl_forward(m3, plain_d[k], [devs[k][0]], [devs[k][1]], mem=A.MEM_DEVICE)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t1
    same = all(torch.equal(out_d[k], plain_d[k][0]) for k in range(w.nShot))
    out = {"metric": "shots/s, batched nonlinear forward, 128 shots on windows nx 300-1500 at nz 400, nts 3000",
           "value": w.nShot / t, "unit": "shots/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": w.name, "shots": w.nShot, "nts": p["nts"], "sub": p["sub"],
                      "window_nx": [min(w.nx), max(w.nx)], "tuned": tuned, "tune_s": tune_s,
                      "graph_capture_ms": capture_ms, "launch_groups": info["launch_groups"],
                      "inputs": "wavelet and data resident in HBM"},
           "gpts_per_s": pts / t / 1e9,
           "roofline": {"bound": "hbm", "kernel": "k_step_win (MODE_NL_FWD)", "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "avg_launch_ms": avg_launch_ms, "launches_per_call": steps_per_call,
                        "bytes_per_point": 16.0, "timing": "HIP events around each call on the batch stream"},
           "plain_one_shot_per_launch": {"shots_per_s": w.nShot / t_plain, "s": t_plain,
                                         "gpts_per_s": pts / t_plain / 1e9},
           "speedup_vs_plain": t_plain / t,
           "bitwise_equal_to_plain": bool(same),
           "finite": all(bool(torch.isfinite(d).all().item()) for d in out_d)}
    print(json.dumps(out), flush=True)


def bench_serve(a):
    """Serving: --requests one-shot requests, mixed nonlinear modeling
    (30 %), Born forward (30 %), Born adjoint (25 %) and 21-offset
    extended-Born adjoint (15 %), on 4 resident models from 200 x 400
    (nts 4000) to 1000 x 3000 (nts 1000), the smaller grids drawn more often.
    Host buffers in and out (each request brings its data and takes its
    result).  Three runs of the same seeded requests:

    * capacity, continuous batching: every request queued at once, served by
      the batching server (--slots concurrent calls, up to 32 shots per call,
      admission by a 200 GB HBM budget); `value` is its requests (shots) per
      second;
    * capacity, plain: the same queue served one request per call, each call
      after the previous one; the speedup is the ratio of the two capacities;
    * latency: the requests arriving as a seeded Poisson stream (mean --rate
      per second) at the batching server; p50/p99 from arrival to completion.
      A stream caps throughput at its arrival rate, so its shots/s measure
      the load, not the server."""
    from acoustic_isotropic_operators_amd.serve import Request, Server, run_stream
    from acoustic_isotropic_operators_amd.workloads import serving_models
    models = serving_models()
    kinds = ["nl_forward", "born_forward", "born_adjoint", "bornext_adjoint"]
    refl_of = {}
    for mid, params, vel, rec, wav, (iz, x0, x1) in models:
        refl = np.zeros((params["nx"], params["nz"]), np.float32)
        refl[x0:x1, iz + 40:iz + 43] = 0.1
        refl_of[mid] = refl

    def one(srv, mid, kind, ix, rng, arrival=0.0):
        params = srv.models[mid].params
        iz = [m for m in models if m[0] == mid][0][5][0]
        r = Request(kind, mid, (iz, ix), arrival=arrival)
        if kind == "born_forward":
            r.payload, r.reflectivity_id = refl_of[mid], mid
        elif kind != "nl_forward":
            r.payload = rng.standard_normal((srv.models[mid].rec[0].nDev, params["nts"])).astype(np.float32)
        return r

    def stream(srv, burst):
        rng = np.random.default_rng(2026)
        arrivals = np.cumsum(rng.exponential(1.0 / a.rate, a.requests))
        reqs = []
        for t in arrivals:
            m = models[rng.choice(4, p=[0.4, 0.3, 0.2, 0.1])]
            kind = kinds[rng.choice(4, p=[0.3, 0.3, 0.25, 0.15])]
            _, _, _, _, _, (iz, x0, x1) = m
            reqs.append(one(srv, m[0], kind, int(rng.integers(x0, x1)), rng, 0.0 if burst else float(t)))
        return reqs, 0.0 if burst else float(arrivals[-1])

    opts = {k: int(v) for k, v in (kv.split("=") for kv in a.opt)}

    def run(slots, max_shots, threads, burst):
        srv = Server(max_models=4, max_shots=max_shots, slots=slots, options=opts)
        for mid, params, vel, rec, wav, _ in models:
            srv.add_model(mid, params, vel, rec, wav)
        rng = np.random.default_rng(7)
        # untimed: one call of every (model, kind) on each slot's context, so
        # buffers and geometries exist
        warm = [one(srv, m[0], kind, m[5][1], rng) for m in models for kind in kinds for _ in range(slots)]
        srv.max_shots = 1                  # one call per warm-up request, spread over the slots
        if threads:
            srv.start()
        for r in warm:
            srv.submit(r)
        if threads:
            srv.drain()
        else:
            while srv.pending():
                srv.step()
        srv.max_shots = max_shots
        reqs, span = stream(srv, burst)
        calls0 = srv.calls
        t = run_stream(srv, reqs)
        srv.close()
        lat = np.array([r.done - r.arrival for r in reqs]) * 1e3
        print(f"serve: slots {slots}, max shots {max_shots}, {'burst' if burst else 'stream'}: "
              f"{len(reqs) / t:.1f} shots/s over {t:.2f} s", flush=True)
        return {"shots_per_s": len(reqs) / t, "s": t, "stream_s": span, "calls": srv.calls - calls0,
                "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                "max_calls_running": srv.max_running,
                "errors": sum(r.error is not None for r in reqs),
                "finite": all(r.result is not None and np.isfinite(r.result).all() for r in reqs)}, reqs

    cap_b, reqs = run(a.slots, 32, True, True)
    cap_p, _ = run(1, 1, False, True)
    lat_b, _ = run(a.slots, 32, True, False)
    by_kind = {k: sum(r.kind == k for r in reqs) for k in kinds}
    out = {"metric": "shots/s at capacity, serving mix (NL, Born fwd/adj, 21-offset extended Born adj), "
                     "4 models 200x400-1000x3000",
           "value": cap_b["shots_per_s"], "unit": "shots/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": cap_b["s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "serve_mix_4models", "requests": len(reqs), "arrival_rate_per_s": a.rate,
                      "requests_by_kind": by_kind, "slots": a.slots, "max_shots_per_call": 32,
                      "hbm_budget_gb": 200, "engine_options": opts},
           "latency_ms": lat_b["latency_ms"],
           "capacity_continuous_batching": cap_b,
           "capacity_plain_one_request_at_a_time": cap_p,
           "speedup_vs_plain": cap_b["shots_per_s"] / cap_p["shots_per_s"],
           "stream_continuous_batching": lat_b,
           "finite": cap_b["finite"] and cap_p["finite"] and lat_b["finite"],
           "errors": cap_b["errors"] + cap_p["errors"] + lat_b["errors"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
