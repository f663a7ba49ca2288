#!/usr/bin/env python3
"""Throughput of the 2M-particle density+force step on one MI355X.

Workload: 3D anisotropic diffusion, 128^3 = 2,097,152 particles on the
staggered close-packed-like lattice in the periodic unit cube, M6 kernel,
`fab` second derivative, kappa = kappa_par b b + kappa_perp (I - b b) with
b = (1,1,0)/sqrt(2), kappa_par = 1, kappa_perp = 1e-3, T = sin sin sin,
Newton h-iteration in FP64.  One "step" is the reference's iterate()
(iterator.f90:44): neighbour search + loop 1 (Newton sweeps) + EOS + loop 2,
on particles resident in HBM.

Metric: pair-interactions/s = P / step time, P = sum_i |N_i| (ordered pairs
within krad*h_i, periodic images included).  Both paths are timed: the FP64
path (reference arithmetic) and the mixed path (FP32 pair math, FP64
separations/Newton/sums); `value` is the mixed path, whose density error
against the FP64 path is reported and must stay <= 1e-5.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic work per pair / per particle (fixed figures used for the roofline)
FLOP_PER_PAIR_LOOP2_FAB = 84      # loop 2, fab, per-particle kappa
FLOP_PER_PAIR_LOOP1_GRAD = 50     # loop 1 sweep with grad T
PEAK_FP64_TFLOPS = 78.6           # MI355X FP64 (vector and dense matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the 128^3 workload's kernels, measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE in separate runs, scaled by what the counters read
# for a plain 1 GiB copy (x2 and x1): profiles/hbm_traffic_diffusion3d_128cube.md
MEASURED_TRAFFIC_128 = {"neigh": 1.204e9, "loop1_mixed": 1.525e9, "loop2_mixed": 1.632e9}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_path(state, st, precision, steps, warmup, torch):
    from sph_playground_amd.engine import Engine
    eng = Engine(state, kernel="m6", precision=precision, tensor="full", kappa_perp=1e-3)
    eng.load(st)
    for _ in range(warmup):
        eng.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the per-kernel breakdown (the library's HIP events) from a few more
    # steps of the same run, outside the timed region
    per = []
    for _ in range(min(steps, 5)):
        eng.step()
        per.append(eng.stats())
    return eng, per, t1 - t0


def adjoint_cost(size=64, s=5, nsteps=20, every=5, reps=3):
    """Adjoint/forward cost of the differentiable diffusion: 64^3 particles,
    20 RKL2 super-steps of s stages, checkpoints every 5 steps, loss
    0.5 |T - T*|^2 with a seeded target; gradients w.r.t. kappa_par,
    kappa_perp, the field angles and T0."""
    import math
    import numpy as np
    import torch
    from sph_playground_amd import autograd as AG
    from sph_playground_amd import consts as C
    from sph_playground_amd import problems as P
    from sph_playground_amd.engine import Engine
    state, st = P.diffusion3d(size, ddw="fab", b=(1.0, 1.0, 0.0), disotropic="no")
    eng = Engine(state, kernel="m6", tensor="full", kappa_perp=1e-3)
    eng.load(st)
    AG.prepare(eng)
    T0 = torch.from_numpy(st[:, C.es_t - 1].copy()).cuda().requires_grad_(True)
    target = torch.from_numpy(P.gaussian(5, st.shape[0]) * 0.1).cuda()
    params = [torch.tensor(v, dtype=torch.float64, requires_grad=True)
              for v in (1.0, 1e-3, math.pi / 2, math.pi / 4)]
    dt = 1e-6
    fw, bw = [], []
    for _ in range(reps):
        K = AG.conduction_tensor(*params)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        T = AG.sts(T0, K, eng, s, dt, nsteps, every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss = 0.5 * torch.sum((T - target) ** 2)
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fw.append(t1 - t0)
        bw.append(t2 - t1)
    f, b = float(np.median(fw)), float(np.median(bw))
    return {"workload": "%d^3 = %d particles, %d super-steps x %d stages, checkpoint every %d, FP64"
                        % (size, size ** 3, nsteps, s, every),
            "forward_ms": f * 1e3, "backward_ms": b * 1e3, "adjoint_over_forward": b / f}


def batched_cost(nprob=256, s=10, nsteps=5, reps=3):
    """The batched 2D anisotropic-diffusion study: nprob problems with N
    log-uniform in [32^2, 512^2] on periodic square lattices, field angle
    U[0, pi), kappa_perp / kappa_par in {1, 0.1, 0.01, 0.001}, sine T, M6, fab;
    nsteps RKL2 super-steps of s stages over all problems (one launch per
    stage, a replayed hipGraph per super-step), against the same super-steps
    run problem by problem (sph_diff_forward)."""
    import ctypes as ct
    import numpy as np
    import torch
    from sph_playground_amd import autograd as AG
    from sph_playground_amd import consts as C
    from sph_playground_amd import problems as P
    from sph_playground_amd._lib import check, lib
    from sph_playground_amd.batch import Batch
    from sph_playground_amd.engine import Engine
    specs = P.batch2d_specs(nprob)
    engines, Ks, T0s, dts = [], [], [], []
    t0 = time.perf_counter()
    for k, sp in enumerate(specs):
        state, st, K = P.batch2d_problem(sp)
        eng = Engine(state, kernel="m6", tensor="full", kappa_perp=sp["ratio"])
        eng.load(st)
        AG.prepare(eng)
        engines.append(eng)
        Ks.append(K)
        T0s.append(st[:, C.es_t - 1].copy())
        dx = 1.0 / sp["n_side"]
        dts.append(0.05 * dx * dx * (s * s + s - 2) / 4.0)
        if (k + 1) % 64 == 0:
            log("batch: %d problems prepared" % (k + 1))
    t_prep = time.perf_counter() - t0
    batch = Batch(engines, np.stack(Ks))
    info = batch.info()
    T0 = torch.from_numpy(np.concatenate(T0s)).cuda()
    out = torch.empty_like(T0)
    res = {}
    for mode, graph in (("graph", True), ("direct", False)):
        batch.sts(T0, s, dts, nsteps, graph=graph, out=out)  # warm-up (captures the graph)
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            a = time.perf_counter()
            batch.sts(T0, s, dts, nsteps, graph=graph, out=out)
            ts.append(time.perf_counter() - a)
        res[mode] = float(np.median(ts))
    # the same super-steps problem by problem
    T0p = [torch.from_numpy(x).cuda() for x in T0s]
    outp = [torch.empty_like(x) for x in T0p]
    Kh = [np.ascontiguousarray(K) for K in Ks]

    def per_problem():
        for p in range(nprob):
            check(lib().sph_diff_forward(engines[p]._h, Kh[p].ctypes.data_as(ct.POINTER(ct.c_double)), s, dts[p],
                                         nsteps, nsteps, ct.c_void_p(T0p[p].data_ptr()),
                                         ct.c_void_p(outp[p].data_ptr())), "sph_diff_forward")
    per_problem()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        per_problem()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - a)
    t_single = float(np.median(ts))
    same = all(torch.equal(x, y) for x, y in zip(batch.split(out), outp))
    # the same study with FP32-stored pair weights
    b32 = Batch(engines, np.stack(Ks), precision="mixed")
    out32 = torch.empty_like(T0)
    b32.sts(T0, s, dts, nsteps, graph=True, out=out32)
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        b32.sts(T0, s, dts, nsteps, graph=True, out=out32)
        ts.append(time.perf_counter() - a)
    t32 = float(np.median(ts))
    dev32 = max(float((x - y).abs().max() / max((y - z).abs().max(), 1e-300))
                for x, y, z in zip(b32.split(out32), batch.split(out), batch.split(T0)))
    info32 = b32.info()
    b32.close()
    P_ = info["n_pairs"]
    N_ = info["n_total"]
    inter = P_ * s * nsteps
    stage_ms = res["graph"] * 1e3 / (s * nsteps)
    byts = 12 * P_ + 80 * N_
    batch.close()
    for e in engines:
        e.close()
    return {"workload": "%d 2D anisotropic-diffusion problems, N log-uniform in [32^2, 512^2] (total %d), "
                        "theta U[0,pi), kappa_perp/kappa_par in {1,0.1,0.01,0.001}, M6, fab, FP64, "
                        "%d super-steps x %d stages" % (nprob, N_, nsteps, s),
            "n_total": N_, "pairs_per_stage": P_, "pair_interactions_per_s": inter / res["graph"],
            "ms_graph": res["graph"] * 1e3, "ms_direct": res["direct"] * 1e3,
            "ms_problem_by_problem": t_single * 1e3, "speedup_vs_problem_by_problem": t_single / res["graph"],
            "bitwise_equal_to_problem_by_problem": same,
            "stage_ms": stage_ms, "stage_algorithmic_bytes": byts,
            "stage_hbm_gbs": byts / (stage_ms * 1e-3) / 1e9,
            "autotune_block": info["block"], "autotune_ms_per_4_stages": info["tune"],
            "prepare_s": t_prep,
            "fp32_weights": {"ms_graph": t32 * 1e3, "pair_interactions_per_s": inter / t32,
                             "stage_ms": t32 * 1e3 / (s * nsteps),
                             "stage_algorithmic_bytes": 8 * P_ + 80 * N_,
                             "stage_hbm_gbs": (8 * P_ + 80 * N_) / (t32 / (s * nsteps)) / 1e9,
                             "max_dev_rel_to_change_vs_fp64": dev32, "autotune_block": info32["block"]}}


def cpu_baseline(size):
    """Reference Fortran hot path (oracle/_ref, built from the reference
    sources) on a bounded sample of the same workload, all host cores."""
    from oracle import refrun
    from sph_playground_amd import problems as P
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    state, st = P.diffusion3d(size, ddw="fab", b=(1.0, 1.0, 0.0), disotropic="no")
    if refrun.available(1):
        r = refrun.run(state, st, kernel=1, mode=103, threads=cores)
        pairs = int(r["n_pairs"])
        return dict(value=pairs / r["seconds"], unit="pair-interactions/s", cores=cores, kind="reference",
                    sample="%d^3 = %d particles of the same lattice/kernel/fab workload (reference diag(b^2) "
                           "tensor), second iterate() of the reference OpenMP build, %.3f s, P=%d"
                           % (size, size ** 3, r["seconds"], pairs))
    from oracle import oracle
    t0 = time.perf_counter()
    o = oracle.iterate(state, st, kernel=1, tensor_mode=1, kappa_perp=1e-3)
    dt = time.perf_counter() - t0
    pairs = int(o["off"][-1])
    return dict(value=pairs / dt, unit="pair-interactions/s", cores=cores, kind="port",
                sample="%d^3 particles, C restatement, %.3f s" % (size, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--cpu-size", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the step runs on one GPU; runs on more than one GPU are not supported "
              "(pass --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from sph_playground_amd import consts as C
    from sph_playground_amd import problems as P

    state, st = P.diffusion3d(args.size, ddw="fab", b=(1.0, 1.0, 0.0), disotropic="no")
    n = st.shape[0]
    log("workload %d particles; FP64 path" % n)
    e64, per64, t64 = run_path(state, st, "fp64", args.steps, args.warmup, torch)
    log("mixed path")
    emx, permx, tmx = run_path(state, st, "mixed", args.steps, args.warmup, torch)
    pairs = per64[-1]["n_pairs"]
    assert all(p["n_pairs"] == pairs for p in per64 + permx), "pair counts differ between paths"
    v64 = pairs * args.steps / t64
    vmx = pairs * args.steps / tmx
    out64 = e64.save()
    outmx = emx.save()
    den64 = out64[:, C.es_den - 1]
    den_err = float(np.max(np.abs(outmx[:, C.es_den - 1] - den64) / np.abs(den64)))
    ddt64 = out64[:, C.es_ddt - 1]
    ddt_err = float(np.max(np.abs(outmx[:, C.es_ddt - 1] - ddt64)) / np.max(np.abs(ddt64)))

    def avg(per, k):
        return float(np.mean([p[k] for p in per]))

    k64 = {k: avg(per64, k) for k in ("ms_k_neigh", "ms_k_loop1", "ms_k_loop2", "ms_total")}
    kmx = {k: avg(permx, k) for k in ("ms_k_neigh", "ms_k_loop1", "ms_k_loop2", "ms_total")}
    sweeps = avg(permx, "sweeps")

    def traffic(key):  # measured HBM bytes per launch (only for the default 128^3 workload)
        return MEASURED_TRAFFIC_128[key] if args.size == 128 else None

    # dominant kernel of the headline (mixed) path
    dom = max(("ms_k_neigh", "ms_k_loop1", "ms_k_loop2"), key=lambda k: kmx[k])
    if dom == "ms_k_loop2":
        flops = pairs * FLOP_PER_PAIR_LOOP2_FAB
        ach = flops / (kmx[dom] * 1e-3) / 1e12
        roof = dict(kernel="k_loop2_fab<3,M6,float> (loop 2, fab, full tensor)", bound="valu", achieved=ach,
                    peak=PEAK_FP32_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP32_TFLOPS, traffic=traffic("loop2_mixed"),
                    flop_per_pair=FLOP_PER_PAIR_LOOP2_FAB, pairs_per_launch=pairs, ms_per_launch=kmx[dom])
    elif dom == "ms_k_loop1":
        p1 = avg(permx, "loop1_pairs") / max(sweeps, 1)
        flops = p1 * FLOP_PER_PAIR_LOOP1_GRAD
        ms = kmx[dom] / max(sweeps, 1)
        ach = flops / (ms * 1e-3) / 1e12
        roof = dict(kernel="k_loop1<3,M6,grad,float> (one Newton sweep)", bound="valu", achieved=ach,
                    peak=PEAK_FP32_TFLOPS, unit="TFLOP/s", frac=ach / PEAK_FP32_TFLOPS, traffic=traffic("loop1_mixed"),
                    flop_per_pair=FLOP_PER_PAIR_LOOP1_GRAD, pairs_per_launch=p1, ms_per_launch=ms)
    else:
        # neighbour build: algorithmic bytes = read x,y,z,m,h (40 B) + write the list (4 B/pair) + count
        byts = n * 44 + pairs * 4
        ach = byts / (kmx[dom] * 1e-3) / 1e9
        roof = dict(kernel="k_neigh<3> (cell-list neighbour build)", bound="hbm", achieved=ach,
                    peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS, traffic=traffic("neigh"),
                    bytes_per_launch=byts, ms_per_launch=kmx[dom])

    adj = None
    if not args.no_adjoint:
        log("adjoint cost")
        adj = adjoint_cost()

    bat = None
    if not args.no_batch:
        log("batched problems")
        bat = batched_cost()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        try:
            cpu = cpu_baseline(args.cpu_size)
        except Exception as ex:  # reported, not fatal
            cpu = dict(value=None, unit="pair-interactions/s", cores=None, kind=None, sample="failed: %s" % ex)

    line = {
        "metric": "pair-interactions/s, 2M-particle density+force step; mixed vs FP64; adjoint/fwd",
        "value": vmx,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmx * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic (f64 positions, separations, Newton h, sums)",
        "data": "synthetic (seeded lattice, T = sin sin sin)",
        "config": {"workload": "3D anisotropic diffusion density+force step, %d^3 = %d particles, M6, fab, "
                               "kappa_par=1 kappa_perp=1e-3 b=(1,1,0)/sqrt2, periodic unit cube" % (args.size, n),
                   "pairs_per_step": pairs, "newton_sweeps_per_step": sweeps},
        "fp64": {"value": v64, "ms_per_step": t64 * 1e3 / args.steps, "dtype": "f64", "kernels_ms": k64},
        "mixed": {"value": vmx, "kernels_ms": kmx, "density_max_rel_err_vs_fp64": den_err,
                  "ddt_max_err_rel_to_max": ddt_err},
        "mixed_vs_fp64": vmx / v64,
        "adjoint": adj,
        "batched": bat,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
