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

# algorithmic work per pair / per particle (add/sub/mul/div/sqrt = 1 flop,
# fma = 2; per-particle constants hoisted)
FLOP_PER_PAIR_LOOP2_FAB = 72      # loop 2, fab, one tensor for every pair (84 with a per-particle one)
FLOP_PER_PAIR_LOOP1_GRAD = 50     # loop 1 sweep with grad T
# compulsory HBM bytes per particle of one step of this workload: search
# (read x, y, z, h; write key and index: 40), loop 1 sweep with grad T (read
# x, y, z, m, h, T, rho, Omega; write rho, Omega, h, eps: 92), loop 2 fab with
# one tensor (read x, y, z, m, rho, h, T; write dT/dt: 64)
COMPULSORY_BYTES_PER_PARTICLE = {"search": 40, "loop1_per_sweep": 92, "loop2": 64}
PEAK_FP64_TFLOPS = 78.6           # MI355X FP64 (vector and dense matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
GPU_CLOCK_HZ = 2.4e9              # MI355X peak engine clock
N_SIMD = 1024                     # 256 CUs x 4 SIMDs
N_TA = 256                        # one texture address unit per CU
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILE = "hbm_traffic_diffusion3d_128cube.md"
COUNTERS_FILE = "counters_diffusion3d_128cube.md"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def profile_table(name):
    """Rows of the first markdown table in profiles/<name>: {kernel: [cells]},
    the kernel name with its template arguments as rocprofv3 prints it."""
    rows = {}
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if line.startswith("| `"):
                    cells = [c.strip() for c in line.strip().strip("|").split("|")]
                    rows[cells[0].strip("`")] = cells[1:]
    except OSError:
        pass
    return rows


def kernel_roofs(kname, ms, alg_bytes, alg_flops, peak_tflops):
    """Both roofs of one kernel at its measured launch time: algorithmic and
    measured HBM bytes (FETCH_SIZE/WRITE_SIZE per launch, copy-calibrated, from
    the traffic profile) against 8 TB/s; algorithmic flops against the VALU
    peak; VALU and address-unit busy from the counter profile (a wave64 VALU
    instruction occupies its SIMD 4 cycles)."""
    t = ms * 1e-3
    out = {"ms_per_launch": ms, "algorithmic_bytes": alg_bytes,
           "algorithmic_hbm_frac": alg_bytes / t / (PEAK_HBM_GBS * 1e9)}
    tr = profile_table(TRAFFIC_FILE).get(kname)
    if tr:
        out["measured_bytes"] = float(tr[-1])
        out["measured_hbm_frac"] = out["measured_bytes"] / t / (PEAK_HBM_GBS * 1e9)
    if alg_flops:
        out["algorithmic_flops"] = alg_flops
        out["algorithmic_tflops"] = alg_flops / t / 1e12
        out["valu_peak_tflops"] = peak_tflops
        out["algorithmic_valu_frac"] = alg_flops / t / 1e12 / peak_tflops
    ct = profile_table(COUNTERS_FILE).get(kname)
    if ct:
        cyc = t * GPU_CLOCK_HZ
        out["valu_busy"] = float(ct[1]) * 4.0 / (N_SIMD * cyc)
        out["ta_busy"] = float(ct[6]) / (N_TA * cyc)
    busy = {"hbm": out.get("measured_hbm_frac", out["algorithmic_hbm_frac"]), "valu": out.get("valu_busy", 0.0),
            "address_unit": out.get("ta_busy", 0.0)}
    out["measured_bound"] = max(busy, key=busy.get)
    return out


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
    # the per-kernel breakdown (the library's HIP events, recorded only with
    # timing on) from a few more steps of the same run, outside the timed region
    eng.set_timing(True)
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
    from sph_playground_amd import batch as B
    specs = P.batch2d_specs(nprob)
    built = [P.batch2d_problem(sp) for sp in specs]
    triples = [(state, st, sp["ratio"]) for (state, st, _), sp in zip(built, specs)]
    Ks = [K for _, _, K in built]
    T0s = [st[:, C.es_t - 1].copy() for _, st, _ in built]
    dts = [0.05 * (1.0 / sp["n_side"]) ** 2 * (s * s + s - 2) / 4.0 for sp in specs]
    t0 = time.perf_counter()
    cells, cell_table = B.tune_cells(triples)
    t_tune = time.perf_counter() - t0
    log("batch: search cells per size bucket %s (tuned in %.2f s)" % (cells, t_tune))
    engines, t_prep = B.prepare_problems(triples, cells)
    log("batch: %d problems prepared in %.2f s" % (nprob, t_prep))
    # the same preparation one problem after another on one host thread
    eng1, t_prep_serial = B.prepare_problems(triples, cells, threads=1)
    for e in eng1:
        e.close()
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
            "prepare_s": t_prep, "prepare_s_one_thread": t_prep_serial,
            "cell_autotune": {"cells_r_over": {str(4 ** b * 1024): d for b, d in cells.items()},
                              "step_ms": {str(4 ** b * 1024): {str(d): ms for d, ms in t.items()}
                                          for b, t in cell_table.items()},
                              "tune_s": t_tune},
            "fp32_weights": {"ms_graph": t32 * 1e3, "pair_interactions_per_s": inter / t32,
                             "stage_ms": t32 * 1e3 / (s * nsteps),
                             "stage_algorithmic_bytes": 8 * P_ + 80 * N_,
                             "stage_hbm_gbs": (8 * P_ + 80 * N_) / (t32 / (s * nsteps)) / 1e9,
                             "max_dev_rel_to_change_vs_fp64": dev32, "autotune_block": info32["block"]}}


def rel_error_stats(ref, got, floor=1e-3):
    """Max and RMS over particles of |got - ref| / max(|ref|, floor rms(ref))."""
    import numpy as np
    rms = float(np.sqrt(np.mean(ref * ref)))
    if rms == 0.0:
        d = np.abs(got - ref)
        return {"max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d)))}
    d = np.abs(got - ref) / np.maximum(np.abs(ref), floor * rms)
    return {"max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d)))}


def field_errors(ref, got, nr, names):
    from sph_playground_amd import consts as C
    out = {}
    for f in names:
        k = getattr(C, "es_" + f) - 1
        if float(abs(ref[:nr, k]).max()) > 0.0:
            out[f] = rel_error_stats(ref[:nr, k], got[:nr, k])
    return out


def moving_cost(n=1048576, steps=20, warmup=3):
    """Steps in which particles move: leapfrog passes (main.f90:241-362:
    time step, predictor, iterate, corrector) of the 3D shock tube, 1,048,576
    uniform-random particles in [-1,1] x [0,0.125]^2, periodic in y and z,
    fixed x walls, B = (0.75, 1, 0), opacity 1; its MHD half and its FLD half,
    both precisions.  Every pass wraps, re-keys and checks the cell order of
    the moved particles (and re-sorts them when the order broke) before the
    search.  Pairs per pass are the iterate's P."""
    import numpy as np
    import torch
    from sph_playground_amd import problems as P
    from sph_playground_amd.engine import Engine
    out = {"workload": "3D shock tube, %d uniform-random particles in [-1,1]x[0,0.125]^2, periodic y/z, "
                       "fixed x walls, B=(0.75,1,0), opacity 1, LP limiter; %d leapfrog passes after %d"
                       % (n, steps, warmup)}
    for half in ("mhd", "fld"):
        state, st = P.shock_tube3d(n, half=half, velocity_seed=2)
        res = {}
        saved = {}
        for prec in ("fp64", "mixed"):
            eng = Engine(state, kernel="m6", precision=prec, neigh_cap=400)
            eng.load(st)
            for _ in range(warmup):
                eng.leapfrog(1.0)
            pairs, sorted_, sweeps = [], 0, []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                eng.leapfrog(1.0)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            s = eng.stats()
            saved[prec] = eng.save()
            res[prec] = {"ms_per_step": t * 1e3 / steps, "pairs_per_step": s["n_pairs"],
                         "pair_interactions_per_s": s["n_pairs"] * steps / t,
                         "newton_sweeps_last_step": s["sweeps"]}
            eng.close()
        nr = int(state[24])
        res["mixed_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
        res["mixed_field_errors_vs_fp64_after_%d_steps" % (warmup + steps)] = field_errors(
            saved["fp64"], saved["mixed"], nr, ("den", "h", "om", "ax", "ay", "az", "du", "dh", "dbx", "dby",
                                                 "dbz", "ddt", "dtdx", "vfld"))
        out[half] = res
    return out


def sts_cost(size=128, s=5, nsteps=10):
    """The anisotropic-diffusion workload's 10 RKL2 super-steps
    (sts_integrator.f90:87-124, every stage a full iterate) at its size, both
    precisions, against the closed-form decay of its two eigenmodes."""
    import numpy as np
    import torch
    from sph_playground_amd import consts as C
    from sph_playground_amd import problems as P
    from sph_playground_amd.engine import Engine
    state, st = P.diffusion3d(size, ddw="fab", b=(1.0, 1.0, 0.0), disotropic="no")
    h = float(st[0, C.es_h - 1])
    dt = (s * s + s - 2) / 4.0 * 0.2 * h * h
    T0 = st[:, C.es_t - 1]
    exact = P.anisotropic_decay_exact(st[:, 0:3], nsteps * dt)
    change = float(np.abs(exact - T0).max())
    res = {"workload": "%d^3 anisotropic diffusion, %d RKL2 super-steps x %d stages, dt = %.3e "
                       "((s^2+s-2)/4 x 0.2 h^2 / kappa_par)" % (size, nsteps, s, dt),
           "t_final": nsteps * dt, "change_of_T": change}
    outs = {}
    for prec in ("fp64", "mixed"):
        eng = Engine(state, kernel="m6", precision=prec, tensor="full", kappa_perp=1e-3)
        eng.load(st)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            eng.sts(s, dt)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        T = eng.save()[:, C.es_t - 1]
        eng.close()
        outs[prec] = T
        d = T - exact
        res[prec] = {"ms_per_super_step": t * 1e3 / nsteps,
                     "err_vs_closed_form_max_over_change": float(np.abs(d).max()) / change,
                     "err_vs_closed_form_rms_over_change": float(np.sqrt(np.mean(d * d))) / change}
    res["mixed_vs_fp64_max_over_change"] = float(np.abs(outs["mixed"] - outs["fp64"]).max()) / change
    return res


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
    ap.add_argument("--no-moving", action="store_true")
    ap.add_argument("--no-sts", action="store_true")
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

    # both roofs of the three main kernels of each path; the profiles hold
    # the counters of the 128^3 workload, so smaller sizes report no
    # measured bytes or busy fractions
    sweeps64 = avg(per64, "sweeps")
    p1 = avg(permx, "loop1_pairs") / max(sweeps, 1)
    p164 = avg(per64, "loop1_pairs") / max(sweeps64, 1)
    names = {"mixed": ("k_neigh<3, true>", "k_loop1<3, 1, 1, float>", "k_loop2_fab<3, 1, float, HIP_vector_type<float, 4u> >"),
             "fp64": ("k_neigh<3, true>", "k_loop1<3, 1, 1, double>",
                      "k_loop2_fab<3, 1, double, HIP_vector_type<double, 4u> >")}
    roofs = {}
    for path, km, sw, pp, peak in (("mixed", kmx, sweeps, p1, PEAK_FP32_TFLOPS),
                                   ("fp64", k64, sweeps64, p164, PEAK_FP64_TFLOPS)):
        nk, n1, n2 = names[path] if args.size == 128 else ("", "", "")
        roofs[path] = {
            "k_neigh": kernel_roofs(nk, km["ms_k_neigh"], n * 44 + pairs * 4, 0, peak),
            "k_loop1": kernel_roofs(n1, km["ms_k_loop1"] / max(sw, 1),
                                    n * COMPULSORY_BYTES_PER_PARTICLE["loop1_per_sweep"],
                                    pp * FLOP_PER_PAIR_LOOP1_GRAD, peak),
            "k_loop2_fab": kernel_roofs(n2, km["ms_k_loop2"], n * COMPULSORY_BYTES_PER_PARTICLE["loop2"],
                                        pairs * FLOP_PER_PAIR_LOOP2_FAB, peak)}
    # traffic of a whole steady mixed step: every kernel of it in the profile
    # (search prep, search, Newton sweep, post, loop 2, close) against the
    # compulsory bytes
    step_kernels = ("k_prep", "k_neigh<3, true>", "k_loop1<3, 1, 1, float>", "k_post1<HIP_vector_type<float, 4u> >",
                    "k_loop2_fab<3, 1, float, HIP_vector_type<float, 4u> >", "k_step_close")
    tr = profile_table(TRAFFIC_FILE)
    step_measured = sum(float(tr[k][-1]) for k in step_kernels if k in tr) if args.size == 128 else None
    cb = COMPULSORY_BYTES_PER_PARTICLE
    step_compulsory = n * (cb["search"] + cb["loop1_per_sweep"] * sweeps + cb["loop2"])
    # the contract's block: the dominant kernel of the headline (mixed) path
    dom = max(("ms_k_neigh", "ms_k_loop1", "ms_k_loop2"), key=lambda k: kmx[k])
    if dom == "ms_k_neigh":
        r = roofs["mixed"]["k_neigh"]
        roof = dict(kernel="k_neigh<3, true> (cell-list neighbour search, extended grid)", bound="hbm",
                    achieved=r["algorithmic_bytes"] / (r["ms_per_launch"] * 1e-3) / 1e9, peak=PEAK_HBM_GBS,
                    unit="GB/s", frac=r["algorithmic_hbm_frac"], traffic=r.get("measured_bytes"),
                    bytes_per_launch=r["algorithmic_bytes"], ms_per_launch=r["ms_per_launch"],
                    measured_bound=r["measured_bound"], valu_busy=r.get("valu_busy"), ta_busy=r.get("ta_busy"))
    else:
        key = "k_loop1" if dom == "ms_k_loop1" else "k_loop2_fab"
        r = roofs["mixed"][key]
        roof = dict(kernel=key + " (FP32 pair loop)", bound="valu", achieved=r["algorithmic_tflops"],
                    peak=PEAK_FP32_TFLOPS, unit="TFLOP/s", frac=r["algorithmic_valu_frac"],
                    traffic=r.get("measured_bytes"), ms_per_launch=r["ms_per_launch"],
                    measured_bound=r["measured_bound"], valu_busy=r.get("valu_busy"), ta_busy=r.get("ta_busy"))
    roof["kernels"] = roofs
    roof["step_measured_bytes"] = step_measured
    roof["step_compulsory_bytes"] = step_compulsory
    roof["profiles"] = [TRAFFIC_FILE, COUNTERS_FILE]

    adj = None
    if not args.no_adjoint:
        log("adjoint cost")
        adj = adjoint_cost()

    sts = None
    if not args.no_sts:
        log("STS super-steps at full size")
        sts = sts_cost(args.size)

    mov = None
    if not args.no_moving:
        log("moving particles (shock tube)")
        mov = moving_cost()

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
        "dtype": "f32 pair arithmetic and per-particle sums (separations from 31-bit fixed-point positions, "
                 "f64 Newton update and energy sum)",
        "data": "synthetic (seeded lattice, T = sin sin sin)",
        "config": {"workload": "3D anisotropic diffusion density+force step, %d^3 = %d particles, M6, fab, "
                               "kappa_par=1 kappa_perp=1e-3 b=(1,1,0)/sqrt2, periodic unit cube" % (args.size, n),
                   "pairs_per_step": pairs, "newton_sweeps_per_step": sweeps},
        "fp64": {"value": v64, "ms_per_step": t64 * 1e3 / args.steps, "dtype": "f64", "kernels_ms": k64},
        "mixed": {"value": vmx, "kernels_ms": kmx, "density_max_rel_err_vs_fp64": den_err,
                  "ddt_max_err_rel_to_max": ddt_err,
                  "field_errors_vs_fp64": field_errors(out64, outmx, n, ("den", "h", "om", "ddt", "dtdx", "dtdy",
                                                                         "dtdz"))},
        "mixed_vs_fp64": vmx / v64,
        "adjoint": adj,
        "batched": bat,
        "moving": mov,
        "sts": sts,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
