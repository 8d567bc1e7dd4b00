"""Throughput of the ladder hot path on one MI355X.

One step = one pass over the synthetic 64x64 workload (U=2, beta=30, n=0.9,
Nν=120, Nω=121, νmax=45) with every input resident in HBM:
    bubble -> BSE (m) -> BSE (d) -> λ₀ -> EoM Σ (with the λ sums of its tail)
Throughput is (q,ω)-points per second (Nq*Nω = 561*121 points per step).

The headline path (--precision, default mixed: FP32 inversion on MFMA with
FP64 iterative refinement) is timed first, then the other path (FP64) on the
same inputs; the JSON carries both, their ratio and the mixed path's relative
error against the FP64 path on Σ, χ and γ (the 1e-10 gate on Σ).

Prints ONE JSON line.  The roofline entry is for the dominant kernel (the BSE),
timed with HIP events on the stream it runs on; the CPU baseline times the
oracle (numpy/scipy FP64) on a bounded sample of the same work, on all the
host cores available (at most 16) and on one.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "ladderdga.jl_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = "BSE+EoM (q,ω)-points/s: mixed+IR vs FP64 GPU at Σ rel err ≤1e-10"
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (MFMA and vector)
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32

# HBM bytes per BSE launch, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this script, each
# scaled by what the counter reads for a 1024 MiB device copy (--copy-calib): fetch x2.000, write x1.000.
# Summary and commands: profiles/bse_mixed_fp64_2d64.txt.
MEASURED_TRAFFIC = {("2d64", "mixed"): 301.5e6, ("2d64", "fp64"): 292.2e6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="2d64")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"],
                    help="the headline path; the other one is timed beside it")
    ap.add_argument("--numax", type=int, default=-1, help="EoM νmax (default: the λdm window min(n_iν, 3Nω/8))")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-points", type=int, default=960, help="(q,ω) points in the CPU BSE sample")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibrates the HBM byte counters)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy of a known size: reads and writes mib MiB each, three launches."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calib_bytes_read": n * 8, "copy_calib_bytes_written": n * 8}), file=_STDOUT)


def _cpu_share(conn, barrier, spec):
    """One process of the CPU baseline: its share of each stage (items i with i % nsplit == rank),
    each repetition timed between barriers shared by all the processes of the run."""
    import numpy as np
    from multiprocessing import shared_memory
    from threadpoolctl import threadpool_limits
    from oracle import ldga_oracle as O

    shms, arr = [], {}
    for key, (name, shape, dt) in spec["arrays"].items():
        m = shared_memory.SharedMemory(name=name)  # the parent owns and unlinks it
        shms.append(m)
        arr[key] = np.ndarray(shape, dtype=dt, buffer=m.buf, order="F")
    D, Ns, t, wedge, U, mu, beta, n, ekin, n_iw, n_iν, shift, numax = spec["params"]
    oK = O.KGrid(D, Ns, t, wedge_coords=wedge)
    omP = O.ModelParameters(U, mu, beta, n, 0, 0, ekin)
    osP = O.SimulationParameters(n_iw, n_iν, 0, shift)
    rank, nsplit, reps = spec["rank"], spec["nsplit"], spec["reps"]
    Nq, Nν, Nω = arr["chi0"].shape
    pts = spec["points"][rank::nsplit]
    ωs = set(range(rank, Nω, nsplit))
    ones = np.broadcast_to(np.complex128(1.0), (Nq, Nν, Nω))
    cm = np.full((Nq, Nω), 0.1)
    times = {"bse": [], "bubble": [], "eom": []}

    def bse():
        for q, w in pts:
            for Γ in (arr["Gm"], arr["Gd"]):
                A = np.asfortranarray(Γ[:, :, w] + np.diag(1.0 / arr["chi0"][q, :, w]))
                inv, _, _ = O.inv_lapack(A)
                _ = np.real(inv.sum()), inv.sum(axis=0)

    stages = {"bse": bse,
              "bubble": lambda: O.calc_bubble(arr["gf"], arr["gr"], oK, omP, osP, ωsel=ωs),
              "eom": lambda: O.calc_Σ_ω(O.EOM["full"], cm, ones, cm, ones, arr["gr"], ones, U, oK, osP, numax,
                                        ωsel=ωs)}
    with threadpool_limits(limits=1):
        for k, f in stages.items():
            for _ in range(reps):
                barrier.wait()
                t0 = time.perf_counter()
                f()
                times[k].append(time.perf_counter() - t0)
    conn.send(times)
    conn.close()
    del arr
    for m in shms:
        m.close()


def _cpu_run(spec0, nproc, nsplit):
    """Run the CPU baseline's stages on nproc processes (ranks 0..nproc-1 of an nsplit-way split);
    per stage and repetition the wall time is the slowest process's."""
    import multiprocessing as mp
    import numpy as np
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(nproc, timeout=300)
    procs, pipes = [], []
    for r in range(nproc):
        a, b = ctx.Pipe(duplex=False)
        p = ctx.Process(target=_cpu_share, args=(b, barrier, dict(spec0, rank=r, nsplit=nsplit)), daemon=True)
        p.start()
        procs.append(p)
        pipes.append(a)
    res = [a.recv() for a in pipes]
    for p in procs:
        p.join(60)
    return {k: float(np.median(np.max([r[k] for r in res], axis=0))) for k in res[0]}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, numax, npoints_sample, χ0, reps=5):
    """Oracle (numpy/scipy FP64) on a bounded sample of the step, on all the host cores this
    process may use (at most 16) and on one core.

    BSE: zgetrf+zgetri on `npoints_sample` random (q,ω) points, both channels.  Bubble and EoM: the
    full oracle passes (their ω loop split over the processes), divided by the number of points.
    One process per core (the LAPACK and FFT calls hold the interpreter lock), one BLAS thread each;
    per stage the median over `reps` repetitions of the slowest process.  The one-core value runs
    one process on one share of the same split and scales it by the number of shares.
    """
    import numpy as np
    from multiprocessing import shared_memory

    cfg = inp.cfg
    Nq, Nω = inp.kG.Nq, 2 * cfg.n_iω + 1
    npts = Nq * Nω
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    rng = np.random.default_rng(0)
    points = list(zip(rng.integers(0, Nq, npoints_sample).tolist(), rng.integers(0, Nω, npoints_sample).tolist()))
    src = {"gf": inp.G_fft, "gr": inp.G_rfft, "Gm": inp.Γ_m, "Gd": inp.Γ_d, "chi0": χ0}
    shms, arrays = [], {}
    try:
        for k, a in src.items():
            a = np.asfortranarray(a)
            m = shared_memory.SharedMemory(create=True, size=max(a.nbytes, 1))
            shms.append(m)
            np.ndarray(a.shape, dtype=a.dtype, buffer=m.buf, order="F")[...] = a
            arrays[k] = (m.name, a.shape, a.dtype.str)
        mp_ = inp.mP
        spec0 = {"arrays": arrays, "points": points, "reps": reps,
                 "params": (cfg.D, cfg.Ns, cfg.t, inp.kG.wedge, mp_.U, mp_.μ, mp_.β, mp_.n, mp_.Ekin_DMFT,
                            cfg.n_iω, cfg.n_iν, cfg.shift, numax)}
        t_all = _cpu_run(spec0, cores, cores)
        t_one = {k: v * cores for k, v in _cpu_run(spec0, 1, cores).items()}
    finally:
        for m in shms:
            m.close()
            m.unlink()

    def rate(t):
        return 1.0 / (t["bse"] / npoints_sample + (t["bubble"] + t["eom"]) / npts)

    per_stage = {f"{k}_{lab}_s": round(t[k], 4) for lab, t in (("all_cores", t_all), ("one_core", t_one)) for k in t}
    return {"value": rate(t_all), "unit": "(q,ω)-points/s", "cores": cores, "kind": "port",
            "value_one_core": rate(t_one), "cpu_model": cpu_model(), "repetitions": reps,
            "stages": per_stage,
            "sample": f"oracle on the {cfg.Ns}^{cfg.D} workload, {cores} processes x 1 BLAS thread: zgetrf+zgetri "
                      f"(both channels) on {npoints_sample} random (q,ω) points + the full FFT bubble + the full FFT "
                      f"EoM at νmax={numax} (ω loops split over the processes), per point over {npts}; median of "
                      f"{reps}; one-core value from one process on 1/{cores} of the same work"}


def lambdadm_step_costs(L, ctx, o, λ0, inp, numax, Nq, Nν, Nω, reps=5):
    """Cost of one Newton step of λdm_correction on this step's ladder arrays.

    The reference's 2D Newton (LambdaCorrection/helpers.jl:160-208) takes a
    forward finite-difference Jacobian: three residual evaluations per step.
    The adjoint gives the residual and its Jacobian in one evaluation.  Both
    are timed at a λ inside the admissible region (2 |λ_min| + 1 per channel).
    """
    import numpy as np
    Fo = lambda a, shape: np.asfortranarray(a.cpu().numpy().reshape(shape, order="F"))
    χm = L.χT(Fo(o["χm"], (Nq, Nω)), inp.mP.β, tail_c=[0, 0, inp.mP.Ekin_DMFT])
    χd = L.χT(Fo(o["χd"], (Nq, Nω)), inp.mP.β, tail_c=[0, 0, inp.mP.Ekin_DMFT])
    γm, γd = L.γT(Fo(o["γm"], (Nq, Nν, Nω))), L.γT(Fo(o["γd"], (Nq, Nν, Nω)))
    prob = L.LambdaDmProblem(χm, γm, χd, γd, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, Fo(λ0, (Nq, Nν, Nω)), inp.kG,
                             inp.mP, inp.sP, numax, ctx=ctx)
    lam = [2 * abs(L.get_λ_min(χm.data)) + 1.0, 2 * abs(L.get_λ_min(χd.data)) + 1.0]
    prob.eval(lam)
    prob.eval(lam, want_jac=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        prob.eval(lam)
    t_res = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        r = prob.eval(lam, want_jac=True)
    t_jac = (time.perf_counter() - t0) / reps
    prob.free()
    return {"lambda": lam, "residual_ms": t_res * 1e3, "residual_and_adjoint_jacobian_ms": t_jac * 1e3,
            "fd_newton_step_ms": 3 * t_res * 1e3, "adjoint_newton_step_ms": t_jac * 1e3,
            "step_speedup": 3 * t_res / t_jac, "mu_iters": int(r["mu_iters"])}


def lambdadm_convergence(L, ctx, o, λ0, inp, numax, Nq, Nν, Nω):
    """λdm_correction (conditions_new.jl:207-324) solved to convergence from λ = (0, 0) on this step's
    ladder arrays, once with the reference's forward finite-difference Jacobian (three residuals per
    Newton step) and once with the adjoint Jacobian; wall time of the whole call (including the one
    upload of χ, γ, λ₀), Newton steps, residual evaluations and whether the root was found."""
    import numpy as np
    Fo = lambda a, shape: np.asfortranarray(a.cpu().numpy().reshape(shape, order="F"))
    χm = L.χT(Fo(o["χm"], (Nq, Nω)), inp.mP.β, tail_c=[0, 0, inp.mP.Ekin_DMFT])
    χd = L.χT(Fo(o["χd"], (Nq, Nω)), inp.mP.β, tail_c=[0, 0, inp.mP.Ekin_DMFT])
    γm, γd = L.γT(Fo(o["γm"], (Nq, Nν, Nω))), L.γT(Fo(o["γd"], (Nq, Nν, Nω)))
    l0 = Fo(λ0, (Nq, Nν, Nω))
    out = {}
    for jac in ("fd", "adjoint"):
        t0 = time.perf_counter()
        r = L.λdm_correction(χm, γm, χd, γd, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, l0, inp.kG, inp.mP, inp.sP,
                             νmax=numax, jacobian=jac, ctx=ctx)
        out[jac] = {"wall_ms": (time.perf_counter() - t0) * 1e3, "converged": bool(r.converged),
                    "newton_steps": int(r.newton_steps), "residual_evals": int(r.residual_evals),
                    "lambda": [float(r.λm), float(r.λd)]}
    return out


def lambdadm_convergence_rooted(L, ctx):
    """The same on a 2D 64x64 system whose λdm conditions have a root (U=2, β=10, n=0.9, Nν=60, Nω=61,
    vertex noise 0.1): its chain on the GPU in FP64, then both Newton variants to convergence."""
    import warnings
    from ladderdga_hip import synthetic
    cfg = synthetic.tiny_config(D=2, Ns=64, n_iν=30, n_iω=30, n=0.9, U=2.0, β=10.0, noise=0.1)
    inp = synthetic.build_inputs(cfg)
    χ0 = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    ch = {}
    for c, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        ch["χ" + c], ch["γ" + c] = L.calc_χγ(c, Γ, χ0, inp.kG, inp.mP, inp.sP, precision="fp64", ctx=ctx)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        λ0 = L.calc_λ0(χ0, inp.F_m, inp.mP, inp.sP, inp.kG, ctx=ctx)
    out = {"workload": "2D square 64x64 U=2 β=10 n=0.9 Nν=60 Nω=61 (vertex noise 0.1)"}
    for jac in ("fd", "adjoint"):
        t0 = time.perf_counter()
        r = L.λdm_correction(ch["χm"], ch["γm"], ch["χd"], ch["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, λ0,
                             inp.kG, inp.mP, inp.sP, jacobian=jac, ctx=ctx)
        out[jac] = {"wall_ms": (time.perf_counter() - t0) * 1e3, "converged": bool(r.converged),
                    "newton_steps": int(r.newton_steps), "residual_evals": int(r.residual_evals),
                    "lambda": [float(r.λm), float(r.λd)]}
    return out


def run_sc_iteration_cost(L, ctx, o, λ0, inp, numax, Nq, Nν, Nω):
    """Cost of one iteration of the self-consistency loop (run_sc!, conditions_new.jl:479-491) on this
    step's ladder arrays: EoM, mixing, μ fit and energies, G-table rebuild, all resident on the device.
    Taken as (median t(9 iterations) - median t(1 iteration)) / 8 over 7 calls each of run_sc type :fix,
    so the one-time uploads and their host-side jitter cancel; conv_abs = 0 keeps the loop from
    stopping early."""
    import numpy as np
    Fo = lambda a, shape: np.asfortranarray(a.cpu().numpy().reshape(shape, order="F"))
    χm = L.χT(Fo(o["χm"], (Nq, Nω)), inp.mP.β, tail_c=[0, 0, inp.mP.Ekin_DMFT])
    χd = L.χT(Fo(o["χd"], (Nq, Nω)), inp.mP.β, tail_c=[0, 0, inp.mP.Ekin_DMFT])
    γm, γd = L.γT(Fo(o["γm"], (Nq, Nν, Nω))), L.γT(Fo(o["γd"], (Nq, Nν, Nω)))
    l0 = Fo(λ0, (Nq, Nν, Nω))
    lam = (2 * abs(L.get_λ_min(χm.data)) + 1.0, 2 * abs(L.get_λ_min(χd.data)) + 1.0)

    def t(maxit):
        t0 = time.perf_counter()
        r = L.run_sc(χm, γm, χd, γd, l0, inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, inp.kG, inp.mP, inp.sP,
                     type="fix", νmax=numax, maxit=maxit, mixing=0.2, conv_abs=0.0, λm=lam[0], λd=lam[1],
                     ctx=ctx)
        # conv_abs = 0 never stops early; a failed μ fit would (NaN μ) and is reported
        return time.perf_counter() - t0, maxit if np.isfinite(r.μ) else 0

    t(1)
    t(9)
    r1 = [t(1) for _ in range(7)]
    r9 = [t(9) for _ in range(7)]
    n1, n9 = min(n for _, n in r1), min(n for _, n in r9)
    t1 = float(np.median([x for x, _ in r1]))
    t9 = float(np.median([x for x, _ in r9]))
    return {"iteration_ms": (t9 - t1) / max(n9 - n1, 1) * 1e3, "iterations_timed": n9 - n1,
            "one_iteration_call_ms": t1 * 1e3, "lambda": list(lam)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the ω-sharded multi-GPU design is described but not launched)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    import ctypes as C

    import numpy as np
    import torch
    import ladderdga_hip as L
    from ladderdga_hip import _lib, synthetic
    from ladderdga_hip.api import _freq, _model

    cfg = synthetic.CONFIGS[args.config]
    inp = synthetic.build_inputs(cfg)
    kG, mP, sP = inp.kG, inp.mP, inp.sP
    Nq, Nν, Nω = kG.Nq, 2 * cfg.n_iν, 2 * cfg.n_iω + 1
    Nk = cfg.Ns ** cfg.D
    Nfft = inp.G_fft.shape[-1]
    numax = args.numax if args.numax > 0 else min(cfg.n_iν, (3 * Nω) // 8)
    dev = torch.device("cuda:0")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev)

    Gf, Gr = up(inp.G_fft), up(inp.G_rfft)
    Γm, Γd, Fm = up(inp.Γ_m), up(inp.Γ_d), up(inp.F_m)
    cplx = torch.complex128
    χ0 = torch.empty(Nq * Nν * Nω, dtype=cplx, device=dev)
    out = {}
    for prec in ("mixed", "fp64"):
        out[prec] = dict(χm=torch.empty(Nq * Nω, dtype=torch.float64, device=dev),
                         χd=torch.empty(Nq * Nω, dtype=torch.float64, device=dev),
                         γm=torch.empty(Nq * Nν * Nω, dtype=cplx, device=dev),
                         γd=torch.empty(Nq * Nν * Nω, dtype=cplx, device=dev),
                         Σ=torch.empty(Nq * numax, dtype=cplx, device=dev))
    λ0 = torch.empty(Nq * Nν * Nω, dtype=cplx, device=dev)

    ctx = L.Context(0)
    ctx.set_kgrid(kG)
    Lb = _lib.lib()
    fr, mo = _freq(sP), _model(mP)
    # a dedicated stream (default priority): the library treats a NULL stream as its own
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    frp, mop = C.byref(fr), C.byref(mo)
    PREC = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}

    def ev():
        return torch.cuda.Event(enable_timing=True)

    def step(prec, rec):
        o = out[prec]
        e = [ev() for _ in range(5)] if rec is not None else None
        if e: e[0].record(stream)
        ctx.check(Lb.ldga_bubble_dev(ctx.handle, frp, mP.β, Gf.data_ptr(), Gr.data_ptr(), χ0.data_ptr(), sp))
        if e: e[1].record(stream)
        for s, Γ, χ, γ in ((1, Γm, o["χm"], o["γm"]), (-1, Γd, o["χd"], o["γd"])):
            ctx.check(Lb.ldga_chi_gamma_dev(ctx.handle, s, PREC[prec], frp, mP.U, mP.β, Γ.data_ptr(), χ0.data_ptr(),
                                            χ.data_ptr(), γ.data_ptr(), None, sp))
        if e: e[2].record(stream)
        ctx.check(Lb.ldga_lambda0_dev(ctx.handle, frp, mP.β, Fm.data_ptr(), χ0.data_ptr(), λ0.data_ptr(), sp))
        if e: e[3].record(stream)
        ctx.check(Lb.ldga_sigma_dev(ctx.handle, frp, mop, numax, 0.0, 0.0, 1, 0, inp.χloc_m_sum, o["χm"].data_ptr(),
                                    o["γm"].data_ptr(), o["χd"].data_ptr(), o["γd"].data_ptr(), λ0.data_ptr(),
                                    Gr.data_ptr(), o["Σ"].data_ptr(), sp))
        if e:
            e[4].record(stream)
            rec.append(e)

    def run(prec, steps, warmup):
        for _ in range(warmup):
            step(prec, None)
        torch.cuda.synchronize()
        rec = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(prec, rec)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ph = np.mean([[a.elapsed_time(b) for a, b in zip(e[:-1], e[1:])] for e in rec], axis=0)
        return wall * 1e3 / steps, dict(zip(["bubble", "bse_2ch", "lambda0", "eom"], ph.tolist()))

    main_prec = args.precision
    other = "fp64" if main_prec == "mixed" else "mixed"
    ms_main, ph_main = run(main_prec, args.steps, args.warmup)
    # above Nν = 128 the FP64 BSE is the global-memory kernel (seconds per
    # step at 32^3): one untimed-warmup-free step, for the accuracy comparison
    slow = other == "fp64" and Nν > 128
    other_steps = 1 if slow else args.steps
    ms_other, ph_other = run(other, other_steps, 0 if slow else args.warmup)
    pts = Nq * Nω

    def rel(a, b):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    errs = {k: rel(out["mixed"][k], out["fp64"][k]) for k in ("Σ", "χm", "χd", "γm", "γd")}
    bse_ms = ph_main["bse_2ch"] / 2                       # one channel per launch
    bse_flops = Nq * Nω * (8.0 / 3.0) * Nν ** 3          # LU-equivalent, one channel
    peak = FP64_PEAK_TFLOPS if main_prec == "fp64" else FP32_PEAK_TFLOPS
    achieved = bse_flops / (bse_ms * 1e-3) / 1e12
    bub_bytes = 2 * Nk * Nfft * 16 + Nq * Nν * Nω * 16
    eom_bytes = 3 * Nq * numax * Nω * 16 + 2 * Nq * Nω * 8 + Nk * (numax + Nω) * 16 + Nq * numax * 16
    result = {
        "metric": METRIC,
        "value": pts / (ms_main * 1e-3),
        "unit": "(q,ω)-points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_main,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64-IR",
        "data": "synthetic (seeded atomic-limit Σ, noisy local vertices)",
        "config": {"workload": f"{cfg.D}D simple-cubic Ns={cfg.Ns} (Nq={Nq}) U={cfg.U} β={cfg.β} n={cfg.n} "
                               f"Nν={Nν} Nω={Nω} νmax={numax} shift={cfg.shift}, plain sums",
                   "precision": main_prec, "points_per_step": pts},
        "roofline": {"bound": "mfma", "kernel": f"bse {main_prec} (one channel per launch)", "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": MEASURED_TRAFFIC.get((args.config, main_prec)), "avg_launch_ms": bse_ms,
                     "flops_per_launch": bse_flops},
        "phases_ms": {k: round(v, 4) for k, v in ph_main.items()},
        "hbm_GBps": {"bubble": bub_bytes / (ph_main["bubble"] * 1e-3) / 1e9,
                     "eom": eom_bytes / (ph_main["eom"] * 1e-3) / 1e9},
        f"{other}_gpu": {"value": pts / (ms_other * 1e-3), "ms_per_step": ms_other, "steps": other_steps,
                         "phases_ms": {k: round(v, 4) for k, v in ph_other.items()}},
        "mixed_vs_fp64_gpu": (ms_other / ms_main) if main_prec == "mixed" else (ms_main / ms_other),
        "mixed_rel_err_vs_fp64": errs,
    }
    # refinement statistics of the timed mixed BSE (one extra call per channel with a report)
    bse_rep = {}
    for s, Γ, ch in ((1, Γm, "m"), (-1, Γd, "d")):
        rp = _lib.BseReport()
        o = out["mixed"]
        ctx.check(Lb.ldga_chi_gamma_dev(ctx.handle, s, _lib.PREC_MIXED, frp, mP.U, mP.β, Γ.data_ptr(),
                                        χ0.data_ptr(), o["χ" + ch].data_ptr(), o["γ" + ch].data_ptr(),
                                        C.byref(rp), sp))
        bse_rep[ch] = {"ir_steps_mean": rp.ir_steps_mean, "ir_steps_max": rp.ir_steps_max,
                       "n_fallback": rp.n_fallback, "n_fp64_resolved": rp.n_fp64_resolved,
                       "max_rel_resid": rp.max_rel_resid}
    result["mixed_bse_report"] = bse_rep
    result["lambdadm_newton_step"] = lambdadm_step_costs(L, ctx, out[main_prec], λ0, inp, numax, Nq, Nν, Nω)
    result["lambdadm_convergence"] = {"workload": cfg.name, **lambdadm_convergence(L, ctx, out[main_prec], λ0, inp,
                                                                                    numax, Nq, Nν, Nω)}
    result["lambdadm_convergence_rooted"] = lambdadm_convergence_rooted(L, ctx)
    result["run_sc_iteration"] = run_sc_iteration_cost(L, ctx, out[main_prec], λ0, inp, numax, Nq, Nν, Nω)
    # G tables (helpers.jl:153-183) from the wedge G_loc, as the self-consistency loop rebuilds them:
    # algorithmic bytes = wedge G read + both full tables written
    gw = up(inp.gLoc)
    ga, gb = torch.empty_like(Gf), torch.empty_like(Gr)
    ts = []
    for it in range(4):
        e0, e1 = ev(), ev()
        e0.record(stream)
        ctx.check(Lb.ldga_g_tables_dev(ctx.handle, frp, gw.data_ptr(), ga.data_ptr(), gb.data_ptr(), sp))
        e1.record(stream)
        torch.cuda.synchronize()
        if it:
            ts.append(e0.elapsed_time(e1))
    gt_ms = float(np.mean(ts))
    gt_bytes = Nq * Nfft * 16 + 2 * Nk * Nfft * 16
    result["g_tables"] = {"ms": gt_ms, "GBps": gt_bytes / (gt_ms * 1e-3) / 1e9,
                          "max_rel_err_vs_input_tables": max(rel(ga, Gf), rel(gb, Gr))}
    if not args.no_cpu:
        χ0h = χ0.cpu().numpy().reshape((Nq, Nν, Nω), order="F")
        result["cpu_baseline"] = cpu_baseline(inp, numax, args.cpu_points, χ0h)
    ctx.close()
    print(json.dumps(result, ensure_ascii=False), file=_STDOUT)


# the JSON line is the only output on stdout: everything else the run prints goes to stderr
_STDOUT = sys.stdout

if __name__ == "__main__":
    import contextlib
    with contextlib.redirect_stdout(sys.stderr):
        main()
