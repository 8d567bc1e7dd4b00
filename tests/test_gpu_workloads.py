"""The workloads of the benchmark description at their real sizes, on the GPU against the oracle:

* 3D 24^3, Nν = 80 (the precision form's 3D half): the whole chain bubble -> BSE -> λ₀ -> Σ in FP64 and
  mixed precision against the oracle's FP64 chain (LAPACK BSE on every (q, ω)), with a per-output error
  line (χ, γ, Σ, n, E_pot) for each precision;
* 2D 64^2, Nν = 120: the chained Σ (GPU χ, γ, λ₀) against the oracle's EoM, the BSE on 500 sampled
  (q, ω) per channel against LAPACK;
* 3D 32^3, Nν = 160, Nω = 161 (the largest one-GPU case): mixed BSE sampled against LAPACK, bubble on
  sampled ω slices, Σ at νmax = 60 on sampled ν, FP64 vs mixed, and the peak device memory;
* 2D 48^2, Nν = 100 (the differentiation form): the adjoint λdm Jacobian against central differences
  of the oracle's residual at an admissible λ.
"""
import json
import math

import numpy as np
import pytest

from oracle import ldga_oracle as O

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _setup(name):
    from test_gpu_parity import setup
    return setup(name)


def _mem_used_gb():
    import torch
    free, total = torch.cuda.mem_get_info(0)
    return (total - free) / 1e9, total / 1e9


def _gpu_chain(ctx, inp, precision, numax, mem=None):
    """bubble -> BSE (m, d) -> λ₀ -> Σ through the host-array entry points."""
    import warnings
    import ladderdga_hip as L
    out = {}
    out["χ0"] = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    for ch, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        rep = {}
        out["χ" + ch], out["γ" + ch] = L.calc_χγ(ch, Γ, out["χ0"], inp.kG, inp.mP, inp.sP, precision=precision,
                                                 ctx=ctx, report=rep)
        out["rep" + ch] = rep
        if mem is not None:
            mem.append(_mem_used_gb()[0])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out["λ0"] = L.calc_λ0(out["χ0"], inp.F_m, inp.mP, inp.sP, inp.kG, ctx=ctx)
    out["Σ"] = L.calc_Σ(out["χm"], out["γm"], out["χd"], out["γd"], inp.χloc_m_sum, out["λ0"], inp.G_rfft, inp.kG,
                        inp.mP, inp.sP, νmax=numax, ctx=ctx)
    if mem is not None:
        mem.append(_mem_used_gb()[0])
    return out


def _bse_point(Γ, χ0, q, w, s, U, β):
    A = Γ[:, :, w] + np.diag(1.0 / χ0[q, :, w])
    inv, _, _ = O.inv_lapack(np.asfortranarray(A))
    χ = np.real(inv.sum()) / β ** 2
    return χ, inv.sum(axis=0) / (χ0[q, :, w] * (1 + s * U * χ))


def _bse_sampled_err(inp, χ0, out, pts):
    worst = {"χ": 0.0, "γ": 0.0}
    for ch, Γ, s in (("m", inp.Γ_m, 1), ("d", inp.Γ_d, -1)):
        for q, w in pts:
            rχ, rγ = _bse_point(Γ, χ0, q, w, s, inp.mP.U, inp.mP.β)
            worst["χ"] = max(worst["χ"], abs(out["χ" + ch].data[q, w] - rχ) / abs(rχ))
            worst["γ"] = max(worst["γ"], relerr(out["γ" + ch].data[q, :, w], rγ))
    return worst


def _χγ_slice(args):
    from threadpoolctl import threadpool_limits
    ch, Γ, χ0, D, Ns, t, wedge, mP, sP = args
    with threadpool_limits(limits=1):   # one LAPACK thread per worker process
        return O.calc_χγ(ch, Γ, χ0, O.KGrid(D, Ns, t, wedge_coords=wedge), mP, sP)


def _oracle_χγ_parallel(ch, Γ, χ0, oK, omP, osP, workers=8):
    """The oracle's calc_χγ (one LAPACK zgetrf+zgetri per (q, ω)) over ω slices in fresh CPU processes
    (spawned: they import numpy and the oracle only), joined along ω."""
    import multiprocessing as mp
    import os
    from concurrent.futures import ProcessPoolExecutor
    Nω = χ0.shape[2]
    n = max(1, min(workers, os.cpu_count() or 1, Nω))
    cuts = np.linspace(0, Nω, n + 1).astype(int)
    jobs = [(ch, np.ascontiguousarray(Γ[:, :, a:b]), np.ascontiguousarray(χ0[:, :, a:b]), oK.D, oK.Ns, oK.t, oK.wedge,
             omP, osP) for a, b in zip(cuts[:-1], cuts[1:])]
    with ProcessPoolExecutor(n, mp_context=mp.get_context("spawn")) as ex:
        parts = list(ex.map(_χγ_slice, jobs))
    return np.concatenate([p[0] for p in parts], axis=1), np.concatenate([p[1] for p in parts], axis=2)


def _omodel(inp):
    return O.ModelParameters(inp.mP.U, inp.mP.μ, inp.mP.β, inp.mP.n, 0, 0, inp.mP.Ekin_DMFT)


def _admissible_λ(χm, χd):
    import ladderdga_hip as L
    return [2 * abs(L.get_λ_min(χm)) + 1.0, 2 * abs(L.get_λ_min(χd)) + 1.0]


def test_chain_3d24_per_output_errors(ctx):
    """3D 24^3 (U = 1.5, β = 20, Nν = 80, Nω = 81): both GPU chains against the oracle's FP64 chain, every
    output: χ and γ on all (q, ω) against LAPACK, λ₀, Σ at νmax = 30, and n, E_pot (E_pot = U⟨n↑n↓⟩) of
    the λdm residual at an admissible λ.  Bar: 1e-10 (normwise relative) for both precisions."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _setup("3d24")
    numax = min(inp.sP.n_iν, (3 * (2 * inp.sP.n_iω + 1)) // 8)
    print("3d24: oracle bubble", flush=True)
    rχ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    print("3d24: oracle BSE (LAPACK on every (q, ω), both channels)", flush=True)
    ref = {}
    ref["χm"], ref["γm"] = _oracle_χγ_parallel("m", inp.Γ_m, rχ0, oK, omP, osP)
    ref["χd"], ref["γd"] = _oracle_χγ_parallel("d", inp.Γ_d, rχ0, oK, omP, osP)
    ref["λ0"] = O.calc_λ0(rχ0, inp.F_m, omP, osP)
    c2 = omP.Ekin_DMFT
    oχm, oχd = O.ChiT(ref["χm"], omP.beta, tail_c=[0, 0, c2]), O.ChiT(ref["χd"], omP.beta, tail_c=[0, 0, c2])
    print("3d24: oracle EoM", flush=True)
    ref["Σ"] = O.calc_Σ(oχm, ref["γm"], oχd, ref["γd"], inp.χloc_m_sum, ref["λ0"], inp.G_rfft, oK, omP, osP,
                        νmax=numax)
    λ = _admissible_λ(ref["χm"], ref["χd"])
    r = O.λdm_residual_vals(np.array(λ), oχm, ref["γm"], oχd, ref["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                            ref["λ0"], oK, omP, osP, numax, omP.mu)
    ref["n"], ref["μ"], ref["E_pot"] = r[0], r[1], r[3]
    lines = {}
    for prec in ("fp64", "mixed"):
        print(f"3d24: GPU chain {prec}", flush=True)
        g = _gpu_chain(ctx, inp, prec, numax)
        assert g["repm"]["n_zero_pivot"] == 0 and g["repd"]["n_zero_pivot"] == 0
        prob = L.LambdaDmProblem(g["χm"], g["γm"], g["χd"], g["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, g["λ0"],
                                 inp.kG, inp.mP, inp.sP, numax, ctx=ctx)
        res = prob.eval(λ)
        prob.free()
        err = {"χ0": relerr(g["χ0"].data, rχ0),
               "χ": max(relerr(g["χm"].data, ref["χm"]), relerr(g["χd"].data, ref["χd"])),
               "γ": max(relerr(g["γm"].data, ref["γm"]), relerr(g["γd"].data, ref["γd"])),
               "λ0": relerr(g["λ0"], ref["λ0"]), "Σ": relerr(g["Σ"], ref["Σ"]),
               "Σ-Hartree": relerr(g["Σ"] - inp.mP.n * inp.mP.U / 2, ref["Σ"] - omP.n * omP.U / 2),
               "n": abs(res["n"] - ref["n"]) / abs(ref["n"]), "μ": abs(res["mu"] - ref["μ"]) / abs(ref["μ"]),
               "E_pot": abs(res["Epot1"] - ref["E_pot"]) / abs(ref["E_pot"])}
        lines[prec] = err
        print("per-output relative error vs the FP64 oracle, 3D 24^3 " + prec + ": " + json.dumps(err), flush=True)
        for k, v in err.items():
            assert v <= 1e-10, (prec, k, v)
        if prec == "mixed":
            assert g["repm"]["n_fallback"] == g["repm"]["n_fp64_resolved"]


def test_chain_2d64_chained_sigma(ctx):
    """2D 64^2 (Nν = 120, Nω = 121, νmax = 45): Σ of the GPU chain against the oracle's EoM on the GPU's own
    χ, γ and λ₀ (FP64 and mixed), λ₀ on sampled ω against the oracle, and the BSE on 500 sampled (q, ω) per
    channel against LAPACK."""
    inp, oK, omP, osP = _setup("2d64")
    numax = 45
    rng = np.random.default_rng(64)
    Nq, Nω = inp.kG.Nq, 2 * inp.sP.n_iω + 1
    pts = list(zip(rng.integers(0, Nq, 500), rng.integers(0, Nω, 500)))
    c2 = omP.Ekin_DMFT
    for prec in ("fp64", "mixed"):
        print(f"2d64: GPU chain {prec}", flush=True)
        g = _gpu_chain(ctx, inp, prec, numax)
        χ0 = g["χ0"].data
        w = _bse_sampled_err(inp, χ0, g, pts)
        assert w["χ"] <= 1e-10 and w["γ"] <= 1e-10, w
        ws = [0, 60, 120]
        rl = O.calc_λ0(χ0[:, :, ws], inp.F_m[:, :, ws], omP, osP)
        assert relerr(g["λ0"][:, :, ws], rl) <= 1e-10
        print(f"2d64: oracle EoM on the {prec} chain", flush=True)
        rΣ = O.calc_Σ(O.ChiT(g["χm"].data, omP.beta, tail_c=[0, 0, c2]), g["γm"].data,
                      O.ChiT(g["χd"].data, omP.beta, tail_c=[0, 0, c2]), g["γd"].data, inp.χloc_m_sum, g["λ0"],
                      inp.G_rfft, oK, omP, osP, νmax=numax)
        err = relerr(g["Σ"], rΣ)
        print(f"2d64 {prec}: BSE sampled χ {w['χ']:.2e} γ {w['γ']:.2e}, Σ {err:.2e}", flush=True)
        assert err <= 1e-10


def test_largest_3d32(ctx):
    """3D 32^3 (U = 2, β = 25, Nν = 160, Nω = 161, Nq = 969): bubble on sampled (ν, ω) against the oracle's
    FFT convolution; mixed BSE (FP32 MFMA + FP64 refinement) on 40 sampled (q, ω) per channel against
    LAPACK; λ₀ on sampled ω; Σ at νmax = 60 on sampled ν against the oracle's EoM on the GPU's χ, γ, λ₀;
    FP64 (the global-memory kernel at this Nν) against mixed; the peak device memory against 288 GB."""
    inp, oK, omP, osP = _setup("3d32")
    numax = 60
    assert (inp.kG.Nq, 2 * inp.sP.n_iν, 2 * inp.sP.n_iω + 1) == (969, 160, 161)
    mem = [_mem_used_gb()[0]]
    print("3d32: GPU chain mixed", flush=True)
    g = _gpu_chain(ctx, inp, "mixed", numax, mem=mem)
    χ0 = g["χ0"].data
    # bubble: sampled (ν, ω) through the oracle's convolution
    f0 = osP.fft_range.start
    worst = 0.0
    for wi in (0, 80, 160):
        ωn = wi - inp.sP.n_iω
        for νi in (0, 41, 79, 80, 121, 159):
            νn = -inp.sP.n_iν + νi - O.trunc_div2(ωn)
            r = -omP.beta * O.conv_fft(oK, inp.G_fft[..., νn - f0], inp.G_rfft[..., νn + ωn - f0])
            worst = max(worst, relerr(χ0[:, νi, wi], r))
    assert worst <= 1e-10, worst
    rng = np.random.default_rng(32)
    pts = list(zip(rng.integers(0, inp.kG.Nq, 40), rng.integers(0, 161, 40))) + [(0, 80), (968, 0)]
    w = _bse_sampled_err(inp, χ0, g, pts)
    assert w["χ"] <= 1e-10 and w["γ"] <= 1e-10, w
    for ch in "md":
        rep = g["rep" + ch]
        assert rep["n_zero_pivot"] == 0 and rep["n_fallback"] == rep["n_fp64_resolved"]
    ws = [0, 80, 160]
    assert relerr(g["λ0"][:, :, ws], O.calc_λ0(χ0[:, :, ws], inp.F_m[:, :, ws], omP, osP)) <= 1e-10
    νsel = [0, 29, 59]
    c2 = omP.Ekin_DMFT
    print("3d32: oracle EoM on sampled ν", flush=True)
    rΣ = O.calc_Σ(O.ChiT(g["χm"].data, omP.beta, tail_c=[0, 0, c2]), g["γm"].data,
                  O.ChiT(g["χd"].data, omP.beta, tail_c=[0, 0, c2]), g["γd"].data, inp.χloc_m_sum, g["λ0"],
                  inp.G_rfft, oK, omP, osP, νmax=numax, νsel=νsel)
    eΣ = relerr(g["Σ"][:, νsel], rΣ[:, νsel])
    assert eΣ <= 1e-10
    print("3d32: GPU chain fp64 (global-memory BSE)", flush=True)
    f = _gpu_chain(ctx, inp, "fp64", numax, mem=mem)
    e = {k: relerr(g[k].data if hasattr(g[k], "data") else g[k], f[k].data if hasattr(f[k], "data") else f[k])
         for k in ("χm", "χd", "γm", "γd", "Σ")}
    peak, total = max(mem), _mem_used_gb()[1]
    print("3d32: mixed vs FP64 " + json.dumps(e) + f"; Σ vs oracle {eΣ:.2e}; BSE sampled {w}; "
          f"peak device memory {peak:.2f} GB of {total:.1f} GB", flush=True)
    assert max(e.values()) <= 1e-10
    assert peak < 0.1 * total


def test_adjoint_jacobian_2d48(ctx):
    """2D 48^2 (U = 2.5, β = 15, n = 0.85, Nν = 100, Nω = 101, νmax = 37): the hand-written adjoint Jacobian of
    the λdm residual (∂(res_pp, res_epot)/∂(λm, λd), through the EoM, χ_λ and the μ fit) against central
    differences of the oracle's residual, on the GPU's FP64 χ, γ, λ₀, at an admissible λ.

    On these inputs (past the magnetic instability: χm < 0 almost everywhere) the filling condition has
    more than one root in μ, and χ_λ has poles at 0.12 from this λm.  So the differences continue the μ
    branch of λ (each fit starts from μ(λ), as the reference's Newton loop restarts from its last μ) and
    use a step of 1e-7 relative, (h/0.12)² ≈ 1e-6 away from the derivative."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _setup("2d48")
    numax = min(inp.sP.n_iν, (3 * (2 * inp.sP.n_iω + 1)) // 8)
    print("2d48: GPU chain fp64", flush=True)
    g = _gpu_chain(ctx, inp, "fp64", numax)
    λ = _admissible_λ(g["χm"].data, g["χd"].data)
    prob = L.LambdaDmProblem(g["χm"], g["γm"], g["χd"], g["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, g["λ0"],
                             inp.kG, inp.mP, inp.sP, numax, ctx=ctx)
    got = prob.eval(λ, want_jac=True)
    prob.free()
    c2 = omP.Ekin_DMFT
    oχm, oχd = O.ChiT(g["χm"].data, omP.beta, tail_c=[0, 0, c2]), O.ChiT(g["χd"].data, omP.beta, tail_c=[0, 0, c2])
    rhs = omP.n / 2 * (1 - omP.n / 2)

    def res(l, μ0):
        r = O.λdm_residual_vals(np.array(l, float), oχm, g["γm"].data, oχd, g["γd"].data, inp.Σ_loc, inp.G_rfft,
                                inp.χloc_m_sum, g["λ0"], oK, omP, osP, numax, μ0)
        return np.array([r[5] - rhs, r[3] - r[4]]), r[1]

    print("2d48: oracle residual and central differences", flush=True)
    r0, μλ = res(λ, omP.mu)
    assert np.allclose(got["residual"], r0, rtol=1e-10, atol=1e-12 * np.max(np.abs(r0)))
    assert got["mu"] == pytest.approx(μλ, rel=1e-10)
    Jfd = np.empty((2, 2))
    for j in range(2):
        h = 1e-7 * max(1.0, abs(λ[j]))
        lp, lm = list(λ), list(λ)
        lp[j] += h
        lm[j] -= h
        (rp, μp), (rm, μm) = res(lp, μλ), res(lm, μλ)
        assert abs(μp - μm) <= 1e-3 * abs(μλ)   # the same μ branch on both sides
        Jfd[:, j] = (rp - rm) / (2 * h)
    print(f"2d48: λ = {λ}, adjoint J = {got['jac'].tolist()}, central FD = {Jfd.tolist()}", flush=True)
    assert np.allclose(got["jac"], Jfd, rtol=1e-5, atol=0.0)
