"""GPU tests of the bosonic sums on usable ranges, the χ tail coefficients, the failed μ fit and the
asynchronous mixed BSE, each against the oracle (or the synchronous call) on seeded inputs."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import ldga_oracle as O

pytestmark = pytest.mark.gpu


def _setup(name="2d16", **kw):
    from test_gpu_parity import setup
    return setup(name, **kw)


@pytest.mark.parametrize("rng_", [(0, 40), (5, 29), (17, 23), (20, 20)])
@pytest.mark.parametrize("lam", [0.0, 0.37])
def test_sum_ω_and_sum_ωk_usable_ranges(ctx, rng_, lam):
    """sum_ω (DataTypes.jl:277-308) per q and sum_ωk (:257-264) on full and strict usable ranges, with and
    without λ, against the oracle, which follows the reference's pairing of the usable slice with the
    leading frequencies of the grid; sum_kω keeps raising on a strict range, as the reference's does."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _setup()
    Nω = 2 * inp.sP.n_iω + 1
    lo, hi = rng_
    d = np.asfortranarray(np.random.default_rng(9).uniform(0.05, 0.8, (inp.kG.Nq, Nω)))
    u = range(lo, hi + 1)
    χ = L.χT(d, inp.mP.β, tail_c=[0.0, 0.0, 0.31], usable_ω=u)
    oχ = O.ChiT(d, inp.mP.β, tail_c=[0.0, 0.0, 0.31], usable_ω=u)
    print(f"sum_ω usable {lo}..{hi} λ={lam}", flush=True)
    if lam == 0.0:
        got = L.sum_ω(χ, ctx=ctx)
        ref = O.sum_ω(oχ)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
        assert np.allclose(L.sum_ω(χ, force_full_range=True, ctx=ctx), O.sum_ω(oχ, force_full_range=True),
                           rtol=1e-13, atol=1e-15)
    got_k = L.sum_ωk(inp.kG, χ, λ=lam, ctx=ctx)
    ref_k = O.sum_ωk(oK, oχ, λ=lam)
    assert got_k == pytest.approx(ref_k, rel=1e-12, abs=1e-14)
    if (lo, hi) == (0, Nω - 1):
        # on the full range the two orders of summation agree (test/DataTypes.jl:66)
        assert got_k == pytest.approx(L.sum_kω(inp.kG, χ, λ=lam if lam else math.nan, ctx=ctx), rel=1e-12)
    else:
        with pytest.raises(IndexError):
            L.sum_kω(inp.kG, χ, ctx=ctx)
    with pytest.raises(ValueError):
        L.sum_ω(L.χT(d, inp.mP.β, tail_c=[0.0, 0.1, 0.31]), ctx=ctx)


def test_calc_Σ_and_residual_after_update_tail(ctx):
    """update_tail! (DataTypes.jl:315-345) changes χm.tail_c[3] away from Ekin_DMFT: calc_Σ's tail
    (sum_kω(χm) with χm's own c₂, ladderDGATools_singleCore.jl:114-136) and calc_Σ_parts follow it, and
    the λdm residual's χ sums use each χ's own c₂ -- all against the oracle, which reads χ.tail_c."""
    import ladderdga_hip as L
    from test_gpu_parity import _ladder_inputs, relerr
    inp, oK, omP, osP, _, cm, cd, gm, gd, l0 = _ladder_inputs("tiny", dict(D=2, Ns=6, n_iν=5, n_iω=4, n=0.9, U=1.4))
    c2 = inp.mP.Ekin_DMFT
    χm = L.χT(cm.copy(order="F"), inp.mP.β, tail_c=[0.0, 0.0, c2])
    χd = L.χT(cd.copy(order="F"), inp.mP.β, tail_c=[0.0, 0.0, c2])
    L.update_tail(χm, [0.0, 0.0, c2 + 0.37])
    L.update_tail(χd, [0.0, 0.0, c2 - 0.11])
    assert χm.tail_c[2] != c2 and χd.tail_c[2] != c2
    oχm = O.ChiT(np.array(χm.data), omP.beta, tail_c=list(χm.tail_c))
    oχd = O.ChiT(np.array(χd.data), omP.beta, tail_c=list(χd.tail_c))
    νmax = 3
    print("calc_Σ after update_tail", flush=True)
    Σ = L.calc_Σ(χm, L.γT(gm), χd, L.γT(gd), 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, νmax=νmax, λm=0.2,
                 λd=0.05, ctx=ctx)
    ref = O.calc_Σ(oχm, gm, oχd, gd, 0.123, l0, inp.G_rfft, oK, omP, osP, νmax=νmax, λm=0.2, λd=0.05)
    assert relerr(Σ, ref) <= 1e-10
    # the tail is the part that moved: with the old c₂ the result differs
    ref_old = O.calc_Σ(O.ChiT(oχm.data, omP.beta, tail_c=[0, 0, c2]), gm, oχd, gd, 0.123, l0, inp.G_rfft, oK, omP,
                       osP, νmax=νmax, λm=0.2, λd=0.05)
    assert relerr(ref_old, ref) > 1e-6
    P = L.calc_Σ_parts(χm, L.γT(gm), χd, L.γT(gd), 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, λm=0.2, λd=0.05,
                       ctx=ctx)
    full = L.calc_Σ(χm, L.γT(gm), χd, L.γT(gd), 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, λm=0.2, λd=0.05,
                    ctx=ctx)
    assert relerr(P.sum(axis=2), full) <= 1e-10
    # the λdm residual's sums and E_pot,2 with each χ's own c₂
    λ = (0.3, 0.2)
    prob = L.LambdaDmProblem(χm, L.γT(gm), χd, L.γT(gd), inp.Σ_loc, inp.G_rfft, 0.123, l0, inp.kG, inp.mP, inp.sP,
                             νmax, ctx=ctx)
    got = prob.eval(λ)
    prob.free()
    r = O.λdm_residual_vals(np.array(λ), oχm, gm, oχd, gd, inp.Σ_loc, inp.G_rfft, 0.123, l0, oK, omP, osP, νmax,
                            omP.mu)
    for k, v in (("n", r[0]), ("mu", r[1]), ("Ekin", r[2]), ("Epot1", r[3]), ("Epot2", r[4]), ("lhs_c1", r[5])):
        assert got[k] == pytest.approx(v, rel=1e-10, abs=1e-12), k


def test_mu_fit_failure_gives_nan(ctx):
    """A μ fit that cannot reach |f| <= 1e-8 (the reference's find_zero tolerance) returns μ = NaN, as
    G_from_Σladder! does (GFTools.jl:171-176): a Σ_ladder holding a NaN, on the GPU and in the oracle.
    The same Σ without the NaN fits n."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _setup("tiny", D=2, Ns=6, n_iν=5, n_iω=4, n=0.9, U=1.4)
    νmax = 3
    Σ = np.asfortranarray(np.tile(inp.Σ_loc[:νmax], (inp.kG.Nq, 1)))
    print("μ fit: finite Σ", flush=True)
    μg, _ = L.G_from_Σladder(Σ, inp.Σ_loc, inp.kG, inp.mP, inp.sP, fix_n=True, ctx=ctx)
    μr, _ = O.G_from_Σladder(Σ, inp.Σ_loc, oK, omP, osP, fix_n=True)
    assert math.isfinite(μg) and μg == pytest.approx(μr, rel=1e-10, abs=1e-12)
    Σ[2, 1] = np.nan
    print("μ fit: Σ with a NaN", flush=True)
    μg, G = L.G_from_Σladder(Σ, inp.Σ_loc, inp.kG, inp.mP, inp.sP, fix_n=True, ctx=ctx)
    assert math.isnan(μg) and math.isnan(O.G_from_Σladder(Σ, inp.Σ_loc, oK, omP, osP, fix_n=True)[0])


def test_run_sc_trace_filling_and_sums(ctx):
    """The run_sc trace row (conditions_new.jl:507-515): n is the filling of the whole G_ladder
    (filling(G_ladder, kG, U, μ, β), GFTools.jl:234-241) and cs_m2/cs_d2 are sum_ωk, against the oracle
    on the G_ladder of the same iteration."""
    import ladderdga_hip as L
    from test_gpu_lambda import ladder
    inp, oK, omP, osP, o, p, νmax = ladder("tiny")
    λ = (94.0, -26.0)
    print("run_sc trace", flush=True)
    r = L.run_sc(p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                 inp.kG, inp.mP, inp.sP, type="fix", νmax=νmax, maxit=2, mixing=0.2, λm=λ[0], λd=λ[1], trace=True,
                 ctx=ctx)
    ref = O.run_sc_fix(o["χm"], o["γm"], o["χd"], o["γd"], o["λ0"], omP.mu, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                       oK, omP, osP, νmax, λ[0], λ[1], maxit=2, mixing=0.2)
    last = r.trace[-1]
    n_full = O.filling_kG(ref["G"], oK, omP.U, ref["μ"], omP.beta)
    assert last["n"] == pytest.approx(n_full, rel=1e-9)
    # the filling over the whole fft_range differs from the ν < νmax filling the result reports
    assert abs(n_full - ref["n"]) > 1e-6
    assert last["cs_m2"] == pytest.approx(O.sum_ωk(oK, o["χm"], λ=λ[0]), rel=1e-12)
    assert last["cs_d2"] == pytest.approx(O.sum_ωk(oK, o["χd"], λ=λ[1]), rel=1e-12)


def test_calc_gen_χ_with_shell(ctx):
    """calc_gen_χ (BSETools.jl:259-275) with ν_shell_size = 2: the diagonal takes χ₀[q, shell+1:end-shell]."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _setup("tiny", D=2, Ns=4, n_iν=4, n_iω=3)
    sh = 2
    osPs = O.SimulationParameters(inp.sP.n_iω, inp.sP.n_iν, sh, inp.sP.shift, fft_range=osP.fft_range)
    χ0s = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osPs)
    assert χ0s.shape[1] == 2 * inp.sP.n_iν + 2 * sh
    print("gen_chi shell 2", flush=True)
    X = L.calc_gen_χ(inp.Γ_m, L.χ0T(np.asfortranarray(χ0s), inp.mP.β, ν_shell_size=sh), inp.kG, ctx=ctx)
    ref = O.calc_gen_χ(inp.Γ_m, χ0s, sh=sh)
    assert np.max(np.abs(X - ref)) <= 1e-10 * np.max(np.abs(ref))
    # the shell columns hold other values: ignoring the offset would not match
    assert np.max(np.abs(O.calc_gen_χ(inp.Γ_m, χ0s[:, :-2 * sh, :], sh=0) - ref)) > 1e-6 * np.max(np.abs(ref))


def test_mixed_bse_without_report_is_asynchronous_and_equal(ctx):
    """ldga_chi_gamma_dev in mixed mode without a report makes no host synchronisation: the FP64 re-solve
    of the refinement failures reads their count on the device.  Its χ and γ equal the reported call's
    bit for bit, re-solved matrices included (an ill-conditioned Γ forces one)."""
    import ladderdga_hip as L
    from ladderdga_hip import _lib
    from ladderdga_hip.api import _freq
    from test_gpu_parity import _make_ill_conditioned
    inp, oK, omP, osP = _setup("tiny", D=2, Ns=4, n_iν=4, n_iω=3)
    χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    Γ = _make_ill_conditioned(inp.Γ_m, χ0, [(0, 0)])
    rep = {}
    print("mixed BSE with report", flush=True)
    χa, γa = L.calc_χγ("m", Γ, L.χ0T(np.asfortranarray(χ0), inp.mP.β), inp.kG, inp.mP, inp.sP, precision="mixed",
                       ctx=ctx, report=rep)
    assert rep["n_fallback"] >= 1 and rep["n_fp64_resolved"] == rep["n_fallback"]
    ctx.set_kgrid(inp.kG)
    dΓ, dχ0 = ctx.to_device(np.asfortranarray(Γ)), ctx.to_device(np.asfortranarray(χ0))
    dχ, dγ = ctx.empty(χa.data.shape, np.float64), ctx.empty(γa.data.shape, np.complex128)
    print("mixed BSE without report (_dev)", flush=True)
    rc = _lib.lib().ldga_chi_gamma_dev(ctx.handle, 1, _lib.PREC_MIXED, C.byref(_freq(inp.sP)), inp.mP.U, inp.mP.β,
                                       dΓ.ptr, dχ0.ptr, dχ.ptr, dγ.ptr, None, None)
    ctx.check(rc)
    ctx.sync()
    assert np.array_equal(dχ.to_host(), χa.data) and np.array_equal(dγ.to_host(), γa.data)
    for a in (dΓ, dχ0, dχ, dγ):
        a.free()
