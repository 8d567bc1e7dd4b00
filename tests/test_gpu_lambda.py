"""GPU λ corrections against the oracle: the λdm residual (conditions_new.jl:
250-287), its adjoint Jacobian (against central finite differences of the
oracle residual), the λdm and λm roots."""
import math

import numpy as np
import pytest

from oracle import ldga_oracle as O

pytestmark = pytest.mark.gpu

_cache = {}

TINY = {"tiny": dict(D=2, Ns=6, n_iν=6, n_iω=6, n=0.9, U=1.4, β=4.0),
        # a system on which the self-consistency loop contracts (weak coupling, hot, no vertex noise)
        "conv": dict(D=2, Ns=6, n_iν=6, n_iω=6, n=0.9, U=1.0, β=2.0, noise=0.0),
        # a 3D grid whose octant transforms take the register-resident pass
        "reg3d16": dict(D=3, Ns=16, n_iν=4, n_iω=4, n=0.9, U=1.4, β=4.0, t=0.4)}


def ladder(name):
    """Oracle ladder quantities (bubble -> BSE -> λ0) and the product's view of them."""
    if name in _cache:
        return _cache[name]
    import ladderdga_hip as L
    from ladderdga_hip import synthetic
    cfg = synthetic.CONFIGS[name] if name in synthetic.CONFIGS else synthetic.tiny_config(**TINY[name])
    inp = synthetic.build_inputs(cfg)
    oK = O.KGrid(cfg.D, cfg.Ns, cfg.t, wedge_coords=inp.kG.wedge)
    omP = O.ModelParameters(inp.mP.U, inp.mP.μ, inp.mP.β, inp.mP.n, 0, 0, inp.mP.Ekin_DMFT)
    osP = O.SimulationParameters(cfg.n_iω, cfg.n_iν, 0, cfg.shift)
    χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χm, γm = O.calc_χγ("m", inp.Γ_m, χ0, oK, omP, osP)
    χd, γd = O.calc_χγ("d", inp.Γ_d, χ0, oK, omP, osP)
    λ0 = np.asfortranarray(O.calc_λ0(χ0, inp.F_m, omP, osP))
    c2 = omP.Ekin_DMFT
    o = dict(χm=O.ChiT(χm, omP.beta, tail_c=[0, 0, c2]), χd=O.ChiT(χd, omP.beta, tail_c=[0, 0, c2]), γm=γm, γd=γd,
             λ0=λ0)
    p = dict(χm=L.χT(np.asfortranarray(χm), inp.mP.β, tail_c=[0, 0, c2]), γm=L.γT(np.asfortranarray(γm)),
             χd=L.χT(np.asfortranarray(χd), inp.mP.β, tail_c=[0, 0, c2]), γd=L.γT(np.asfortranarray(γd)))
    _, νmax = O.gen_νω_indices(o["χm"], o["χd"], omP, osP)
    _cache[name] = (inp, oK, omP, osP, o, p, νmax)
    return _cache[name]


def oracle_res(name, λ, μ0=None):
    inp, oK, omP, osP, o, p, νmax = ladder(name)
    r = O.λdm_residual_vals(np.array(λ, float), o["χm"], o["γm"], o["χd"], o["γd"], inp.Σ_loc, inp.G_rfft,
                            inp.χloc_m_sum, o["λ0"], oK, omP, osP, νmax, omP.mu if μ0 is None else μ0)
    rhs = omP.n / 2 * (1 - omP.n / 2)
    return dict(n=r[0], mu=r[1], Ekin=r[2], Epot1=r[3], Epot2=r[4], lhs=r[5], Σ=r[6], G=r[7],
                residual=np.array([r[5] - rhs, r[3] - r[4]]))


def assert_oracle_root(name, λ, tol=1e-8):
    """λ is a root of the oracle's λdm residual.  On these synthetic inputs the residual has several
    roots (the oracle reaches (94.29, -26.56) from (0, 0) and from (94.2859, -30), (98.94, -26.56) from
    (94, -26)), and the Newton path from (0, 0) runs λm out to ~1e7 and resets after a NaN: which root it
    reaches turns on last-bit differences of the residual (such as the order of a k-ω sum).  A root the
    product reaches must be one of the oracle's."""
    r = oracle_res(name, λ)["residual"]
    assert np.all(np.abs(r) <= tol), (λ, r)


def problem(ctx, name):
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder(name)
    return L.LambdaDmProblem(p["χm"], p["γm"], p["χd"], p["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, o["λ0"],
                             inp.kG, inp.mP, inp.sP, νmax, ctx=ctx)


@pytest.mark.parametrize("name,λ", [("tiny", (60.0, -20.0)), ("tiny", (94.0, -26.0)), ("tiny", (150.0, 3.0)),
                                     ("2d16", (192.0, -29.0)), ("2d16", (250.0, 10.0))])
def test_lambdadm_residual_parity(ctx, name, λ):
    print(f"residual {name} {λ}", flush=True)
    prob = problem(ctx, name)
    got = prob.eval(λ, arrays=True)
    prob.free()
    ref = oracle_res(name, λ)
    for k, kk in (("n", "n"), ("mu", "mu"), ("Ekin", "Ekin"), ("Epot1", "Epot1"), ("Epot2", "Epot2"),
                  ("lhs_c1", "lhs")):
        assert got[k] == pytest.approx(ref[kk], rel=1e-10, abs=1e-12), k
    assert got["mu_converged"] == 1
    Σ, G = got["Σ_ladder"], got["G_ladder"]
    assert np.max(np.abs(Σ - ref["Σ"])) <= 1e-10 * np.max(np.abs(ref["Σ"]))
    assert np.max(np.abs(G - ref["G"])) <= 1e-10 * np.max(np.abs(ref["G"]))


def test_self_consistency_rebuild(ctx):
    """One rebuild of the self-consistency loop (conditions_new.jl:485-492) on the GPU: G_ladder from the
    λdm residual, its G tables (helpers.jl:153-183) and the next bubble, against the oracle's chain."""
    import ladderdga_hip as L
    name, λ = "tiny", (94.0, -26.0)
    print(f"self-consistency rebuild {name} {λ}", flush=True)
    inp, oK, omP, osP, o, p, νmax = ladder(name)
    prob = problem(ctx, name)
    got = prob.eval(λ, arrays=True)
    prob.free()
    ref = oracle_res(name, λ)
    rel = lambda a, b: float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))
    f, r = L.G_fft(np.asfortranarray(got["G_ladder"]), inp.kG, inp.sP, ctx=ctx)
    rf, rr = O.G_fft_from_wedge(ref["G"], oK)
    assert rel(f, rf) <= 1e-10 and rel(r, rr) <= 1e-10
    χ0 = L.calc_bubble(f, r, inp.kG, inp.mP, inp.sP, ctx=ctx)
    assert rel(χ0.data, O.calc_bubble(rf, rr, oK, omP, osP)) <= 1e-10


def λdm_root(name):
    """The oracle's λdm root (cached)."""
    key = ("root", name)
    if key not in _cache:
        inp, oK, omP, osP, o, p, νmax = ladder(name)
        _cache[key] = tuple(O.λdm_correction(o["χm"], o["γm"], o["χd"], o["γd"], inp.Σ_loc, inp.G_rfft,
                                             inp.χloc_m_sum, o["λ0"], oK, omP, osP)[0])
    return _cache[key]


@pytest.mark.parametrize("at,mixing,maxit", [("fixed", 0.2, 6), ("fixed", 0.0, 4), ("root", 0.5, 6)])
def test_run_sc_fix(ctx, at, mixing, maxit):
    """run_sc (conditions_new.jl:342-560, type :fix) with the loop on the GPU against the oracle's loop:
    the same iterations, μ, filling, energies, Σ_ladder and G_ladder, at an arbitrary λ and at the λdm
    root.  On these synthetic inputs the loop does not converge for any mixing (the random vertices are
    not a self-consistent solution), so the iteration amplifies rounding and only the first few steps
    are comparable; the loop length here keeps the two trajectories within 1e-9."""
    import ladderdga_hip as L
    name = "tiny"
    λ = (94.0, -26.0) if at == "fixed" else λdm_root(name)
    print(f"run_sc fix λ={λ} mixing={mixing} maxit={maxit}", flush=True)
    inp, oK, omP, osP, o, p, νmax = ladder(name)
    r = L.run_sc(p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                 inp.kG, inp.mP, inp.sP, type="fix", νmax=νmax, maxit=maxit, mixing=mixing, λm=λ[0], λd=λ[1],
                 trace=True, ctx=ctx)
    ref = O.run_sc_fix(o["χm"], o["γm"], o["χd"], o["γd"], o["λ0"], omP.mu, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                       oK, omP, osP, νmax, λ[0], λ[1], maxit=maxit, mixing=mixing)
    assert len(r.trace) == ref["it"] and r.sc_converged == ref["sc_converged"]
    for a, b in ((r.μ, ref["μ"]), (r.n, ref["n"]), (r.EKin, ref["E_kin"]), (r.EPot_p1, ref["E_pot_1"]),
                 (r.EPot_p2, ref["E_pot_2"]), (r.PP_p2, ref["lhs_c1"])):
        assert a == pytest.approx(b, rel=1e-9, abs=1e-11)
    rel = lambda a, b: float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))
    assert rel(r.Σ_ladder, ref["Σ"]) <= 1e-9 and rel(r.G_ladder, ref["G"]) <= 1e-9
    # χ is left untouched
    assert np.array_equal(p["χm"].data, np.asfortranarray(o["χm"].data))
    # as the reference's run_sc!, :fix reports its loop-local λ = 0 and logs the λ used in the trace
    assert (r.λm, r.λd) == (0.0, 0.0)
    assert all((row["λm"], row["λd"]) == tuple(λ) for row in r.trace)
    # the trace's checksums of the last iteration (conditions_new.jl:507-515)
    last = r.trace[-1]
    assert last["cs_Σ"] == pytest.approx(abs(ref["Σ"].sum()), rel=1e-9)
    assert last["cs_G"] == pytest.approx(abs(ref["G"].sum()), rel=1e-9)
    assert last["cs_m"] == pytest.approx(O.sum_kω(oK, O.ChiT(O.χ_λ(o["χm"].data, λ[0]), omP.beta,
                                                             tail_c=o["χm"].tail_c)), rel=1e-12)


@pytest.mark.parametrize("mixing,iters", [(0.0, 12), (0.5, 20)])
def test_run_sc_converges(ctx, mixing, iters):
    """A synthetic system on which the :fix loop is a contraction (U = 1, β = 2, no vertex noise): the
    GPU loop stops at the same iteration as the oracle's, with sc_converged, and both reach the same
    fixed point.  The last energy changes are 4.9e-10 and 9.6e-9 against conv_abs = 1e-8."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("conv")
    print(f"run_sc converging mixing={mixing}", flush=True)
    r = L.run_sc(p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                 inp.kG, inp.mP, inp.sP, type="fix", νmax=νmax, maxit=100, mixing=mixing, trace=True, ctx=ctx)
    ref = O.run_sc_fix(o["χm"], o["γm"], o["χd"], o["γd"], o["λ0"], omP.mu, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                       oK, omP, osP, νmax, 0.0, 0.0, maxit=100, mixing=mixing)
    assert ref["sc_converged"] and ref["it"] == iters
    assert r.sc_converged and len(r.trace) == iters
    for a, b in ((r.μ, ref["μ"]), (r.n, ref["n"]), (r.EKin, ref["E_kin"]), (r.EPot_p1, ref["E_pot_1"])):
        assert a == pytest.approx(b, rel=1e-10, abs=1e-12)
    rel = lambda a, b: float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))
    assert rel(r.Σ_ladder, ref["Σ"]) <= 1e-10 and rel(r.G_ladder, ref["G"]) <= 1e-10
    # the result's validation flag follows the sum rules at the fixed point (conditions_new.jl:397)
    assert r.converged == (abs(r.PP_p1 - r.PP_p2) <= 1e-8 and abs(r.EPot_p1 - r.EPot_p2) <= 1e-8)


@pytest.mark.parametrize("mixing", [0.0, 0.5])
def test_run_sc_pre_dm_converges(ctx, mixing):
    """Full lDΓA_dm self-consistency (run_sc :pre_dm) on the contracting input: λdm is re-solved with
    every rebuilt G_rfft and the loop converges to a point where both sum rules hold, so the result
    validates (converged = true).  Against the oracle's loop (O.run_sc_dm) from its λdm root: same
    fixed point.  Each λ re-solve stops at the Newton tolerance (1e-8 on the residual), so the two
    loops may stop an iteration apart and agree to that tolerance, not to rounding."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("conv")
    print(f"run_sc pre_dm converging mixing={mixing}", flush=True)
    r = L.run_sc(p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                 inp.kG, inp.mP, inp.sP, type="pre_dm", νmax=νmax, maxit=40, mixing=mixing, trace=True, ctx=ctx)
    root = λdm_root("conv")
    ref = O.run_sc_dm(o["χm"], o["γm"], o["χd"], o["γd"], o["λ0"], omP.mu, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                      oK, omP, osP, νmax, root[0], root[1], maxit=40, mixing=mixing)
    assert ref["sc_converged"] and r.sc_converged
    assert abs(len(r.trace) - ref["it"]) <= 2
    assert np.allclose([r.λm, r.λd], [ref["λm"], ref["λd"]], rtol=1e-6, atol=1e-6)
    assert np.allclose([r.trace[0]["λm"], r.trace[0]["λd"]], root, rtol=1e-6, atol=1e-6)
    for a, b in ((r.μ, ref["μ"]), (r.EPot_p1, ref["E_pot_1"]), (r.EPot_p2, ref["E_pot_2"]), (r.PP_p2, ref["lhs_c1"])):
        assert a == pytest.approx(b, rel=1e-6, abs=1e-7)
    # the result validates when both sum rules hold at the last iterate (conditions_new.jl:397).  Without
    # mixing the last iterate is a λdm root (E_pot gap 0 in the oracle); with mixing 0.5 the mixed Σ lags
    # by the last step (gap 6.5e-9 in the oracle), too close to 1e-8 to require either way
    assert r.converged == (abs(r.PP_p1 - r.PP_p2) <= 1e-8 and abs(r.EPot_p1 - r.EPot_p2) <= 1e-8)
    assert abs(r.PP_p1 - r.PP_p2) <= 1e-8
    if mixing == 0.0:
        assert r.converged and abs(r.EPot_p1 - r.EPot_p2) <= 1e-9


def test_run_sc_pre_dm(ctx):
    """run_sc type :pre_dm: λ from λdm_correction first (a root of the oracle's residual), then re-solved
    with every rebuilt G_rfft.  The first iteration runs at the root; on these synthetic inputs a later
    re-solve finds no root, and the loop then stops with λ = NaN and converged = false, as the reference's
    does (conditions_new.jl:465-471)."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("tiny")
    print("run_sc pre_dm", flush=True)
    r = L.run_sc(p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                 inp.kG, inp.mP, inp.sP, type="pre_dm", νmax=νmax, maxit=6, mixing=0.2, trace=True, ctx=ctx)
    assert 1 <= len(r.trace) <= 6
    first = r.trace[0]
    assert_oracle_root("tiny", (first["λm"], first["λd"]))
    assert np.isfinite([first["μ"], first["n"], first["EPot"], first["EPot_c2"]]).all()
    if not (np.isfinite(r.λm) and np.isfinite(r.λd)):
        assert not r.converged and not r.sc_converged
    elif r.sc_converged:
        assert r.converged == (abs(r.PP_p1 - r.PP_p2) <= 1e-8 and abs(r.EPot_p1 - r.EPot_p2) <= 1e-8)


def test_run_sc_pre_m(ctx):
    """run_sc type :pre_m (conditions_new.jl:365-369): λm from λm_correction with the plain-mode λm_rhs
    (2 imp_density - Σ χd, helpers.jl:234-255), λd = 0, then the loop with λ kept fixed.  imp_density is
    chosen so that the rhs is Σ χm at λ = 35.  The synthetic Σ χm^λ is not monotonic in λ and has several
    roots; from this rhs the Newton path (12 steps, to the root near 76) is stable under 1e-13 changes of
    the sums, while from other targets it wanders for 100 steps and ends where rounding sends it.  The
    root must meet the condition, and the loop must follow the oracle's :fix loop at that λ."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("tiny")
    imp = (O.sum_kω(oK, o["χm"], λ=35.0) + O.sum_kω(oK, o["χd"])) / 2
    rhs = 2 * imp - O.sum_kω(oK, o["χd"])
    assert L.λm_rhs(imp, p["χm"], p["χd"], inp.kG, inp.mP, inp.sP, ctx=ctx) == pytest.approx(rhs, rel=1e-12)
    λm = O.λm_correction(o["χm"], rhs, oK, omP, osP)
    assert O.sum_kω(oK, o["χm"], λ=λm) == pytest.approx(rhs, rel=1e-10)
    print(f"run_sc pre_m λm={λm}", flush=True)
    r = L.run_sc(p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], inp.mP.μ, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                 inp.kG, inp.mP, inp.sP, type="pre_m", νmax=νmax, maxit=4, mixing=0.2, trace=True,
                 imp_density=imp, ctx=ctx)
    assert r.λm == pytest.approx(λm, rel=1e-9) and r.λd == 0.0
    ref = O.run_sc_fix(o["χm"], o["γm"], o["χd"], o["γd"], o["λ0"], omP.mu, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                       oK, omP, osP, νmax, λm, 0.0, maxit=4, mixing=0.2)
    assert len(r.trace) == ref["it"]
    for a, b in ((r.μ, ref["μ"]), (r.EKin, ref["E_kin"]), (r.EPot_p1, ref["E_pot_1"]), (r.EPot_p2, ref["E_pot_2"])):
        assert a == pytest.approx(b, rel=1e-8, abs=1e-10)
    assert np.max(np.abs(r.Σ_ladder - ref["Σ"])) <= 1e-8 * np.max(np.abs(ref["Σ"]))


@pytest.mark.parametrize("name,λ", [("tiny", (94.0, -26.0)), ("tiny", (130.0, -10.0)), ("tiny", (40.0, 5.0))])
def test_lambdadm_adjoint_jacobian(ctx, name, λ):
    """Adjoint Jacobian vs central differences (h = 1e-6 relative) of the oracle residual."""
    prob = problem(ctx, name)
    got = prob.eval(λ, want_jac=True)
    prob.free()
    J = got["jac"]
    Jfd = np.empty((2, 2))
    for j in range(2):
        h = 1e-6 * max(1.0, abs(λ[j]))
        lp, lm = list(λ), list(λ)
        lp[j] += h
        lm[j] -= h
        Jfd[:, j] = (oracle_res(name, lp)["residual"] - oracle_res(name, lm)["residual"]) / (2 * h)
    assert np.allclose(J, Jfd, rtol=1e-5, atol=1e-7 * np.max(np.abs(Jfd))), (J, Jfd)


@pytest.mark.parametrize("name", ["2d16", "reg3d16"])
def test_lambdadm_adjoint_jacobian_pass_kinds(ctx, name):
    """The adjoint Jacobian on grids whose EoM and adjoint transforms take the register-resident pass (2D
    16^2 with its corner sum, 3D 16^3 with the lines x >= y), against central differences of the oracle
    residual at an admissible λ.  On 2D 16^2 a pole of χ_λ = χ/(1 + λχ) lies within 1e-4 relative of that
    λ (a difference step of 1e-6 is still 2e-4 off the derivative), so the step is 1e-8 relative, where the
    differences have converged to 1e-8."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder(name)
    λ = [2 * abs(L.get_λ_min(o["χm"].data)) + 1.0, 2 * abs(L.get_λ_min(o["χd"].data)) + 1.0]
    print(f"adjoint {name} {λ}", flush=True)
    prob = problem(ctx, name)
    got = prob.eval(λ, want_jac=True)
    prob.free()
    ref = oracle_res(name, λ)
    assert np.allclose(got["residual"], ref["residual"], rtol=1e-10, atol=1e-12 * np.max(np.abs(ref["residual"])))
    Jfd = np.empty((2, 2))
    for j in range(2):
        h = 1e-8 * max(1.0, abs(λ[j]))
        lp, lm = list(λ), list(λ)
        lp[j] += h
        lm[j] -= h
        Jfd[:, j] = (oracle_res(name, lp, ref["mu"])["residual"] - oracle_res(name, lm, ref["mu"])["residual"]) / (2 * h)
    assert np.allclose(got["jac"], Jfd, rtol=1e-5, atol=1e-7 * np.max(np.abs(Jfd))), (got["jac"], Jfd)


def test_lambdadm_correction_roots(ctx):
    """λdm root with the reference's FD Jacobian and with the adjoint.  From the reference's start (0, 0)
    both reach a root of the oracle's residual where the sum rules hold; from a start in the basin of the
    oracle's root (the oracle reaches it from there too) both converge to that root, the adjoint with
    fewer residual evaluations."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("tiny")
    near = (94.2859, -30.0)
    oroot, _ = O.λdm_correction(o["χm"], o["γm"], o["χd"], o["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, o["λ0"],
                                oK, omP, osP, λinit=near)
    res = {}
    for jac in ("fd", "adjoint"):
        r0 = L.λdm_correction(p["χm"], p["γm"], p["χd"], p["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, o["λ0"],
                              inp.kG, inp.mP, inp.sP, jacobian=jac, ctx=ctx)
        assert r0.converged
        assert abs(r0.PP_p1 - r0.PP_p2) < 1e-8 and abs(r0.EPot_p1 - r0.EPot_p2) < 1e-8
        assert_oracle_root("tiny", (r0.λm, r0.λd))
        r = L.λdm_correction(p["χm"], p["γm"], p["χd"], p["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, o["λ0"],
                             inp.kG, inp.mP, inp.sP, λinit=near, jacobian=jac, ctx=ctx)
        assert r.converged
        assert abs(r.PP_p1 - r.PP_p2) < 1e-8 and abs(r.EPot_p1 - r.EPot_p2) < 1e-8
        assert np.allclose([r.λm, r.λd], oroot, rtol=1e-7, atol=1e-7)
        res[jac] = r
    assert res["adjoint"].residual_evals < res["fd"].residual_evals
    # the returned Σ_ladder is the EoM at the root (full_run.jl:155-160)
    ref = oracle_res("tiny", [res["fd"].λm, res["fd"].λd])
    assert np.max(np.abs(res["fd"].Σ_ladder - ref["Σ"])) <= 1e-9 * np.max(np.abs(ref["Σ"]))
    # χ is left untouched (checksum invariance, full_run.jl:65-77)
    assert np.array_equal(p["χm"].data, np.asfortranarray(o["χm"].data))


def test_lambdam_correction(ctx):
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("2d16")
    rhs = omP.n * (1 - omP.n / 2) - O.sum_kω(oK, o["χd"])
    ref = O.λm_correction(o["χm"], rhs, oK, omP, osP)
    got, ok = L.λm_correction(p["χm"], rhs, inp.kG, inp.mP, inp.sP, ctx=ctx)
    assert got == pytest.approx(ref, rel=1e-10)
    assert ok


def test_lambda_correction_dispatch(ctx):
    """λ_correction (conditions_new.jl:60-85) through an lDΓAHelper: "m" is λm_correction_full
    (:174-190: native rhs, λm root, Σ_ladder over ν < n_iν at (λm, 0), μ fit, energies; the result's
    flags stored as the reference stores them), "dm" the λdm root, "sc" and "sc_m" the loop at λ = 0."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("tiny")
    imp = (O.sum_kω(oK, o["χm"], λ=35.0) + O.sum_kω(oK, o["χd"])) / 2
    h = L.lDΓAHelper(inp.kG, inp.mP, inp.sP, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, imp_density=imp)
    print("λ_correction m", flush=True)
    r = L.λ_correction("m", p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], h, ctx=ctx)
    λm = O.λm_correction(o["χm"], 2 * imp - O.sum_kω(oK, o["χd"]), oK, omP, osP)
    assert r.type == "m" and r.λm == pytest.approx(λm, rel=1e-9) and r.λd == 0.0
    assert r.converged and r.sc_converged            # the λm validation, stored in sc_converged
    ref = O.λdm_residual_vals(np.array([λm, 0.0]), o["χm"], o["γm"], o["χd"], o["γd"], inp.Σ_loc, inp.G_rfft,
                              inp.χloc_m_sum, o["λ0"], oK, omP, osP, inp.sP.n_iν, omP.mu)
    for a, b in ((r.μ, ref[1]), (r.EKin, ref[2]), (r.EPot_p1, ref[3]), (r.EPot_p2, ref[4]), (r.PP_p2, ref[5])):
        assert a == pytest.approx(b, rel=1e-8, abs=1e-10)
    assert r.Σ_ladder.shape == (inp.kG.Nq, inp.sP.n_iν) and r.n == inp.mP.n
    assert np.max(np.abs(r.Σ_ladder - ref[6])) <= 1e-8 * np.max(np.abs(ref[6]))
    print("λ_correction dm", flush=True)
    r = L.λ_correction(":dm", p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], h, ctx=ctx)
    assert r.type == "dm" and r.converged
    assert_oracle_root("tiny", (r.λm, r.λd))
    print("λ_correction sc, sc_m", flush=True)
    inp, oK, omP, osP, o, p, νmax = ladder("conv")
    h = L.lDΓAHelper(inp.kG, inp.mP, inp.sP, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum)
    ref = O.run_sc_fix(o["χm"], o["γm"], o["χd"], o["γd"], o["λ0"], omP.mu, inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum,
                       oK, omP, osP, νmax, 0.0, 0.0, maxit=100, mixing=0.2)
    for t in ("sc", "sc_m"):
        r = L.λ_correction(t, p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], h, ctx=ctx)
        assert r.sc_converged == ref["sc_converged"] and r.EPot_p1 == pytest.approx(ref["E_pot_1"], rel=1e-9)
    with pytest.raises(ValueError, match="not recognized"):
        L.λ_correction("x", p["χm"], p["γm"], p["χd"], p["γd"], o["λ0"], h, ctx=ctx)


@pytest.mark.parametrize("fix_n,μ", [(True, None), (False, 0.3), (True, 2.5)])
def test_G_from_Σladder(ctx, fix_n, μ):
    """G_from_Σladder (GFTools.jl:154-186) alone, through the entry point without χ: the μ fit and
    G over fft_range against the oracle, on a ladder Σ of the tiny system."""
    import ladderdga_hip as L
    inp, oK, omP, osP, o, p, νmax = ladder("tiny")
    Σ = oracle_res("tiny", (94.0, -26.0))["Σ"]
    print(f"G_from_Σladder fix_n={fix_n} μ={μ}", flush=True)
    μg, G = L.G_from_Σladder(Σ, inp.Σ_loc, inp.kG, inp.mP, inp.sP, fix_n=fix_n, μ=μ, ctx=ctx)
    μr, Gr = O.G_from_Σladder(Σ, inp.Σ_loc, oK, omP, osP, fix_n=fix_n, μ=μ)
    assert μg == pytest.approx(μr, rel=1e-10, abs=1e-12)
    assert np.max(np.abs(G - Gr)) <= 1e-10 * np.max(np.abs(Gr))
    if fix_n:
        f0 = osP.fft_range.start
        assert O.filling_pos_kG(G[:, -f0:-f0 + Σ.shape[1]], oK, omP.U, μg, omP.beta) == pytest.approx(omP.n, abs=1e-10)


def test_lambdadm_convergence_fd_vs_adjoint(ctx):
    """λdm solved to convergence from (0, 0) on a 16x16 system with roots (U = 1.4, β = 4, n = 0.9,
    Nν = 32, Nω = 33, no vertex noise): the reference's Newton with its forward-difference Jacobian and
    the adjoint Newton each reach a root of the oracle's residual; the adjoint takes one residual
    evaluation per step instead of three.  The system has more than one root (the oracle's own Newton,
    whose μ fits all restart from mP.μ, walks to another one), and the forward-difference path is
    sensitive to the last bits of the residual: with residuals that differ by 1e-15 relative (two
    orderings of the same sums) it ends at (22.63, -26.30) or at (31.46, -26.33).  The adjoint path
    ends at (22.63, -26.30) either way.  Wall times and roots are printed."""
    import time
    import ladderdga_hip as L
    from ladderdga_hip import synthetic
    cfg = synthetic.tiny_config(D=2, Ns=16, n_iν=16, n_iω=16, n=0.9, U=1.4, β=4.0, noise=0.0)
    TINY["conv16"] = dict(D=2, Ns=16, n_iν=16, n_iω=16, n=0.9, U=1.4, β=4.0, noise=0.0)
    inp, oK, omP, osP, o, p, νmax = ladder("conv16")
    assert inp.cfg == cfg
    res = {}
    for jac in ("fd", "adjoint"):
        t0 = time.perf_counter()
        r = L.λdm_correction(p["χm"], p["γm"], p["χd"], p["γd"], inp.Σ_loc, inp.G_rfft, inp.χloc_m_sum, o["λ0"],
                             inp.kG, inp.mP, inp.sP, jacobian=jac, ctx=ctx)
        dt = time.perf_counter() - t0
        print(f"λdm {jac}: {dt * 1e3:.1f} ms, {r.newton_steps} steps, {r.residual_evals} residuals, "
              f"root ({r.λm}, {r.λd})", flush=True)
        assert r.converged
        ref = O.λdm_residual_vals(np.array([r.λm, r.λd]), o["χm"], o["γm"], o["χd"], o["γd"], inp.Σ_loc, inp.G_rfft,
                                  inp.χloc_m_sum, o["λ0"], oK, omP, osP, νmax, r.μ)
        assert abs(ref[5] - omP.n / 2 * (1 - omP.n / 2)) <= 1e-8 and abs(ref[3] - ref[4]) <= 1e-8
        assert ref[1] == pytest.approx(r.μ, abs=1e-10)
        res[jac] = r
    same = np.allclose([res["fd"].λm, res["fd"].λd], [res["adjoint"].λm, res["adjoint"].λd], rtol=1e-6, atol=1e-6)
    print(f"λdm: the two Newtons end at the same root: {same}", flush=True)
    assert np.allclose([res["adjoint"].λm, res["adjoint"].λd], [22.630875764594, -26.299479747835], rtol=1e-8)
    assert res["adjoint"].residual_evals < res["fd"].residual_evals
