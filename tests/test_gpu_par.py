"""The worker-pool API on one GPU, as test/full_run.jl:14-60 exercises it: the
`_par` results equal the serial ones (bubble, χ, γ, tail and usable ranges,
Σ through initialize_EoM + calc_Σ_par / calc_Σ_par!), and both equal the
oracle to the FP64 bar.  The reference leaves its own serial-vs-parallel Σ
check commented out (full_run.jl:50); here it is asserted."""
import numpy as np
import pytest

from oracle import ldga_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _inputs(name, kw):
    from ladderdga_hip import synthetic
    cfg = synthetic.CONFIGS[name] if name in synthetic.CONFIGS else synthetic.tiny_config(**kw)
    inp = synthetic.build_inputs(cfg)
    oK = O.KGrid(cfg.D, cfg.Ns, cfg.t, wedge_coords=inp.kG.wedge)
    omP = O.ModelParameters(inp.mP.U, inp.mP.μ, inp.mP.β, inp.mP.n, 0, 0, inp.mP.Ekin_DMFT)
    osP = O.SimulationParameters(cfg.n_iω, cfg.n_iν, 0, cfg.shift)
    return inp, oK, omP, osP


@pytest.mark.parametrize("name,kw", [("tiny", dict(D=2, Ns=6, n_iν=5, n_iω=4, n=0.9, U=1.4)),
                                     ("tiny", dict(D=3, Ns=4, n_iν=3, n_iω=2, t=0.4)), ("2d16", {})])
def test_par_api_matches_serial_and_oracle(ctx, name, kw):
    import ladderdga_hip as L
    inp, oK, omP, osP = _inputs(name, kw)
    kG, mP, sP = inp.kG, inp.mP, inp.sP
    print(f"par api {name} {kw}", flush=True)
    L.clear_wcache()
    for k, v in (("kG", kG), ("mP", mP), ("sP", sP), ("G_fft", inp.G_fft), ("G_fft_reverse", inp.G_rfft)):
        L.update_wcache(k, v, ctx=ctx)

    # bubble (full_run.jl:17-21)
    bubble = L.calc_bubble(inp.G_fft, inp.G_rfft, kG, mP, sP, ctx=ctx)
    bubble_par = L.calc_bubble_par(kG, mP, sP)
    assert L.calc_bubble_par(kG, mP, sP, collect_data=False) is None
    assert relerr(bubble_par.data, bubble.data) <= 1e-14
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    assert relerr(bubble_par.data, ref_χ0) <= TOL

    # χ, γ (full_run.jl:23-40)
    serial, par = {}, {}
    for ch, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        serial[ch] = L.calc_χγ(ch, Γ, bubble, kG, mP, sP, ctx=ctx)
        par[ch] = L.calc_χγ_par(ch, Γ, kG, mP, sP)
        (χs, γs), (χp, γp) = serial[ch], par[ch]
        assert relerr(χp.data, χs.data) <= 1e-14 and relerr(γp.data, γs.data) <= 1e-14
        assert χp.tail_c == χs.tail_c and list(χp.usable_ω) == list(χs.usable_ω)
        rχ, rγ = O.calc_χγ(ch, Γ, ref_χ0, oK, omP, osP)
        assert relerr(χp.data, rχ) <= TOL and relerr(γp.data, rγ) <= TOL
        assert relerr(L.collect_χ(ch, kG, mP, sP).data, χp.data) == 0.0

    # Σ (full_run.jl:43-60): calc_Σ vs initialize_EoM + calc_Σ_par!, then calc_Σ_par
    with pytest.warns(UserWarning):
        λ0 = L.calc_λ0(bubble, inp.F_m, mP, sP, kG, ctx=ctx)
    (χm, γm), (χd, γd) = serial["m"], serial["d"]
    νmax = sP.n_iν
    Σ_ladder = L.calc_Σ(χm, γm, χd, γd, inp.χloc_m_sum, λ0, inp.G_rfft, kG, mP, sP, νmax=νmax, λm=0.123, λd=1.234,
                        ctx=ctx)
    L.initialize_EoM(inp.G_rfft, inp.χloc_m_sum, λ0, range(0, νmax), kG, mP, sP, χ_m=χm, γ_m=γm, χ_d=χd, γ_d=γd,
                     force_reinit=True)
    with pytest.raises(RuntimeError, match="Cannot collect γ"):
        L.collect_γ("m", kG, mP, sP)
    Σ_inplace = np.empty_like(Σ_ladder)
    assert L.calc_Σ_par_inplace(Σ_inplace, λm=0.123, λd=1.234) is Σ_inplace
    assert relerr(Σ_inplace, Σ_ladder) <= 1e-14
    ref = O.calc_Σ(O.ChiT(χm.data, omP.beta, tail_c=list(χm.tail_c)), γm.data,
                   O.ChiT(χd.data, omP.beta, tail_c=list(χd.tail_c)), γd.data, inp.χloc_m_sum, λ0, inp.G_rfft, oK,
                   omP, osP, νmax=νmax, λm=0.123, λd=1.234, tc=True)
    assert relerr(Σ_inplace, ref) <= TOL
    for lm, ld, tc in ((10.123, 20.223, False), (10.123, 20.223, True), (0.0, 0.0, True)):
        Σp = L.calc_Σ_par(λm=lm, λd=ld, tc=tc)
        Σs = L.calc_Σ(χm, γm, χd, γd, inp.χloc_m_sum, λ0, inp.G_rfft, kG, mP, sP, νmax=νmax, λm=lm, λd=ld, tc=tc,
                      ctx=ctx)
        assert relerr(Σp, Σs) <= 1e-14, (lm, ld, tc)
    assert L.calc_Σ_par(collect_data=False) is None
    # χ is never written (full_run.jl:65-77 checksum invariance)
    assert relerr(L.collect_χ("m", kG, mP, sP).data, χm.data) == 0.0

    # initialize_EoM again from the χ, γ that calc_χγ_par left resident
    L.clear_wcache()
    for k, v in (("kG", kG), ("mP", mP), ("sP", sP), ("G_fft", inp.G_fft), ("G_fft_reverse", inp.G_rfft)):
        L.update_wcache(k, v, ctx=ctx)
    L.calc_bubble_par(kG, mP, sP, collect_data=False)
    assert L.calc_χγ_par("m", inp.Γ_m, kG, mP, sP, collect_data=False) is None
    L.calc_χγ_par("d", inp.Γ_d, kG, mP, sP, collect_data=False)
    L.initialize_EoM(inp.G_rfft, inp.χloc_m_sum, λ0, range(0, νmax), kG, mP, sP)
    assert relerr(L.calc_Σ_par(λm=0.123, λd=1.234), Σ_ladder) <= 1e-14
    L.clear_wcache()
    with pytest.raises(RuntimeError, match="initialize_EoM"):
        L.calc_Σ_par()


def test_par_api_reference_semantics(ctx):
    """set_reference_semantics(True): calc_χγ_par raises as the reference's worker kernel does without the
    BSE_SC helper (ladderDGATools.jl:92), and calc_Σ_par(tc=False) returns the ω sum without the 1/β
    (ladderDGATools.jl:189), the Hartree term of collect_Σ! (:201-215) unscaled: β (Σ - H) + H of the
    default result, checked against the oracle's sum.  tc=True is unchanged."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _inputs("tiny", dict(D=2, Ns=6, n_iν=5, n_iω=4, n=0.9, U=1.4))
    kG, mP, sP = inp.kG, inp.mP, inp.sP
    L.clear_wcache()
    for k, v in (("kG", kG), ("mP", mP), ("sP", sP), ("G_fft", inp.G_fft), ("G_fft_reverse", inp.G_rfft)):
        L.update_wcache(k, v, ctx=ctx)
    bubble = L.calc_bubble_par(kG, mP, sP)
    (χm, γm), (χd, γd) = L.calc_χγ_par("m", inp.Γ_m, kG, mP, sP), L.calc_χγ_par("d", inp.Γ_d, kG, mP, sP)
    with pytest.warns(UserWarning):
        λ0 = L.calc_λ0(bubble, inp.F_m, mP, sP, kG, ctx=ctx)
    νmax = sP.n_iν
    L.initialize_EoM(inp.G_rfft, inp.χloc_m_sum, λ0, range(0, νmax), kG, mP, sP, χ_m=χm, γ_m=γm, χ_d=χd, γ_d=γd,
                     force_reinit=True)
    default = {tc: L.calc_Σ_par(λm=0.5, λd=0.25, tc=tc) for tc in (False, True)}
    try:
        L.set_reference_semantics(True)
        with pytest.raises(RuntimeError, match="BSE_Asym_Helper"):
            L.calc_χγ_par("m", inp.Γ_m, kG, mP, sP)
        ref = {tc: L.calc_Σ_par(λm=0.5, λd=0.25, tc=tc) for tc in (False, True)}
    finally:
        L.set_reference_semantics(False)
    H = mP.n * mP.U / 2.0
    assert relerr(ref[True], default[True]) == 0.0
    assert relerr(ref[False], (default[False] - H) * mP.β + H) <= 1e-14
    o = O.calc_Σ(O.ChiT(χm.data, omP.beta, tail_c=list(χm.tail_c)), γm.data,
                 O.ChiT(χd.data, omP.beta, tail_c=list(χd.tail_c)), γd.data, inp.χloc_m_sum, λ0, inp.G_rfft, oK, omP,
                 osP, νmax=νmax, λm=0.5, λd=0.25, tc=False)
    assert relerr(ref[False] - H, (o - omP.n * omP.U / 2.0) * omP.beta) <= TOL
    L.clear_wcache()


def test_par_api_rejects_partial_usable_range(ctx):
    """initialize_EoM refuses χ whose usable ω range is not the whole interval, as the reference's does
    (parallelization_helpers.jl:242-244)."""
    import ladderdga_hip as L
    inp, oK, omP, osP = _inputs("tiny", dict(D=2, Ns=4, n_iν=4, n_iω=3))
    kG, mP, sP = inp.kG, inp.mP, inp.sP
    Nq, Nν, Nω = kG.Nq, 2 * sP.n_iν, 2 * sP.n_iω + 1
    χm = L.χT(np.asfortranarray(np.full((Nq, Nω), 0.1)), mP.β, tail_c=[0, 0, mP.Ekin_DMFT], usable_ω=range(1, Nω - 1))
    χd = L.χT(np.asfortranarray(np.full((Nq, Nω), 0.1)), mP.β, tail_c=[0, 0, mP.Ekin_DMFT])
    γ = L.γT(np.asfortranarray(np.ones((Nq, Nν, Nω), complex)))
    L.clear_wcache()
    L.update_wcache("kG", kG, ctx=ctx)
    with pytest.raises(RuntimeError, match="usable ranges"):
        L.initialize_EoM(inp.G_rfft, 0.0, np.ones((Nq, Nν, Nω), complex), range(0, sP.n_iν), kG, mP, sP,
                         χ_m=χm, γ_m=γ, χ_d=χd, γ_d=γ)
    L.clear_wcache()
