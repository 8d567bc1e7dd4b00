"""GPU parity: libldga_hip (through its C ABI) against the CPU oracle on the
same seeded inputs.  FP64 bar: normwise relative error <= 1e-10 on χ₀, χ, γ,
λ₀ and Σ_ladder (the tolerance of the benchmark description)."""
import math

import numpy as np
import pytest

from oracle import ldga_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


_cache = {}


def setup(name, **kw):
    """Product inputs + the oracle's view of the same grid/parameters."""
    key = (name, tuple(sorted(kw.items())))
    if key in _cache:
        return _cache[key]
    from ladderdga_hip import synthetic
    cfg = synthetic.CONFIGS[name] if name in synthetic.CONFIGS else synthetic.tiny_config(**kw)
    inp = synthetic.build_inputs(cfg)
    oK = O.KGrid(cfg.D, cfg.Ns, cfg.t, wedge_coords=inp.kG.wedge)
    omP = O.ModelParameters(inp.mP.U, inp.mP.μ, inp.mP.β, inp.mP.n, 0, 0, inp.mP.Ekin_DMFT)
    osP = O.SimulationParameters(cfg.n_iω, cfg.n_iν, 0, cfg.shift)
    _cache[key] = (inp, oK, omP, osP)
    return _cache[key]


TINY = [dict(D=2, Ns=4, n_iν=4, n_iω=3), dict(D=2, Ns=6, n_iν=5, n_iω=4, n=0.9, U=1.4),
        dict(D=3, Ns=4, n_iν=3, n_iω=2, t=0.4), dict(D=2, Ns=8, n_iν=6, n_iω=5, shift=0)]


def case(i):
    return ("tiny", TINY[i])


CASES = [case(i) for i in range(len(TINY))] + [("2d16", {})]


@pytest.mark.parametrize("name,kw", CASES)
def test_g_tables(ctx, name, kw):
    """G_fft/G_rfft (helpers.jl:153-183) on the GPU from a wedge G: the local G of the inputs gives the
    tables the inputs were built with on the full grid; a random k-dependent G matches the oracle."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name, **kw)
    print(f"g tables {name} {kw}", flush=True)
    f, r = L.G_fft(np.asfortranarray(inp.gLoc), inp.kG, inp.sP, ctx=ctx)
    assert relerr(f, inp.G_fft) <= 1e-12 and relerr(r, inp.G_rfft) <= 1e-12
    rng = np.random.default_rng(5)
    Gw = np.asfortranarray(rng.standard_normal(inp.gLoc.shape) + 1j * rng.standard_normal(inp.gLoc.shape))
    f, r = L.G_fft(Gw, inp.kG, inp.sP, ctx=ctx)
    rf, rr = O.G_fft_from_wedge(Gw, oK)
    assert relerr(f, rf) <= 1e-12 and relerr(r, rr) <= 1e-12
    # the tables feed the bubble as the reference's do
    χ0 = L.calc_bubble(f, r, inp.kG, inp.mP, inp.sP, ctx=ctx)
    assert relerr(χ0.data, O.calc_bubble(rf, rr, oK, omP, osP)) <= TOL


@pytest.mark.parametrize("name", ["2d64", "3d24"])
def test_g_tables_large(ctx, name):
    """Full-size G tables from the wedge G_loc against the full-grid numpy tables."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name)
    print(f"g tables large {name}", flush=True)
    f, r = L.G_fft(np.asfortranarray(inp.gLoc), inp.kG, inp.sP, ctx=ctx)
    assert relerr(f, inp.G_fft) <= 1e-12 and relerr(r, inp.G_rfft) <= 1e-12


@pytest.mark.parametrize("name,kw", CASES)
def test_bubble(ctx, name, kw):
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name, **kw)
    print(f"bubble {name} {kw}", flush=True)
    χ0 = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    ref = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    assert relerr(χ0.data, ref) <= TOL


@pytest.mark.parametrize("name,kw", CASES)
@pytest.mark.parametrize("ch", ["m", "d"])
def test_bse_fp64(ctx, name, kw, ch):
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name, **kw)
    print(f"bse {ch} {name} {kw}", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    Γ = inp.Γ_m if ch == "m" else inp.Γ_d
    rep = {}
    χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, ctx=ctx, report=rep)
    rχ, rγ = O.calc_χγ(ch, Γ, ref_χ0, oK, omP, osP)
    assert rep["n_zero_pivot"] == 0
    assert relerr(χ.data, rχ) <= TOL
    assert relerr(γ.data, rγ) <= TOL
    assert χ.tail_c == [0.0, 0.0, inp.mP.Ekin_DMFT]


@pytest.mark.parametrize("name,kw", CASES)
@pytest.mark.parametrize("ch", ["m", "d"])
def test_bse_mixed(ctx, name, kw, ch):
    """FP32 inverse on MFMA + FP64 refinement: same bar as FP64."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name, **kw)
    print(f"bse mixed {ch} {name} {kw}", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    Γ = inp.Γ_m if ch == "m" else inp.Γ_d
    rep = {}
    χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, precision="mixed", ctx=ctx, report=rep)
    rχ, rγ = O.calc_χγ(ch, Γ, ref_χ0, oK, omP, osP)
    assert rep["n_zero_pivot"] == 0 and rep["n_fallback"] == 0
    assert 1 <= rep["ir_steps_max"] <= 6
    assert relerr(χ.data, rχ) <= TOL
    assert relerr(γ.data, rγ) <= TOL



def test_bse_mixed_nnu160(ctx):
    """Nν = 160 (the 32^3 workload's matrix size) on a small 3D grid: mixed precision (FP32 MFMA inverse)
    and FP64 (the global-memory Gauss-Jordan kernel used above Nν = 128) vs LAPACK."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup("tiny", D=3, Ns=4, n_iν=80, n_iω=1, U=2.0, β=25.0, n=1.0, t=0.408248290463863)
    print("bse mixed Nν=160", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    for ch, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        rep = {}
        χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, precision="mixed", ctx=ctx, report=rep)
        rχ, rγ = O.calc_χγ(ch, Γ, ref_χ0, oK, omP, osP)
        assert rep["n_fallback"] == 0
        assert relerr(χ.data, rχ) <= TOL and relerr(γ.data, rγ) <= TOL
        χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, precision="fp64", ctx=ctx)
        assert relerr(χ.data, rχ) <= TOL and relerr(γ.data, rγ) <= TOL


@pytest.mark.parametrize("n_iν", [65, 100])
def test_bse_fp64_global_memory_kernel(ctx, n_iν):
    """FP64 above Nν = 128 (Nν = 130, 200: no tile or panel alignment) runs the global-memory
    Gauss-Jordan kernel; both channels against LAPACK."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup("tiny", D=3, Ns=4, n_iν=n_iν, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863)
    print(f"bse fp64 global memory Nν={2 * n_iν}", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    for ch, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        rep = {}
        χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, precision="fp64", ctx=ctx, report=rep)
        rχ, rγ = O.calc_χγ(ch, Γ, ref_χ0, oK, omP, osP)
        assert rep["n_zero_pivot"] == 0
        assert relerr(χ.data, rχ) <= TOL and relerr(γ.data, rγ) <= TOL


@pytest.mark.parametrize("n_iν", [1, 2, 13, 28, 36, 44, 52, 59, 60, 68])
def test_bse_every_tile_count(ctx, n_iν):
    """Every kernel instance NT = 3..9 (NT = max(3, ⌈Nν/16⌉); Nν = 26 and 118 leave a partial last
    tile and a partial last panel; Nν = 2 and 4 process a single panel) against LAPACK, both
    channels, FP64 up to Nν = 128 and mixed."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup("tiny", D=3, Ns=4, n_iν=n_iν, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863)
    Nν = 2 * n_iν
    print(f"bse tile count Nν={Nν}", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    precs = ["fp64", "mixed"] if Nν <= 128 else ["mixed"]
    for ch, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        rχ, rγ = O.calc_χγ(ch, Γ, ref_χ0, oK, omP, osP)
        for prec in precs:
            rep = {}
            χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, precision=prec, ctx=ctx, report=rep)
            assert relerr(χ.data, rχ) <= TOL and relerr(γ.data, rγ) <= TOL, (ch, prec)
            if prec == "mixed":
                assert rep["n_fallback"] == 0


def _make_ill_conditioned(Γ, χ0, targets, κ=1e9):
    """Γ with A(q, ω) = Γ[:,:,ω] + diag(1/χ₀[q,:,ω]) of condition number κ at each (q, ω) in targets
    (one per ω): a rank-1 change of Γ[:,:,ω] replaces A's smallest singular value by σ_max/κ."""
    Γ = np.array(Γ, order="F", copy=True)
    for q, w in targets:
        U, σ, Vh = np.linalg.svd(Γ[:, :, w] + np.diag(1.0 / χ0[q, :, w]))
        Γ[:, :, w] += (σ[0] / κ - σ[-1]) * np.outer(U[:, -1], Vh[-1, :])
    return Γ


# targets whose rank-1 change leaves every other matrix of its ω at κ < 5e4 (q points related by lattice
# symmetry share χ₀ and so their A; the test finds those through κ)
RESOLVE_CASES = [(dict(D=2, Ns=4, n_iν=4, n_iω=3), [(0, 0)]),
                 (dict(D=3, Ns=4, n_iν=60, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863), [(9, 2), (8, 0)]),
                 (dict(D=3, Ns=4, n_iν=80, n_iω=1, U=2.0, β=25.0, n=1.0, t=0.408248290463863), [(9, 1)])]


@pytest.mark.parametrize("kw,targets", RESOLVE_CASES)
def test_bse_mixed_refinement_failure_resolved_in_fp64(ctx, kw, targets):
    """Matrices of condition number 1e9 are beyond an FP32 inverse: their refinement fails.  The library
    re-solves exactly those with the FP64 kernel of that size (blocked MFMA up to Nν = 128, global memory
    above), so their χ, γ equal the FP64 call's bit for bit.  Every other (q, ω) meets the FP64 bar
    against LAPACK."""
    import warnings
    import ladderdga_hip as L
    inp, oK, omP, osP = setup("tiny", **kw)
    Nν = 2 * inp.sP.n_iν
    print(f"bse mixed re-solve Nν={Nν} targets={targets}", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    Γ = _make_ill_conditioned(inp.Γ_m, ref_χ0, targets)
    κ = np.array([[np.linalg.cond(Γ[:, :, w] + np.diag(1.0 / ref_χ0[q, :, w])) for w in range(Γ.shape[2])]
                  for q in range(ref_χ0.shape[0])])
    bad = [tuple(qw) for qw in np.argwhere(κ > 1e8)]
    assert set(targets) <= set(bad) and np.max(κ[κ <= 1e8]) < 5e4
    rep = {}
    with warnings.catch_warnings(record=True) as warned:
        warnings.simplefilter("always")
        χ, γ = L.calc_χγ("m", Γ, χ0, inp.kG, inp.mP, inp.sP, precision="mixed", ctx=ctx, report=rep)
    rχ, rγ = O.calc_χγ("m", Γ, ref_χ0, oK, omP, osP)
    assert rep["n_fallback"] == len(bad) and rep["n_zero_pivot"] == 0
    assert rep["n_fp64_resolved"] == rep["n_fallback"] and not warned
    χf, γf = L.calc_χγ("m", Γ, χ0, inp.kG, inp.mP, inp.sP, precision="fp64", ctx=ctx)
    for q, w in bad:
        assert χ.data[q, w] == χf.data[q, w] and np.array_equal(γ.data[q, :, w], γf.data[q, :, w])
        # FP64 against LAPACK at κ = 1e9: both are off the exact solution by up to ~κ 2^-53 Nν
        assert relerr(χ.data[q, w], rχ[q, w]) <= 1e-4 and relerr(γ.data[q, :, w], rγ[q, :, w]) <= 1e-4
    keep = np.ones(rχ.shape, dtype=bool)
    for q, w in bad:
        keep[q, w] = False
    assert relerr(χ.data[keep], rχ[keep]) <= TOL
    assert relerr(γ.data.transpose(0, 2, 1)[keep], rγ.transpose(0, 2, 1)[keep]) <= TOL


ZERO_PIVOT_CASES = [("tiny", TINY[0], 0), ("tiny", TINY[0], 1),
                    ("tiny", dict(D=3, Ns=4, n_iν=65, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863), 0),
                    ("tiny", dict(D=3, Ns=4, n_iν=65, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863), 1)]


@pytest.mark.parametrize("name,kw,precision", ZERO_PIVOT_CASES)
def test_bse_zero_pivot_reports_and_gives_nan(ctx, name, kw, precision):
    """A = Γ + diag(1/χ₀) exactly singular -> LDGA_ERR_ZERO_PIVOT status and NaN outputs: FP64 in the blocked
    kernel (Nν = 8) and the global-memory kernel (Nν = 130), and mixed, whose zero-pivot matrices are
    re-solved in FP64."""
    import ctypes as C
    import ladderdga_hip as L
    from ladderdga_hip import _lib
    from ladderdga_hip.api import _freq
    inp, oK, omP, osP = setup(name, **kw)
    print(f"bse zero pivot Nν={2 * inp.sP.n_iν} precision={precision}", flush=True)
    χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    Nν, Nω = 2 * inp.sP.n_iν, 2 * inp.sP.n_iω + 1
    Γ = np.zeros((Nν, Nν, Nω), dtype=complex, order="F")
    # q = 0 column of ω index 0: Γ = -diag(1/χ0(q=0)) makes that matrix exactly zero
    Γ[np.arange(Nν), np.arange(Nν), 0] = -1.0 / χ0[0, :, 0]
    χ = np.empty((inp.kG.Nq, Nω), order="F")
    γ = np.empty((inp.kG.Nq, Nν, Nω), dtype=complex, order="F")
    rep = _lib.BseReport()
    ctx.set_kgrid(inp.kG)
    χ0f = np.asfortranarray(χ0)  # keep the Fortran copy alive across the call
    rc = _lib.lib().ldga_chi_gamma(ctx.handle, 1, precision, C.byref(_freq(inp.sP)), inp.mP.U, inp.mP.β,
                                   Γ.ctypes.data, χ0f.ctypes.data, χ.ctypes.data, γ.ctypes.data, C.byref(rep))
    assert rc == _lib.LDGA_ERR_ZERO_PIVOT
    assert rep.n_zero_pivot >= 1
    if precision == 1:
        assert rep.n_fp64_resolved == rep.n_fallback >= 1
    # A(q, ω=0) = diag(1/χ0[q] - 1/χ0[0]): singular where a diagonal entry vanishes.  Two different χ0
    # values can have the same rounded reciprocal, so only clearly nonzero diagonals must give finite χ.
    dmin = np.min(np.abs(1.0 / χ0[:, :, 0] - 1.0 / χ0[0:1, :, 0]), axis=1)
    same = np.any(χ0[:, :, 0] == χ0[0:1, :, 0], axis=1)
    safe = dmin > 1e-8 * np.max(np.abs(1.0 / χ0[0, :, 0]))
    assert np.all(np.isnan(χ[same, 0]))
    assert np.all(np.isfinite(χ[safe, 0])) and np.all(np.isfinite(χ[:, 1:]))
    assert safe.sum() >= 1
    with pytest.warns(UserWarning):
        L.calc_χγ("m", Γ, L.χ0T(np.asfortranarray(χ0), inp.mP.β), inp.kG, inp.mP, inp.sP, ctx=ctx,
                  precision=["fp64", "mixed"][precision])


@pytest.mark.parametrize("name,kw", CASES[:3] + [("2d16", {})])
def test_lambda0(ctx, name, kw):
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name, **kw)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    with pytest.warns(UserWarning):
        l0 = L.calc_λ0(L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β), inp.F_m, inp.mP, inp.sP, inp.kG, ctx=ctx)
    ref = O.calc_λ0(ref_χ0, inp.F_m, omP, osP)
    assert relerr(l0, ref) <= TOL


@pytest.mark.parametrize("lam", [0.0, 0.37, -0.05])
def test_sum_kω(ctx, lam):
    import ladderdga_hip as L
    from ladderdga_hip import _lib
    inp, oK, omP, osP = setup("2d16")
    rng = np.random.default_rng(5)
    Nω = 2 * inp.sP.n_iω + 1
    d = np.asfortranarray(rng.uniform(0.05, 0.8, (inp.kG.Nq, Nω)))
    χ = L.χT(d, inp.mP.β, tail_c=[0.0, 0.0, 0.31])
    oχ = O.ChiT(d, inp.mP.β, tail_c=[0.0, 0.0, 0.31])
    got, deriv = L.sum_kω(inp.kG, χ, λ=lam, ctx=ctx, with_derivative=True)
    ref = O.sum_kω(oK, oχ, λ=lam if lam != 0 else None)
    assert got == pytest.approx(ref, rel=1e-13, abs=1e-14)
    # sum_ωk (DataTypes.jl:257-264): the oracle sums ω first per q, with sum_ω's tail handling
    assert L.sum_ωk(inp.kG, χ, λ=lam, ctx=ctx) == pytest.approx(O.sum_ωk(oK, oχ, λ=lam), rel=1e-12, abs=1e-14)
    h = 1e-6
    fd = (O.sum_kω(oK, oχ, λ=lam + h) - O.sum_kω(oK, oχ, λ=lam - h)) / (2 * h)
    assert deriv == pytest.approx(fd, rel=1e-6)
    # a partial usable range: the tail runs over the same range (as in λm_correction, conditions_new.jl:147-153)
    lo, hi = 5, 29
    idx = np.arange(-inp.sP.n_iω, inp.sP.n_iω + 1)[lo:hi + 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        tail = np.where(idx != 0, np.real(0.31 / (2j * np.pi * idx / inp.mP.β) ** 2), 0.0)
    out = np.zeros(3)
    ctx.check(_lib.lib().ldga_sum_kw(ctx.handle, d.ctypes.data, Nω, lo, hi, 0.31, inp.mP.β, lam, out.ctypes.data))
    tr = (lambda x: O.χ_λ(x, lam)) if lam != 0 else None
    assert out[0] == pytest.approx(O.sum_kω_raw(oK, d[:, lo:hi + 1], inp.mP.β, 0.31, tail, tr), rel=1e-13)
    dref = O.sum_kω_raw(oK, d[:, lo:hi + 1], inp.mP.β, 0.31, tail, lambda x: O.dχ_λ(x, lam))
    assert out[1] == pytest.approx(dref, rel=1e-13)
    # the reference's sum_kω(χT) cannot slice its tail to a partial usable range (DataTypes.jl:223-226)
    with pytest.raises(Exception):
        L.sum_kω(inp.kG, L.χT(d, inp.mP.β, tail_c=[0.0, 0.0, 0.31], usable_ω=range(lo, hi + 1)), ctx=ctx)


def _ladder_inputs(name, kw, seed=3):
    inp, oK, omP, osP = setup(name, **kw)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    Nq, Nν, Nω = inp.kG.Nq, 2 * inp.sP.n_iν, 2 * inp.sP.n_iω + 1
    rng = np.random.default_rng(seed)
    cm = np.asfortranarray(rng.uniform(0.05, 0.5, (Nq, Nω)))
    cd = np.asfortranarray(rng.uniform(0.05, 0.5, (Nq, Nω)))
    rc = lambda: np.asfortranarray(rng.standard_normal((Nq, Nν, Nω)) + 1j * rng.standard_normal((Nq, Nν, Nω)))
    return inp, oK, omP, osP, ref_χ0, cm, cd, rc(), rc(), rc()


@pytest.mark.parametrize("name,kw", CASES)
def test_sigma(ctx, name, kw):
    import ladderdga_hip as L
    inp, oK, omP, osP, _, cm, cd, gm, gd, l0 = _ladder_inputs(name, kw)
    print(f"sigma {name} {kw}", flush=True)
    c2 = inp.mP.Ekin_DMFT
    νmax = max(1, min(inp.sP.n_iν, (3 * (2 * inp.sP.n_iω + 1)) // 8))
    for lm, ld, tc in ((0.0, 0.0, True), (0.21, -0.03, True), (0.13, 0.4, False)):
        Σ = L.calc_Σ(L.χT(cm, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gm), L.χT(cd, inp.mP.β, tail_c=[0, 0, c2]),
                     L.γT(gd), 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, νmax=νmax, λm=lm, λd=ld, tc=tc,
                     ctx=ctx)
        ref = O.calc_Σ(O.ChiT(cm, omP.beta, tail_c=[0, 0, c2]), gm, O.ChiT(cd, omP.beta, tail_c=[0, 0, c2]), gd,
                       0.123, l0, inp.G_rfft, oK, omP, osP, νmax=νmax, λm=lm, λd=ld, tc=tc)
        assert relerr(Σ, ref) <= TOL
        assert relerr(Σ - inp.mP.n * inp.mP.U / 2, ref - omP.n * omP.U / 2) <= TOL


@pytest.mark.parametrize("name,kw", [case(0), case(2)])
def test_sigma_parts(ctx, name, kw):
    import ladderdga_hip as L
    inp, oK, omP, osP, _, cm, cd, gm, gd, l0 = _ladder_inputs(name, kw)
    c2 = inp.mP.Ekin_DMFT
    P = L.calc_Σ_parts(L.χT(cm, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gm), L.χT(cd, inp.mP.β, tail_c=[0, 0, c2]),
                       L.γT(gd), 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, λm=0.2, λd=0.1, ctx=ctx)
    ref = O.calc_Σ_parts(O.ChiT(cm, omP.beta, tail_c=[0, 0, c2]), gm, O.ChiT(cd, omP.beta, tail_c=[0, 0, c2]), gd,
                         0.123, l0, inp.G_rfft, oK, omP, osP, λm=0.2, λd=0.1)
    assert relerr(P, ref) <= TOL
    full = L.calc_Σ(L.χT(cm, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gm), L.χT(cd, inp.mP.β, tail_c=[0, 0, c2]),
                    L.γT(gd), 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, λm=0.2, λd=0.1, ctx=ctx)
    assert relerr(P.sum(axis=2), full) <= TOL


@pytest.mark.parametrize("name,kw", [case(0), case(2)])
def test_sigma_omega_every_eom(ctx, name, kw):
    """calc_Σ_ω! (ladderDGATools_singleCore.jl:49-69) with each of the reference's real-χ EoM functions
    (ladderDGATools.jl:260-271, eom_sp_01/_02, eom_sp and eom_ch included): the bare ω sum against the
    oracle's loop nest."""
    import ladderdga_hip as L
    inp, oK, omP, osP, _, cm, cd, gm, gd, l0 = _ladder_inputs(name, kw)
    νmax = inp.sP.n_iν
    for kind in L._lib.EOM_KINDS:
        print(f"calc_Σ_ω {name} {kind}", flush=True)
        got = L.calc_Σ_ω(kind, L.χT(cm, inp.mP.β), L.γT(gm), L.χT(cd, inp.mP.β), L.γT(gd), inp.G_rfft, l0,
                         inp.mP.U, inp.kG, inp.sP, ctx=ctx)
        ref = O.calc_Σ_ω(O.EOM[kind], cm, gm, cd, gd, inp.G_rfft, l0, omP.U, oK, osP, νmax)
        assert relerr(got, ref) <= TOL, kind
    with pytest.raises(ValueError, match="unknown EoM"):
        L.calc_Σ_ω("nope", L.χT(cm, inp.mP.β), L.γT(gm), L.χT(cd, inp.mP.β), L.γT(gd), inp.G_rfft, l0,
                   inp.mP.U, inp.kG, inp.sP, ctx=ctx)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pipeline_2d16(ctx, precision):
    """bubble -> BSE (m, d) -> λ₀ -> Σ on the 16x16 workload, against the oracle's chain."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup("2d16")
    χ0 = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    χm, γm = L.calc_χγ("m", inp.Γ_m, χ0, inp.kG, inp.mP, inp.sP, precision=precision, ctx=ctx)
    χd, γd = L.calc_χγ("d", inp.Γ_d, χ0, inp.kG, inp.mP, inp.sP, precision=precision, ctx=ctx)
    with pytest.warns(UserWarning):
        λ0 = L.calc_λ0(χ0, inp.F_m, inp.mP, inp.sP, inp.kG, ctx=ctx)
    Σ = L.calc_Σ(χm, γm, χd, γd, inp.χloc_m_sum, λ0, inp.G_rfft, inp.kG, inp.mP, inp.sP, νmax=15, ctx=ctx)
    rχ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    rχm, rγm = O.calc_χγ("m", inp.Γ_m, rχ0, oK, omP, osP)
    rχd, rγd = O.calc_χγ("d", inp.Γ_d, rχ0, oK, omP, osP)
    rλ0 = O.calc_λ0(rχ0, inp.F_m, omP, osP)
    c2 = omP.Ekin_DMFT
    rΣ = O.calc_Σ(O.ChiT(rχm, omP.beta, tail_c=[0, 0, c2]), rγm, O.ChiT(rχd, omP.beta, tail_c=[0, 0, c2]), rγd,
                  inp.χloc_m_sum, rλ0, inp.G_rfft, oK, omP, osP, νmax=15)
    for a, b in ((χ0.data, rχ0), (χm.data, rχm), (γm.data, rγm), (χd.data, rχd), (γd.data, rγd), (λ0, rλ0),
                 (Σ, rΣ)):
        assert relerr(a, b) <= TOL
    # χ is left untouched by calc_Σ (full_run.jl:65-77 checksum invariance)
    assert np.array_equal(χm.data, np.asfortranarray(χm.data))


@pytest.mark.parametrize("name", ["2d64", "3d24"])
def test_bubble_large(ctx, name):
    """Full bubble at the 64x64 / 24^3 sizes vs the oracle's FFT path."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name)
    print(f"bubble large {name}", flush=True)
    χ0 = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    ref = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    assert relerr(χ0.data, ref) <= TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bse_large_sampled(ctx, precision):
    """64x64, Nν=120: all (q,ω) on the GPU; 60 sampled matrices per channel through LAPACK."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup("2d64")
    print(f"bse large 2d64 {precision}", flush=True)
    χ0 = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    rng = np.random.default_rng(1)
    Nq, Nω = inp.kG.Nq, 2 * inp.sP.n_iω + 1
    pts = list(zip(rng.integers(0, Nq, 60), rng.integers(0, Nω, 60))) + [(0, inp.sP.n_iω), (Nq - 1, 0)]
    for ch, Γ in (("m", inp.Γ_m), ("d", inp.Γ_d)):
        rep = {}
        χ, γ = L.calc_χγ(ch, Γ, χ0, inp.kG, inp.mP, inp.sP, precision=precision, ctx=ctx, report=rep)
        print(f"  {ch}: report {rep}", flush=True)
        s = 1 if ch == "m" else -1
        worst = 0.0
        for q, w in pts:
            A = Γ[:, :, w] + np.diag(1.0 / χ0.data[q, :, w])
            inv, _, _ = O.inv_lapack(np.asfortranarray(A))
            rχ = np.real(inv.sum()) / omP.beta ** 2
            rγ = inv.sum(axis=0) / (χ0.data[q, :, w] * (1 + s * omP.U * rχ))
            worst = max(worst, abs(χ.data[q, w] - rχ) / abs(rχ), relerr(γ.data[q, :, w], rγ))
        assert worst <= TOL, worst


def test_sigma_large(ctx):
    """Σ at 64x64 (νmax = 45, the λdm window) vs the oracle's FFT EoM."""
    import ladderdga_hip as L
    inp, oK, omP, osP, _, cm, cd, gm, gd, l0 = _ladder_inputs("2d64", {})
    print("sigma large 2d64", flush=True)
    c2 = inp.mP.Ekin_DMFT
    Σ = L.calc_Σ(L.χT(cm, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gm), L.χT(cd, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gd),
                 0.2, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, νmax=45, λm=0.1, λd=0.05, ctx=ctx)
    ref = O.calc_Σ(O.ChiT(cm, omP.beta, tail_c=[0, 0, c2]), gm, O.ChiT(cd, omP.beta, tail_c=[0, 0, c2]), gd, 0.2,
                   l0, inp.G_rfft, oK, omP, osP, νmax=45, λm=0.1, λd=0.05)
    assert relerr(Σ, ref) <= TOL


# One k grid per octant-pass kind and specialisation: the matrix-core pass (2D 48^2; 64^2 is covered by the
# large tests above), the register-resident pass in 2D (24^2, 32^2: the last pass skips the line x = h and
# sums the corner apart) and 3D (16^3, 24^3, 32^3: the last pass transforms the lines x >= y only), and the
# general vector pass (2D 20^2, 3D 12^3, 3D 10^3 with Ns/2 odd).
PASS_KINDS = [dict(D=2, Ns=48, n_iν=3, n_iω=2), dict(D=2, Ns=24, n_iν=4, n_iω=3, shift=0),
              dict(D=2, Ns=32, n_iν=3, n_iω=2), dict(D=3, Ns=16, n_iν=3, n_iω=2, t=0.4),
              dict(D=3, Ns=24, n_iν=2, n_iω=2, t=0.4), dict(D=3, Ns=32, n_iν=2, n_iω=1, t=0.4),
              dict(D=2, Ns=20, n_iν=3, n_iω=2), dict(D=3, Ns=12, n_iν=3, n_iω=2, t=0.4),
              dict(D=3, Ns=10, n_iν=3, n_iω=2, t=0.4)]


@pytest.mark.parametrize("kw", PASS_KINDS)
def test_bubble_sigma_pass_kinds(ctx, kw):
    """Bubble and Σ (with λ shifts and the tail) against the oracle on every octant-pass kind."""
    import ladderdga_hip as L
    inp, oK, omP, osP, _, cm, cd, gm, gd, l0 = _ladder_inputs("tiny", kw)
    print(f"pass kinds {kw}", flush=True)
    χ0 = L.calc_bubble(inp.G_fft, inp.G_rfft, inp.kG, inp.mP, inp.sP, ctx=ctx)
    assert relerr(χ0.data, O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)) <= TOL
    c2 = inp.mP.Ekin_DMFT
    νmax = max(1, min(inp.sP.n_iν, (3 * (2 * inp.sP.n_iω + 1)) // 8))
    Σ = L.calc_Σ(L.χT(cm, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gm), L.χT(cd, inp.mP.β, tail_c=[0, 0, c2]), L.γT(gd),
                 0.123, l0, inp.G_rfft, inp.kG, inp.mP, inp.sP, νmax=νmax, λm=0.21, λd=-0.03, ctx=ctx)
    ref = O.calc_Σ(O.ChiT(cm, omP.beta, tail_c=[0, 0, c2]), gm, O.ChiT(cd, omP.beta, tail_c=[0, 0, c2]), gd, 0.123,
                   l0, inp.G_rfft, oK, omP, osP, νmax=νmax, λm=0.21, λd=-0.03)
    assert relerr(Σ, ref) <= TOL
    f, r = L.G_fft(np.asfortranarray(inp.gLoc), inp.kG, inp.sP, ctx=ctx)
    assert relerr(f, inp.G_fft) <= 1e-12 and relerr(r, inp.G_rfft) <= 1e-12


@pytest.mark.parametrize("kw", [dict(n_iω=3, n_iν=4, shift=1), dict(n_iω=5, n_iν=6, shift=0), dict(n_iω=4, n_iν=3, shift=1)])
@pytest.mark.parametrize("diag", [True, False])
def test_F_from_χ(ctx, kw, diag):
    """F_from_χ (BSETools.jl:40-53) on the GPU against the oracle, random local χ and G over fft_range."""
    import ladderdga_hip as L
    from ladderdga_hip.types import SimulationParameters
    print(f"F_from_χ {kw} diag={diag}", flush=True)
    osP = O.SimulationParameters(kw["n_iω"], kw["n_iν"], 0, kw["shift"])
    sP = SimulationParameters(kw["n_iω"], kw["n_iν"], 0, kw["shift"])
    assert sP.fft_range == osP.fft_range
    rng = np.random.default_rng(21)
    Nν, Nω, nf = 2 * kw["n_iν"], 2 * kw["n_iω"] + 1, len(osP.fft_range)
    χ = np.asfortranarray(rng.standard_normal((Nν, Nν, Nω)) + 1j * rng.standard_normal((Nν, Nν, Nω)))
    G = 1.0 / (1j * (2 * np.arange(osP.fft_range.start, osP.fft_range.stop) + 1) * np.pi / 4.0 + 0.3
               - 0.2 * rng.standard_normal(nf))
    F = L.F_from_χ(χ, G, sP, 4.0, diag_term=diag, ctx=ctx)
    ref = O.F_from_χ(χ, (G, osP.fft_range.start), osP, 4.0, diag_term=diag)
    assert relerr(F, ref) <= 1e-13


@pytest.mark.parametrize("name,kw", CASES[:3] + [("tiny", dict(D=2, Ns=4, n_iν=70, n_iω=1))])
def test_calc_gen_χ(ctx, name, kw):
    """calc_gen_χ (BSETools.jl:259-275): the full inverses against numpy's inv, and tied to calc_χγ:
    Re Σ_νν' χ[ν,ν',q,ω] / β² is the physical χ (BSETools.jl:226-230)."""
    import ladderdga_hip as L
    inp, oK, omP, osP = setup(name, **kw)
    print(f"gen_chi {name} {kw}", flush=True)
    ref_χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
    χ0 = L.χ0T(np.asfortranarray(ref_χ0), inp.mP.β)
    X = L.calc_gen_χ(inp.Γ_m, χ0, inp.kG, ctx=ctx)
    ref = O.calc_gen_χ(inp.Γ_m, ref_χ0)
    assert relerr(X, ref) <= 1e-10
    χ, _ = L.calc_χγ("m", inp.Γ_m, χ0, inp.kG, inp.mP, inp.sP, ctx=ctx)
    assert relerr(np.real(X.sum(axis=(0, 1))) / inp.mP.β ** 2, χ.data) <= 1e-10
