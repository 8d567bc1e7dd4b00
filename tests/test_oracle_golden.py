"""Pins the CPU oracle against the reference's own known-answer tests and
against what the mathematics fixes (closed forms, brute force, identities).

Reference tests restated here: test/GFTools.jl:1-34, test/helpers.jl:1-22,
test/DataTypes.jl:30-66, test/LambdaCorrection/helpers.jl:1-24,
test/LapackWrapper.jl:9-36.
"""
import math

import numpy as np
import pytest

from oracle import ldga_oracle as O

# test/runtests.jl:5-8
mP_1 = O.ModelParameters(1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7)


def sP(n_iw, n_inu, shell, shift):
    return O.SimulationParameters(n_iw, n_inu, shell, int(shift), fft_range=range(1, 4))


# --------------------------------------------------------------------------- GFTools
def test_filling_golden():
    """test/GFTools.jl:166-171: filling(G, kG_1, 1.1, 1.2, 1.3) ≈ 1.1706318843228878.

    Pins the square-lattice dispersion (-2t sum cos k), the wedge weights of
    kintegrate and the filling/G_shell_sum formulas at once.
    """
    kG_1 = O.gen_kgrid("2dsc-1.0", 2)
    G = np.array([[O.G_from_Σ(n, mP_1.beta, mP_1.mu, e, 0.0) for n in range(50)] for e in kG_1.eps_wedge])
    assert O.filling_kG(G, kG_1, 1.1, 1.2, 1.3) == pytest.approx(1.1706318843228878, rel=1e-12)


def test_filling_golden_fails_for_wrong_dispersion():
    """The same golden rejects a dispersion without the factor 2."""
    kG = O.KGrid(2, 2, 0.5)   # -2*0.5 sum cos k == -t sum cos k with t = 1
    G = np.array([[O.G_from_Σ(n, mP_1.beta, mP_1.mu, e, 0.0) for n in range(50)] for e in kG.eps_wedge])
    assert abs(O.filling_kG(G, kG, 1.1, 1.2, 1.3) - 1.1706318843228878) > 1e-3


def test_G_shell_sum_vs_naive():
    """test/GFTools.jl:162-169 (rtol 0.01) and the closed form beta psi1(N+1/2)/(2 pi^2)."""
    β = 11.1
    iν = O.iν_array(β, range(-50, 50))
    naive = np.real(np.sum(1.0 / iν ** 2)) / β + β / 4
    assert O.G_shell_sum(50, β) == pytest.approx(naive, rel=1e-2)
    from scipy.special import polygamma
    assert O.G_shell_sum(50, β) == pytest.approx(β * polygamma(1, 50.5) / (2 * math.pi ** 2), rel=1e-12)


def test_matsubara_arrays():
    """test/GFTools.jl:138-143."""
    assert np.allclose(O.iν_array(5.0, range(3)), 1j * np.array([1, 3, 5]) * math.pi / 5.0)
    assert np.allclose(O.iω_array(5.0, range(3)), 1j * np.array([0, 2, 4]) * math.pi / 5.0)


def test_G_from_Σ():
    """test/GFTools.jl:149-159."""
    mf1 = 3j * math.pi / mP_1.beta
    assert O.G_from_Σ(1, mP_1.beta, 1.2, 1.3, 1.4) == pytest.approx(1 / (mf1 + 1.2 - 1.3 - 1.4))


# --------------------------------------------------------------------------- helpers
def test_reduce_range():
    """test/helpers.jl:1-11."""
    cases = [((1, 3), 1.0, (2, 2)), ((1, 4), 1.0, (2, 3)), ((1, 5), 1.0, (3, 3)), ((1, 6), 1.0, (3, 4)),
             ((1, 10), 0.2, (2, 9)), ((1, 20), 0.1, (2, 19)), ((1, 21), 0.1, (2, 20)), ((1, 20), 0.5, (6, 15)),
             ((1, 21), 0.5, (6, 16))]
    for (a, b), p, (c, d) in cases:
        r = O.reduce_range(range(a, b + 1), p)
        assert (r.start, r.stop - 1) == (c, d)


def test_index_functions():
    """test/helpers.jl:13-22 and the grids of test/runtests.jl:13-16."""
    assert O.νnGrid(4, sP(1, 2, 3, True)) == list(range(-7, 3))
    kG_1 = O.gen_kgrid("2dsc-1.0", 2)
    assert tuple(kG_1.wedge[0]) == (0, 0)                     # q0_index == 1 (1-based)
    for shift in (0, 1):
        s = sP(4, 5, 2, shift)
        for ωi in range(1, 2 * 4 + 2):
            ν0 = O.ν0Index_of_ωIndex(ωi, s)
            ωn = ωi - 4 - 1
            grid = [j - shift * O.trunc_div2(ωn) for j in range(-5, 5)]
            assert grid[ν0 - 1] == 0


# --------------------------------------------------------------------------- chi-lambda, sums, root finding
def test_χ_λ_closed_forms():
    """test/LambdaCorrection/helpers.jl:3-15."""
    assert O.χ_λ(1.0, 0.0) == 1.0
    assert O.χ_λ(1.0, 1.0) == 0.5
    x = np.array([1.0, 2.0, 1.0])
    assert np.allclose(O.χ_λ(O.χ_λ(x, 1.1), -1.1), x)
    assert O.dχ_λ(2.0, 0.3) == pytest.approx(-(2.0 / 1.6) ** 2)


def test_newton_right():
    """test/LambdaCorrection/helpers.jl:17-26 (root of 2x^3 + x - 3 at 1)."""
    f = lambda x: 2 * x ** 3 + x - 3
    df = lambda x: 6 * x ** 2 + 1
    assert O.newton_right_1d(f, df, 0.0, -1.0) == pytest.approx(1.0)
    assert O.newton_right_1d(f, df, -0.9, -1.0) == pytest.approx(1.0)
    f2 = lambda x: np.array([2 * x[0] ** 3 + x[0] - 3, x[1]])
    assert np.allclose(O.newton_right_nd(f2, [0.0, 0.0], [-1.0, -1.0]), [1.0, 0.0])


def _χ_tail_case():
    """test/DataTypes.jl:94-104."""
    ωn = 2j * np.arange(-100, 101) * math.pi / 11.1
    t3 = np.zeros((4, len(ωn)), dtype=complex)
    with np.errstate(divide="ignore", invalid="ignore"):
        row = 1.2 / ωn + 2.3 / ωn ** 2
    t3[:, :] = row[None, :]
    t3[:, 100] = 0.0
    return np.real(t3)


def test_sum_ω_and_sum_kω_reference_cases():
    """test/DataTypes.jl:111-130 on the 3Dsc-1.1 Ns=2 grid."""
    kG = O.gen_kgrid("3Dsc-1.1", 2)
    d = _χ_tail_case()
    χ_test = O.ChiT(d.copy(), 1.0, tail_c=[0.0, 0.0, 2.3])
    χ_test2 = O.ChiT(d.copy(), 1.0, tail_c=[0.0, 0.0, 0.0])
    direct = d.sum(axis=1)
    assert np.allclose(O.sum_ω(χ_test2), direct)
    assert float(kG.kintegrate(direct)) == pytest.approx(O.sum_kω(kG, χ_test2))
    assert float(kG.kintegrate(direct)) == pytest.approx(O.sum_ωk(kG, χ_test2))
    assert float(kG.kintegrate((d ** 3).sum(axis=1))) == pytest.approx(
        O.sum_kω(kG, χ_test2, transform=lambda x: x ** 3))
    assert O.sum_kω(kG, χ_test) == pytest.approx(O.sum_ωk(kG, χ_test))


def test_bosonic_tail_identity():
    """sum_{n != 0} 1/(i w_n)^2 = -beta^2/12: for a pure c2/(i w)^2 function the tail-corrected
    sum equals the exact infinite sum, and the plain truncated sum converges to it."""
    β, c2 = 3.7, 0.8
    n = np.arange(-20000, 20001)
    iω = 2j * math.pi * n / β
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(n != 0, np.real(c2 / iω ** 2), 0.0)
    exact = -c2 * β ** 2 / 12 / β
    χ = O.ChiT(d[None, :].copy(), β, tail_c=[0.0, 0.0, c2])
    assert O.sum_ω(χ)[0] == pytest.approx(exact, rel=1e-14)
    χ2 = O.ChiT(d[None, :].copy(), β, tail_c=[0.0, 0.0, 0.0])
    assert O.sum_ω(χ2)[0] == pytest.approx(exact, rel=1e-4)


# --------------------------------------------------------------------------- LAPACK inverse
def test_inv_lapack():
    """test/LapackWrapper.jl:9-36: zgetrf+zgetri give the inverse (seeded here)."""
    rng = np.random.default_rng(3)
    for N in range(8, 14):
        A = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        inv, info, info2 = O.inv_lapack(np.asfortranarray(A))
        assert info == 0 and info2 == 0
        assert np.allclose(inv @ A, np.eye(N), atol=1e-12)
        assert np.allclose(inv, np.linalg.inv(A), atol=1e-12)


# --------------------------------------------------------------------------- k-space convolution
@pytest.mark.parametrize("D,Ns", [(2, 4), (2, 6), (3, 4)])
def test_conv_fft_matches_brute_force(D, Ns):
    """conv_fft (the reference tables + conv_post shift) == (1/Nk) sum_k a(k) b(k+q)."""
    kG = O.KGrid(D, Ns, 1.0)
    rng = np.random.default_rng(D * 100 + Ns)
    a = kG.expand(rng.standard_normal(kG.Nq) + 1j * rng.standard_normal(kG.Nq))
    b = kG.expand(rng.standard_normal(kG.Nq) + 1j * rng.standard_normal(kG.Nq))
    af = np.fft.fftn(a)
    br = np.fft.fftn(b[(slice(None, None, -1),) * D])
    assert np.allclose(O.conv_fft(kG, af, br), O.conv_brute(kG, a, b), atol=1e-13)


def test_wedge_counts_and_weights():
    """Wedge sizes (Ns/2+1)(Ns/2+2)/2 in 2D, (n+2)(n+1)n/6 in 3D; kMult sums to Nk."""
    for D, Ns, Nq in [(2, 16, 45), (2, 64, 561), (3, 24, 455), (2, 48, 325), (3, 32, 969)]:
        kG = O.KGrid(D, Ns, 1.0)
        assert kG.Nq == Nq
        assert kG.kMult.sum() == Ns ** D
        # every full-grid point maps to a wedge point with the same dispersion
        assert np.allclose(kG.expand(kG.eps_wedge), kG.eps_full)


def _tiny_inputs(D=2, Ns=4, n_inu=3, n_iw=2, U=1.0, β=4.0, seed=11):
    kG = O.KGrid(D, Ns, 1.0)
    sP_ = O.SimulationParameters(n_iw, n_inu, 0, 1)
    mP = O.ModelParameters(U, U / 2, β, 1.0, 0.0, 0.0, 0.3)
    last = sP_.fft_range.stop - 1
    Σ_loc = np.array([U / 2 + U * U / (4j * (2 * n + 1) * math.pi / β) for n in range(last + 1)])
    gLoc, gf, gr = O.G_tables(kG, mP, sP_, Σ_loc)
    return kG, sP_, mP, Σ_loc, gLoc, gf, gr, np.random.default_rng(seed)


@pytest.mark.parametrize("D,Ns", [(2, 4), (2, 6), (3, 4)])
def test_G_fft_from_wedge(D, Ns):
    """G_fft!/G_rfft! from a wedge G: equals the tables built on the full grid (local Σ), a
    brute-force DFT of expand(G) and of its reverse for a k-dependent G, and the shift identity
    G_rfft(R) = exp(2πi ΣR/Ns) G_fft(R) that an even G implies."""
    kG, sP_, mP, Σ_loc, gLoc, gf, gr, rng = _tiny_inputs(D=D, Ns=Ns)
    f, r = O.G_fft_from_wedge(gLoc, kG)
    assert np.allclose(f, gf, rtol=0, atol=1e-13) and np.allclose(r, gr, rtol=0, atol=1e-13)
    Gw = rng.standard_normal((kG.Nq, 3)) + 1j * rng.standard_normal((kG.Nq, 3))
    f, r = O.G_fft_from_wedge(Gw, kG)
    R = np.stack(np.meshgrid(*([np.arange(Ns)] * D), indexing="ij"), -1).reshape(-1, D)
    E = np.exp(-2j * np.pi * (R @ R.T) / Ns)          # E[R, k] = exp(-2πi k.R/Ns)
    for i in range(3):
        g = kG.expand(Gw[:, i]).reshape(-1)
        grev = np.array([kG.expand(Gw[:, i])[tuple(Ns - 1 - k)] for k in R])
        assert np.allclose(f[..., i].reshape(-1), E @ g, atol=1e-12)
        assert np.allclose(r[..., i].reshape(-1), E @ grev, atol=1e-12)
    shift = np.exp(2j * np.pi * R.sum(axis=1) / Ns).reshape((Ns,) * D)
    assert np.allclose(r, f * shift[..., None], atol=1e-12)


def test_mu_root_fallback():
    """The μ root: where a plain secant from x0 runs off (atan(x - 5) from 0 diverges to ~1e20), the
    bracketing fallback (step out, then Illinois) still finds the root; a well-behaved case is
    solved by the secant itself."""
    assert O.secant(lambda x: math.atan(x - 5.0), 0.0) == pytest.approx(5.0, abs=1e-12)
    assert O.secant(lambda x: math.atan(x + 3.0), 2.0) == pytest.approx(-3.0, abs=1e-12)
    assert O.secant(lambda x: x ** 3 + x - 2.0, 0.9) == pytest.approx(1.0, abs=1e-12)


def test_mu_fit_failure_gives_nan():
    """G_from_Σladder! returns NaN when find_zero fails (GFTools.jl:171-176): a Σ_ladder holding a NaN
    makes the filling NaN for every μ, and the fit reports μ = NaN; the same Σ without the NaN fits n."""
    kG = O.KGrid(2, 4, 1.0)
    sP_ = O.SimulationParameters(3, 4, 0, 1)
    mP = O.ModelParameters(1.0, 0.3, 3.0, 0.9)
    last = sP_.fft_range.stop - 1
    Σ_loc = 0.45 + 0.25 / (4j * (2 * np.arange(last + 1) + 1) * math.pi / mP.beta)
    Σ = np.tile(Σ_loc[:3], (kG.Nq, 1))
    μ, _ = O.G_from_Σladder(Σ, Σ_loc, kG, mP, sP_, fix_n=True)
    assert math.isfinite(μ)
    G = O.G_from_Σladder(Σ, Σ_loc, kG, mP, sP_, fix_n=False, μ=μ)[1]
    f0 = sP_.fft_range.start
    assert O.filling_pos_kG(G[:, -f0:-f0 + 3], kG, mP.U, μ, mP.beta) == pytest.approx(mP.n, abs=1e-12)
    Σ[1, 2] = np.nan
    assert math.isnan(O.G_from_Σladder(Σ, Σ_loc, kG, mP, sP_, fix_n=True)[0])


def test_bubble_local_identity():
    """kintegrate_q chi0(q,nu,w) = -beta G_loc(nu) G_loc(nu+w) (normalisation of the convolution)."""
    kG, sP_, mP, Σ_loc, gLoc, gf, gr, _ = _tiny_inputs(Ns=6)
    χ0 = O.calc_bubble(gf, gr, kG, mP, sP_)
    f0 = sP_.fft_range.start
    Gl = kG.kintegrate(gLoc, axis=0)
    for ωi, ωn in enumerate(range(-sP_.n_iw, sP_.n_iw + 1)):
        νr = [v - O.trunc_div2(ωn) for v in range(-sP_.n_inu, sP_.n_inu)]
        for νi, νn in enumerate(νr):
            lhs = kG.kintegrate(χ0[:, νi, ωi])
            assert lhs == pytest.approx(-mP.beta * Gl[νn - f0] * Gl[νn + ωn - f0], rel=1e-12)


def test_bubble_brute_force():
    """Bubble from the FFT tables == -beta (1/Nk) sum_k G(k,nu) G(k+q,nu+w) by brute force."""
    kG, sP_, mP, Σ_loc, gLoc, gf, gr, _ = _tiny_inputs(D=3, Ns=4, n_inu=2, n_iw=1)
    χ0 = O.calc_bubble(gf, gr, kG, mP, sP_)
    f0 = sP_.fft_range.start
    ωn, νi = 1, 2
    νn = -sP_.n_inu + νi - O.trunc_div2(ωn)
    ga = kG.expand(gLoc[:, νn - f0])
    gb = kG.expand(gLoc[:, νn + ωn - f0])
    assert np.allclose(χ0[:, νi, ωn + sP_.n_iw], -mP.beta * O.conv_brute(kG, ga, gb), atol=1e-13)


def test_bse_closed_forms():
    """Gamma = 0: A^-1 = diag(chi0) -> chi = Re sum chi0 / beta^2, gamma = 1/(1 + sU chi).
    Diagonal Gamma: A^-1 = diag(1/(Gamma_ii + 1/chi0_i))."""
    kG, sP_, mP, Σ_loc, gLoc, gf, gr, rng = _tiny_inputs()
    χ0 = O.calc_bubble(gf, gr, kG, mP, sP_)
    Nν, Nω = 2 * sP_.n_inu, 2 * sP_.n_iw + 1
    Γ0 = np.zeros((Nν, Nν, Nω), dtype=complex)
    for s, t in ((1, "m"), (-1, "d")):
        χ, γ = O.calc_χγ(t, Γ0, χ0, kG, mP, sP_)
        assert np.allclose(χ, np.real(χ0.sum(axis=1)) / mP.beta ** 2, rtol=1e-13)
        assert np.allclose(γ, 1.0 / (1 + s * mP.U * χ[:, None, :]), rtol=1e-12)
    dg = rng.standard_normal((Nν, Nω)) + 1j * rng.standard_normal((Nν, Nω))
    Γd = np.zeros((Nν, Nν, Nω), dtype=complex)
    Γd[np.arange(Nν), np.arange(Nν), :] = dg
    χ, γ = O.calc_χγ("m", Γd, χ0, kG, mP, sP_)
    ainv = 1.0 / (dg[None, :, :] + 1.0 / χ0)
    assert np.allclose(χ, np.real(ainv.sum(axis=1)) / mP.beta ** 2, rtol=1e-12)
    assert np.allclose(γ, ainv / (χ0 * (1 + mP.U * χ[:, None, :])), rtol=1e-12)


def test_eom_brute_force_and_identities():
    """EoM: FFT convolution == brute-force q sum; tc toggles exactly tail/(i nu); parts add up (full_run.jl:60-61)."""
    kG, sP_, mP, Σ_loc, gLoc, gf, gr, rng = _tiny_inputs(Ns=4, n_inu=3, n_iw=2)
    Nq, Nν, Nω = kG.Nq, 2 * sP_.n_inu, 2 * sP_.n_iw + 1
    cm = O.ChiT(rng.uniform(0.1, 0.5, (Nq, Nω)), mP.beta, tail_c=[0, 0, 0.3])
    cd = O.ChiT(rng.uniform(0.1, 0.5, (Nq, Nω)), mP.beta, tail_c=[0, 0, 0.3])
    gm = rng.standard_normal((Nq, Nν, Nω)) + 1j * rng.standard_normal((Nq, Nν, Nω))
    gd = rng.standard_normal((Nq, Nν, Nω)) + 1j * rng.standard_normal((Nq, Nν, Nω))
    l0 = rng.standard_normal((Nq, Nν, Nω)) + 1j * rng.standard_normal((Nq, Nν, Nω))
    Σ_t = O.calc_Σ(cm, gm, cd, gd, 0.17, l0, gr, kG, mP, sP_, νmax=3, λm=0.2, λd=0.1, tc=True)
    Σ_f = O.calc_Σ(cm, gm, cd, gd, 0.17, l0, gr, kG, mP, sP_, νmax=3, λm=0.2, λd=0.1, tc=False)
    χml = O.ChiT(O.χ_λ(cm.data, 0.2), cm.beta, tail_c=cm.tail_c)
    tail = -mP.U * (O.sum_kω(kG, χml) - 0.17)
    assert np.allclose(Σ_t - Σ_f, (tail / O.iν_array(mP.beta, range(3)))[None, :], atol=1e-14)
    # brute force at one (k, nu)
    f0 = sP_.fft_range.start
    ν = 1
    acc = np.zeros(Nq, dtype=complex)
    for ωi in range(Nω):
        ωn = ωi - sP_.n_iw
        νi = O.ν0Index_of_ωIndex(ωi + 1, sP_) - 1 + ν
        if νi > Nν - 1:
            continue
        K = O.EOM["full"](mP.U, gm[:, νi, ωi], gd[:, νi, ωi], O.χ_λ(cm.data[:, ωi], 0.2),
                          O.χ_λ(cd.data[:, ωi], 0.1), l0[:, νi, ωi])
        acc += O.conv_brute(kG, kG.expand(K), kG.expand(gLoc[:, ν + ωn - f0]))
    assert np.allclose(Σ_f[:, ν], acc / mP.beta + mP.n * mP.U / 2, atol=1e-12)
    parts = O.calc_Σ_parts(cm, gm, cd, gd, 0.17, l0, gr, kG, mP, sP_, λm=0.2, λd=0.1, tc=True)
    Σ_all = O.calc_Σ(cm, gm, cd, gd, 0.17, l0, gr, kG, mP, sP_, νmax=sP_.n_inu, λm=0.2, λd=0.1, tc=True)
    assert np.allclose(parts.sum(axis=2), Σ_all, atol=1e-12)


def test_eom_split_functions_identities():
    """The split EoM functions of ladderDGATools.jl:260-270 against each other (random complex γ, λ₀ and
    real χ): eom = eom_sp + eom_ch + eom_rest; eom_sp = eom_χ_m + eom_γ_m + 1.5 eom_rest_01; eom_ch =
    eom_χ_d + eom_γ_d - 0.5 eom_rest_01; eom_sp_01 = eom_sp / 3 and eom_sp_02 = 2 eom_sp / 3.  The full
    eom and the χ/γ/rest parts are pinned by the brute-force EoM and the parts sum above, so a wrong
    coefficient or sign in eom_sp_01, eom_sp_02, eom_sp or eom_ch fails here."""
    rng = np.random.default_rng(7)
    n = 64
    U = 1.7
    gm, gd, l0 = (rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(3))
    cm, cd = rng.uniform(-0.5, 0.8, n), rng.uniform(-0.5, 0.8, n)
    E = {k: f(U, gm, gd, cm, cd, l0) for k, f in O.EOM.items()}
    assert np.allclose(E["full"], E["sp"] + E["ch"] + E["rest"], atol=1e-13)
    assert np.allclose(E["sp"], E["χ_m"] + E["γ_m"] + 1.5 * E["rest_01"], atol=1e-13)
    assert np.allclose(E["ch"], E["χ_d"] + E["γ_d"] - 0.5 * E["rest_01"], atol=1e-13)
    assert np.allclose(E["sp_01"], E["sp"] / 3, atol=1e-13)
    assert np.allclose(E["sp_02"], 2 * E["sp"] / 3, atol=1e-13)


def test_calc_E_limits():
    """U -> 0: E_kin -> 2/Nk sum_k eps_k f(eps_k - mu); E_pot with a constant Sigma = Sigma_h * n_sigma."""
    kG = O.KGrid(2, 8, 1.0)
    β, μ = 4.0, 0.3
    sP_ = O.SimulationParameters(2, 2000, 0, 1)
    mP = O.ModelParameters(0.0, μ, β, 1.0)
    nf = len(sP_.fft_range)
    νall = np.arange(sP_.fft_range.start, sP_.fft_range.stop)
    G = 1.0 / (1j * (2 * νall[None, :] + 1) * math.pi / β + μ - kG.eps_wedge[:, None])
    Σ = np.zeros((kG.Nq, 2000), dtype=complex)
    E_kin, E_pot = O.calc_E(G, Σ, μ, kG, mP, sP_)
    fermi = 1.0 / (np.exp(β * (kG.eps_wedge - μ)) + 1)
    assert E_kin == pytest.approx(float(kG.kintegrate(2 * kG.eps_wedge * fermi)), abs=2e-5)
    assert E_pot == pytest.approx(0.0, abs=1e-12)
    del nf


@pytest.mark.parametrize("shift", [0, 1])
def test_F_from_χ_closed_forms(shift):
    """F_from_χ (BSETools.jl:40-53) on χ built from the four G factors it divides by: the local bubble
    χ = -β G(ν)G(ν+ω) δ_νν' gives F = 0, and χ = c G(ν)G(ν+ω)G(ν')G(ν'+ω) gives F = -c (diagonal term
    off).  A wrong ν index, ω offset or shift fails both."""
    sP = O.SimulationParameters(3, 4, 0, shift)
    rng = np.random.default_rng(11 + shift)
    first = sP.fft_range.start
    Gf = rng.standard_normal(len(sP.fft_range)) + 1j * rng.standard_normal(len(sP.fft_range))
    Nν, Nω = 2 * sP.n_inu, 2 * sP.n_iw + 1
    χ0 = np.zeros((Nν, Nν, Nω), complex)
    χc = np.zeros((Nν, Nν, Nω), complex)
    β, c = 3.0, 0.7 - 0.2j
    for ωi in range(1, Nω + 1):
        for νi in range(1, Nν + 1):
            ωn, νn = O.OneToIndex_to_Freq(ωi, νi, sP)
            χ0[νi - 1, νi - 1, ωi - 1] = -β * Gf[νn - first] * Gf[ωn + νn - first]
            for νpi in range(1, Nν + 1):
                _, νpn = O.OneToIndex_to_Freq(ωi, νpi, sP)
                χc[νi - 1, νpi - 1, ωi - 1] = c * (Gf[νn - first] * Gf[ωn + νn - first] * Gf[νpn - first]
                                                   * Gf[ωn + νpn - first])
    assert np.max(np.abs(O.F_from_χ(χ0, (Gf, first), sP, β))) <= 1e-12
    assert np.allclose(O.F_from_χ(χc, (Gf, first), sP, β, diag_term=False), -c, rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------- λ₀ (BSETools.jl:128-141)
@pytest.mark.parametrize("shell", [0, 2])
def test_calc_λ0_closed_forms(shell):
    """λ₀[q,ν',ω] = χ₀[q, shell+ν', ω]/β² Σ_ν F[ν,ν',ω] on F whose ν-sums are known:
    F ≡ c gives c Nν χ₀/β²; F diagonal gives F[ν',ν',ω] χ₀/β²; F nonzero only in row ν = 0 (not symmetric)
    gives F[0,ν',ω] χ₀/β², so a sum over the second index fails.  With shell > 0 the χ₀ columns of the
    shell hold other values, so an ignored or doubled shell offset fails too; a missing 1/β² fails all."""
    rng = np.random.default_rng(17 + shell)
    Nq, Nν, Nω, β = 5, 6, 3, 2.7
    χ0 = rng.standard_normal((Nq, Nν + 2 * shell, Nω)) + 1j * rng.standard_normal((Nq, Nν + 2 * shell, Nω))
    core = χ0[:, shell:shell + Nν, :]
    mP = O.ModelParameters(1.0, 0.0, β, 1.0)
    sP_ = O.SimulationParameters(1, Nν // 2, shell, 1)
    c = 0.3 - 1.1j
    l0 = O.calc_λ0(χ0, np.full((Nν, Nν, Nω), c), mP, sP_)
    assert np.allclose(l0, c * Nν * core / β ** 2, rtol=1e-13, atol=0)
    d = rng.standard_normal((Nν, Nω)) + 1j * rng.standard_normal((Nν, Nω))
    Fd = np.zeros((Nν, Nν, Nω), complex)
    Fd[np.arange(Nν), np.arange(Nν), :] = d
    assert np.allclose(O.calc_λ0(χ0, Fd, mP, sP_), d[None, :, :] * core / β ** 2, rtol=1e-13, atol=0)
    r = rng.standard_normal((Nν, Nω)) + 1j * rng.standard_normal((Nν, Nω))
    Fr = np.zeros((Nν, Nν, Nω), complex)
    Fr[0, :, :] = r
    assert np.allclose(O.calc_λ0(χ0, Fr, mP, sP_), r[None, :, :] * core / β ** 2, rtol=1e-13, atol=0)


# --------------------------------------------------------------------------- sum_ω on a usable slice
def test_sum_ω_strict_usable_range_known_answer():
    """sum_ω (DataTypes.jl:285-305) on a strict usable range, evaluated by hand: β = 2π puts ω_m = i m, so
    Re(c₂/ω_m²) = -c₂/m².  Nω = 5 (ω_m, m = -2..2), usable indices 2:4 (1-based), χ[q, 2:4] = (1, 2, 3),
    c₂ = 1.  The reference pairs slice element n with ωn_arr[n] (m = -2, -1, 0), so the tail terms
    subtracted are -1/4, -1 and none: ((1 + 2 + 3) + 1/4 + 1 - β²/12)/β = (7.25 - π²/3)/(2π).  Pairing
    each element with its own frequency (m = -1, 0, 1) would give (8 - π²/3)/(2π) instead."""
    β = 2 * math.pi
    d = np.array([[9.0, 1.0, 2.0, 3.0, -7.0], [9.0, 1.0, 2.0, 3.0, -7.0]])
    χ = O.ChiT(d, β, tail_c=[0.0, 0.0, 1.0], usable_ω=range(1, 4))
    want = (7.25 - math.pi ** 2 / 3) / (2 * math.pi)
    assert O.sum_ω(χ) == pytest.approx([want, want], rel=1e-15)
    assert abs(want - (8 - math.pi ** 2 / 3) / (2 * math.pi)) > 0.1
    # force_full_range: all five, each with its own frequency: (8 + 2(1/4 + 1) - π²/3)/(2π)
    assert O.sum_ω(χ, force_full_range=True)[0] == pytest.approx((8 + 2.5 - math.pi ** 2 / 3) / (2 * math.pi),
                                                                  rel=1e-15)
    # sum_ωk = kintegrate of the per-q sums; the same value on every q integrates to itself
    kG = O.gen_kgrid("2Dsc-1.0", 2)
    d3 = np.tile(d[0], (kG.Nq, 1))
    assert O.sum_ωk(kG, O.ChiT(d3, β, tail_c=[0.0, 0.0, 1.0], usable_ω=range(1, 4))) == pytest.approx(want,
                                                                                                    rel=1e-15)
    # λ maps χ before the sums: χ_λ = χ/(λχ + 1) on the slice values
    lam = 0.5
    wl = ((1 / 1.5 + 2 / 2 + 3 / 2.5) + 1.25 - math.pi ** 2 / 3) / (2 * math.pi)
    assert O.sum_ωk(kG, O.ChiT(d3, β, tail_c=[0.0, 0.0, 1.0], usable_ω=range(1, 4)), λ=lam) == pytest.approx(
        wl, rel=1e-14)
    with pytest.raises(ValueError):
        O.sum_ω(O.ChiT(d, β, tail_c=[0.0, 1.0, 1.0]))


def test_omega_subsets_partition_bubble_and_eom():
    """The ω-subset options (used to split the CPU baseline over processes) partition the passes:
    the subsets of a 3-way split add up to the full bubble and the full EoM sum, bit for bit
    per ω slice (each slice is computed by the same calls)."""
    kG, sP_, mP, Σ_loc, gLoc, gf, gr, rng = _tiny_inputs(Ns=4, n_inu=3, n_iw=2)
    Nq, Nν, Nω = kG.Nq, 2 * sP_.n_inu, 2 * sP_.n_iw + 1
    full = O.calc_bubble(gf, gr, kG, mP, sP_)
    parts = [O.calc_bubble(gf, gr, kG, mP, sP_, ωsel=set(range(r, Nω, 3))) for r in range(3)]
    for r, p in enumerate(parts):
        sel = list(range(r, Nω, 3))
        assert np.array_equal(p[:, :, sel], full[:, :, sel])
        assert not np.any(np.delete(p, sel, axis=2))
    cm = rng.uniform(0.1, 0.5, (Nq, Nω))
    gm = rng.standard_normal((Nq, Nν, Nω)) + 1j * rng.standard_normal((Nq, Nν, Nω))
    l0 = rng.standard_normal((Nq, Nν, Nω)) + 1j * rng.standard_normal((Nq, Nν, Nω))
    args = (O.EOM["full"], cm, gm, cm, gm, gr, l0, mP.U, kG, sP_, 3)
    per = O.calc_Σ_ω(*args, per_ω=True)
    for r in range(3):
        sel = list(range(r, Nω, 3))
        p = O.calc_Σ_ω(*args, per_ω=True, ωsel=set(sel))
        assert np.array_equal(p[:, :, sel], per[:, :, sel])
        assert not np.any(np.delete(p, sel, axis=2))
    tot = sum(O.calc_Σ_ω(*args, ωsel=set(range(r, Nω, 3))) for r in range(3))
    assert np.allclose(tot, O.calc_Σ_ω(*args), rtol=1e-14, atol=1e-14)
