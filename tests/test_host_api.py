"""Host-side argument checks of the Python mirror that raise before any GPU work, with
the reference's messages where it has them (CPU only)."""
import math

import numpy as np
import pytest


def _χ(Nq=3, Nω=5, **kw):
    import ladderdga_hip as L
    return L.χT(np.asfortranarray(np.full((Nq, Nω), 0.2)), 10.0, tail_c=[0, 0, 0.5], **kw)


def test_par_api_needs_setup():
    """calc_Σ_par before initialize_EoM (ladderDGATools.jl:169-171) and calc_bubble_par without the
    G tables in the worker cache (ladderDGATools.jl:48) raise."""
    import ladderdga_hip as L
    L.clear_wcache()
    with pytest.raises(RuntimeError, match="Call initialize_EoM before calc_Σ_par"):
        L.calc_Σ_par()
    with pytest.raises(RuntimeError, match="Call initialize_EoM"):
        L.calc_Σ_par_inplace(np.zeros((3, 2), complex))
    with pytest.raises(RuntimeError, match="run lDGA_setup"):
        L.calc_bubble_par(None, None, None)
    with pytest.raises(ValueError, match="unsupported field"):
        L.update_wcache("Σ_loc", np.zeros(3))
    with pytest.raises(ValueError, match="νGrid"):
        L.initialize_EoM(None, 0.0, None, range(1, 4), None, None, type("sP", (), {"n_iν": 4})())


def test_run_sc_argument_checks():
    """run_sc rejects unknown types, the χ tail self-consistency and λ-corrected χ
    (conditions_new.jl:342-380)."""
    import ladderdga_hip as L
    χ = _χ()
    args = (χ, None, χ, None, None, 0.0, None, None, 0.0, None, None, None)
    with pytest.raises(ValueError, match="run_sc type"):
        L.run_sc(*args, type="sc")
    with pytest.raises(NotImplementedError, match="update_χ_tail"):
        L.run_sc(*args, update_χ_tail=True)
    with pytest.raises(RuntimeError, match="already set"):
        L.run_sc(_χ(λ=0.3), None, χ, None, None, 0.0, None, None, 0.0, None, None, None)


def test_λm_rhs_argument_checks():
    """λm_rhs (LambdaCorrection/helpers.jl:234-255): λd with an already corrected χd, an unknown rhs
    type, and the native rhs without a finite impurity density raise."""
    import ladderdga_hip as L
    with pytest.raises(RuntimeError, match="λd"):
        L.λm_rhs(0.1, _χ(), _χ(λ=0.5), None, None, None, λd=0.2)
    with pytest.raises(ValueError, match="λ_rhs"):
        L.λm_rhs(0.1, _χ(), _χ(), None, None, None, λ_rhs="dmft")
    with pytest.raises(ValueError, match="imp_density"):
        L.λm_rhs(math.nan, _χ(), _χ(), None, None, None)


def test_Σ_loc_correction():
    """ladderDGATools.jl:249-257: Σ_ladder - Σ_ladderLoc[1, ν] + Σ_loc[ν]; with Σ_ladderLoc equal to Σ_loc
    it is the identity, and a constant shift of Σ_ladderLoc shifts the result by minus that constant."""
    import ladderdga_hip as L
    rng = np.random.default_rng(3)
    Σ = rng.standard_normal((5, 4)) + 1j * rng.standard_normal((5, 4))
    Σloc = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    assert np.allclose(L.Σ_loc_correction(Σ, Σloc[None, :], Σloc), Σ, rtol=0, atol=1e-15)
    out = L.Σ_loc_correction(Σ, Σloc[None, :] + 0.25j, Σloc)
    assert np.allclose(out, Σ - 0.25j, rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        L.Σ_loc_correction(Σ, Σloc[None, :3], Σloc)


def test_find_usable_χ_interval_reference_vectors():
    """The reference's own cases (test/GFFit.jl:43-59), 1-based there, 0-based ranges here."""
    from ladderdga_hip.types import find_usable_χ_interval as f
    one = lambda r: list(range(r.start + 1, r.stop + 1))
    t0 = [0.6, 0.7, 0.8, 0.9, 1.0, 0.9, 0.8, 0.7, 0.6]
    t1 = [0.8, 0.7, 0.8, 0.9, 1.0, 0.9, 0.8, 0.7, 0.8]
    t2 = [-0.8, 0.7, 0.8, 0.9, 1.0, 0.9, 0.8, 0.7, -0.8]
    t3 = [-0.8, 0.7, 0.8, 0.9, 1.0, 0.9, 0.8, 0.7, -0.8, -0.9]
    t4 = [-0.8, 0.3, -0.2]
    assert one(f(t1)) == list(range(2, 9))
    assert one(f(t1, reduce_range_prct=0.0)) == list(range(2, 9))
    assert one(f(t1, reduce_range_prct=0.50)) == list(range(3, 8))
    assert one(f(t2, reduce_range_prct=0.0)) == list(range(2, 9))
    with pytest.raises(ValueError):
        f(t3)
    assert one(f(t0, reduce_range_prct=0.0)) == list(range(1, 10))
    assert one(f(t2, sum_type="full", reduce_range_prct=0.5)) == list(range(1, 10))
    assert one(f(t2, sum_type=(2, 8), reduce_range_prct=0.5)) == list(range(2, 9))
    assert one(f([-1.0, -2.0, -1.0])) == [2]
    assert one(f(t4, reduce_range_prct=0.9)) == [2]
    # over q: the intersection of the rows, each at reduce_range_prct 0.1 (GFFit.jl:71-73)
    import ladderdga_hip as L
    χ = L.χT(np.array([t0, t1]), 5.0, full_range=False)
    assert one(χ.usable_ω) == list(range(2, 9))
    assert list(L.χT(np.array([t0, t1]), 5.0).usable_ω) == list(range(9))


def test_update_tail():
    """update_tail! (DataTypes.jl:315-345): χ[:, ω] gains (c₂' - c₂)/(iω_n)² off ω = 0, the tail
    coefficient follows, a λ on χ is kept; and the tail-corrected sum changes by exactly -βΔ/12 (the
    analytic Σ_n≠0 1/(iω_n)² = -β²/12 of the tail term), checked with the oracle's sum."""
    import ladderdga_hip as L
    from oracle import ldga_oracle as O
    rng = np.random.default_rng(4)
    β, c, Δ = 7.0, 0.4, 0.15
    d = rng.uniform(0.1, 0.5, (6, 9))
    χ = L.χT(np.asfortranarray(d.copy()), β, tail_c=[0.0, 0.0, c])
    L.update_tail(χ, [0.0, 0.0, c + Δ])
    ωn = 2j * np.pi * np.arange(-4, 5) / β
    exp = d.copy()
    exp[:, np.arange(9) != 4] += np.real(Δ / ωn[np.arange(9) != 4] ** 2)
    assert np.allclose(χ.data, exp, rtol=0, atol=1e-15) and χ.tail_c[2] == c + Δ
    kG = O.KGrid(2, 4, 1.0)
    assert kG.Nq == 6
    s0 = O.sum_kω(kG, O.ChiT(d, β, tail_c=[0, 0, c]))
    s1 = O.sum_kω(kG, O.ChiT(np.asarray(χ.data), β, tail_c=[0, 0, c + Δ]))
    assert s1 - s0 == pytest.approx(-β * Δ / 12, rel=1e-12)
    # with a λ on χ: taken off, the tail replaced, put back
    χ2 = L.χT(np.asfortranarray(d.copy()), β, tail_c=[0.0, 0.0, c])
    L.χ_λ_inplace(χ2, 0.3)
    L.update_tail(χ2, [0.0, 0.0, c + Δ])
    assert χ2.λ == 0.3 and np.allclose(χ2.data, L.χ_λ(exp, 0.3), rtol=1e-14, atol=1e-15)
    with pytest.raises(ValueError, match="c_2"):
        L.update_tail(χ2, [0.0, 0.1, c])
