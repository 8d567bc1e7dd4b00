"""The C ABI library: it loads without a GPU, exports every symbol the header
declares, and its k-grid wedge matches the oracle's (CPU only)."""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import ldga_oracle as O

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "ldga_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|double|const char \*)\s*(ldga_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_header_symbols():
    from ladderdga_hip import _lib
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 17
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTED) == syms
    assert L.ldga_version() >= 10000


@pytest.mark.parametrize("D,Ns,Nq", [(2, 2, 3), (2, 16, 45), (2, 64, 561), (3, 24, 455), (2, 48, 325), (3, 32, 969)])
def test_default_wedge_matches_oracle(D, Ns, Nq):
    from ladderdga_hip import _lib
    w, m = _lib.default_wedge(D, Ns)
    assert w.shape == (Nq, D)
    assert m.sum() == Ns ** D
    oK = O.KGrid(D, Ns, 1.0, wedge_coords=w)
    assert np.array_equal(oK.kMult, m)
    assert {tuple(c) for c in w.tolist()} == {tuple(c) for c in O.KGrid(D, Ns, 1.0).wedge.tolist()}


def test_default_wedge_rejects_bad_grid():
    from ladderdga_hip import _lib
    L = _lib.lib()
    assert L.ldga_default_wedge(4, 8, None, None) == _lib.LDGA_ERR_SHAPE
    assert L.ldga_default_wedge(2, 7, None, None) == _lib.LDGA_ERR_SHAPE


def test_synthetic_G_tables_match_oracle():
    """The input builder's G tables equal the reference construction (runHelpers.jl:120-129)."""
    from ladderdga_hip import synthetic
    inp = synthetic.build_inputs(synthetic.tiny_config(D=2, Ns=6, n_iν=3, n_iω=2, n=0.9, U=1.3, β=4.0))
    kG = O.KGrid(2, 6, 1.0, wedge_coords=inp.kG.wedge)
    mP = O.ModelParameters(inp.mP.U, inp.mP.μ, inp.mP.β, inp.mP.n)
    sP = O.SimulationParameters(2, 3, 0, 1)
    gLoc, gf, gr = O.G_tables(kG, mP, sP, inp.Σ_loc)
    assert np.allclose(gf, inp.G_fft, atol=1e-13)
    assert np.allclose(gr, inp.G_rfft, atol=1e-13)
    assert np.allclose(gLoc, inp.gLoc, atol=1e-15)
    # the chemical potential reproduces the requested filling (GFTools.jl:255-273)
    n = O.filling_pos_kG(gLoc[:, -sP.fft_range.start:], kG, mP.U, mP.mu, mP.beta)
    assert n == pytest.approx(0.9, abs=1e-10)


def test_synthetic_local_chain_matches_oracle():
    """χ_DMFT, F and χloc_m_sum of the input builder vs the oracle's local chain."""
    from ladderdga_hip import synthetic
    inp = synthetic.build_inputs(synthetic.tiny_config(D=2, Ns=4, n_iν=3, n_iω=2, U=1.1, β=3.0))
    sP = O.SimulationParameters(2, 3, 0, 1)
    mP = O.ModelParameters(inp.mP.U, inp.mP.μ, inp.mP.β, inp.mP.n, 0, 0, inp.mP.Ekin_DMFT)
    χ0l = O.calc_bubble_local(inp.gImp, mP, sP)
    kl = O.KGrid(2, 2, 1.0)
    kl.Nk = 1
    χm, γm = O.calc_χγ("m", inp.Γ_m, χ0l, kl, mP, sP)
    gen = O.calc_gen_χ(inp.Γ_m, χ0l)[:, :, 0, :]
    assert np.allclose(gen, inp.χDMFT_m, atol=1e-12)
    F = O.F_from_χ(inp.χDMFT_m, (inp.gImp, sP.fft_range.start), sP, mP.beta)
    assert np.allclose(F, inp.F_m, rtol=1e-11)
    s = O.sum_ω(O.ChiT(χm, mP.beta, tail_c=[0, 0, mP.Ekin_DMFT]))[0]
    assert s == pytest.approx(inp.χloc_m_sum, rel=1e-12)


def test_G_shell_sum_native():
    """ldga_G_shell_sum (the library's trigamma) against scipy's polygamma form (GFTools.jl:281-283)."""
    from ladderdga_hip import _lib
    for N in (1, 2, 5, 16, 46, 121, 1000):
        for β in (1.3, 10.0, 30.0):
            assert _lib.lib().ldga_G_shell_sum(N, β) == pytest.approx(O.G_shell_sum(N, β), rel=1e-13)
