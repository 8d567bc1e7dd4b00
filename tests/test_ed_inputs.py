"""The JLD2 reader and the setup_LDGA-style inputs on the reference's ED test data
(tests/golden/b1u2_ED_data.jld2, copied from test/test_data).

Pins: every dataset reads with the shape, type and values the reference's loaders expect
(IO.jl:67-87, runHelpers.jl:89-111); the impurity energies from the file reproduce the ED
golden values of test/thermodynamics.jl:1-5 (ed_out.csv) with the reference's tolerance; the
local Bethe-Salpeter equation closes on the file's data (inv(Γ + diag(1/χ₀)) = χ_DMFT with
Γ = -Γ_file, χ₀ = -β g(ν) g(ν+ω)), which fixes the (ω,ν,ν') -> (ν,ν',ω) permutation, the
sign flip of runHelpers.jl:97-99 and the frequency shift of the bubble.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import ldga_oracle as O

GOLD = Path(__file__).resolve().parent / "golden"
ED = GOLD / "b1u2_ED_data.jld2"


def test_jld2_reader_datasets():
    from ladderdga_hip.jld2 import jldopen
    f = jldopen(ED)
    assert set(f.keys()) == {"Γch", "Γsp", "χDMFTch", "χDMFTsp", "χ_ch_asympt", "χ_sp_asympt", "χ_pp_asympt",
                             "gImp", "g0", "ϵₖ", "Vₖ", "μ", "U", "β", "nden", "E_kin_DMFT", "E_pot_DMFT",
                             "grid_shift", "grid_nBose", "grid_nFermi"}
    for k in ("Γch", "Γsp", "χDMFTch", "χDMFTsp"):
        a = f[k]
        assert a.shape == (21, 20, 20) and a.dtype == np.complex128 and a.flags.f_contiguous
        assert np.all(np.isfinite(a))
    for k in ("gImp", "g0"):
        assert f[k].shape == (84,) and f[k].dtype == np.complex128
    assert f["ϵₖ"].shape == (4,) and f["Vₖ"].dtype == np.float64
    assert (f["U"], f["μ"], f["β"], f["nden"]) == (2.0, 1.0, 1.0, 1.0)
    assert (f["grid_nBose"], f["grid_nFermi"], f["grid_shift"]) == (10, 10, 1)
    assert isinstance(f["grid_nBose"], int) and isinstance(f["U"], float)
    # half filling: particle-hole symmetric G_imp is purely imaginary, Im G < 0 for ν_n > 0
    g = f["gImp"]
    assert np.max(np.abs(g.real)) < 1e-12 and np.all(g.imag < 0)
    # χ(ν,ν',ω) of the file is symmetric under ν <-> ν' (time reversal)
    x = f["χDMFTsp"]
    assert np.max(np.abs(x - np.transpose(x, (0, 2, 1)))) <= 1e-10 * np.max(np.abs(x))


def test_jld2_reader_rejects_other_files(tmp_path):
    from ladderdga_hip.jld2 import jldopen
    p = tmp_path / "x.jld2"
    p.write_bytes(b"\x00" * 4096)
    with pytest.raises(ValueError, match="no HDF5 superblock"):
        jldopen(p)


def test_ed_energies_golden():
    """calc_E_ED (thermodynamics.jl:46-64) on the file's bath and G_imp against the ED values:
    Ek = -0.10448, Ep = U D = 2 x 0.14177 (atol 0.001, as test/thermodynamics.jl:3-4)."""
    from ladderdga_hip.jld2 import jldopen
    f = jldopen(ED)
    ek_ed, docc = (float(v) for v in (GOLD / "ed_out.csv").read_text().split("\n")[1].split())
    Ek, Ep = O.calc_E_ED(f["ϵₖ"], f["Vₖ"], f["gImp"], f["U"], f["nden"], f["μ"], f["β"])
    assert Ek == pytest.approx(ek_ed, abs=1e-3)
    assert Ep == pytest.approx(2 * docc, abs=1e-3)
    # and the file's own DMFT energies agree with them at the same level
    assert f["E_pot_DMFT"] == pytest.approx(Ep, abs=1e-3) and f["E_kin_DMFT"] == pytest.approx(Ek, abs=2e-3)


def test_ed_inputs_local_bse_closes():
    from ladderdga_hip.ed_inputs import setup_LDGA
    inp, imp_density, χDMFT_d = setup_LDGA(ED)
    β, Nν, nB = inp.mP.β, 2 * inp.cfg.n_iν, inp.cfg.n_iω
    assert inp.kG.D == 3 and inp.kG.Ns == 10 and inp.kG.Nq == 56
    assert inp.sP.fft_range == range(-40, 41) and inp.mP.sVk == pytest.approx(0.2496056419, rel=1e-9)
    ωn = np.arange(-nB, nB + 1)
    sh = np.trunc(ωn / 2).astype(int)
    f0 = inp.sP.fft_range.start
    νn = np.arange(Nν)[:, None] - inp.cfg.n_iν - sh[None, :]
    χ0 = -β * inp.gImp[νn - f0] * inp.gImp[νn + ωn[None, :] - f0]
    for Γ, χD in ((inp.Γ_m, inp.χDMFT_m), (inp.Γ_d, χDMFT_d)):
        A = np.transpose(Γ, (2, 0, 1)).copy()
        i = np.arange(Nν)
        A[:, i, i] += (1.0 / χ0).T
        inv = np.transpose(np.linalg.inv(A), (1, 2, 0))
        assert np.max(np.abs(inv - χD)) <= 1e-13 * np.max(np.abs(χD))
    # with the sign of Γ not flipped the equation does not close
    A = np.transpose(-inp.Γ_m, (2, 0, 1)).copy()
    A[:, i, i] += (1.0 / χ0).T
    assert np.max(np.abs(np.transpose(np.linalg.inv(A), (1, 2, 0)) - inp.χDMFT_m)) > 0.1 * np.max(np.abs(inp.χDMFT_m))
    # Σ_loc = 1/g0 - 1/gImp (Σ_Dyson): at half filling Re Σ = U/2
    assert np.max(np.abs(inp.Σ_loc.real - inp.mP.U / 2)) < 1e-9
    # the same-spin density correlator of the plain χ_DMFT sums, and the local Σ_ω χm
    assert 0.2 < imp_density < 0.25 and 0.2 < inp.χloc_m_sum < 0.25
