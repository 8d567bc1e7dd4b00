"""Inputs of the ladder path from a DMFT/ED JLD2 file, as setup_LDGA builds them
(src/IO.jl:67-93 and src/runHelpers.jl:85-187), for plain sums (chi_asympt_method
"nothing", shell 0):

* parameters: U, μ, β, nden, sVk = Σ Vₖ², E_pot_DMFT, E_kin_DMFT from the file;
  n_iω = grid_nBose, n_iν = grid_nFermi, shift = grid_shift and
  fft_range = -2(n_iν + n_iω) : 2(n_iν + n_iω) (IO.jl:67-93);
* vertices: Γ_m = -permutedims(Γsp, (2,3,1)), Γ_d = -permutedims(Γch, (2,3,1)),
  χDMFT_m/d = permutedims(χDMFTsp/ch, (2,3,1)) (runHelpers.jl:93-100): the file
  stores (ω, ν, ν') and the opposite sign convention for Γ;
* Σ_loc = 1/g0 - 1/gImp (Σ_Dyson, runHelpers.jl:101-104, GFTools.jl:207-210);
* G(k, ν) = 1/(iν_n + μ - ε_k - Σ_loc(ν)) over fft_range (Σ_loc(-ν-1) = conj Σ_loc(ν))
  and its tables G_fft = fft(G), G_rfft = fft(reverse(G)) (runHelpers.jl:113-130);
* the local chain at Nk = 1 (runHelpers.jl:132-143): gImp over -rm..rm-1
  (rm = max |fft_range|, the negative half conj-reversed), F_m = F_from_χ(χDMFT_m, gImp),
  χ₀_loc = -β g(ν) g(ν+ω), χ_m_loc = Re Σ_νν' inv(Γ_m + diag(1/χ₀_loc)) / β² and
  χloc_m_sum = Σ_ω(χ_m_loc) with its c₂ = E_kin_DMFT tail (DataTypes.jl:293-301);
* the impurity density correlator for the native λm rhs (runHelpers.jl:165-187).

This runs on the host; it is the input side of the path, not the path.
"""
from __future__ import annotations

import math

import numpy as np

from .jld2 import jldopen
from .kgrid import gen_kGrid
from .synthetic import LadderConfig, LadderInputs, _Σ_at
from .types import ModelParameters, SimulationParameters, find_usable_χ_interval


def _subtract_tail(x, c, iω, power):
    """GFTools.jl:293-312: x - c/iω^power where iω != 0."""
    out = np.array(x, dtype=complex)
    nz = iω != 0
    out[nz] = out[nz] - c / iω[nz] ** power
    return out


def setup_LDGA(path, kgrid: str = "3Dsc-0.2041241452319315", Nk: int = 10):
    """LadderInputs (as synthetic.build_inputs returns them) and the impurity density."""
    f = jldopen(path)
    U, μ, β, n = float(f["U"]), float(f["μ"]), float(f["β"]), float(f["nden"])
    Vk = np.asarray(f["Vₖ"], dtype=float)
    Epot, Ekin = (float(f["E_pot_DMFT"]), float(f["E_kin_DMFT"])) if "E_kin_DMFT" in f else (0.0, 0.0)
    nB, nF, shift = int(f["grid_nBose"]), int(f["grid_nFermi"]), int(f["grid_shift"])
    kG = gen_kGrid(kgrid, Nk)
    sP = SimulationParameters(nB, nF, 0, shift)
    mP = ModelParameters(U, μ, β, n, float(np.sum(Vk ** 2)), Epot, Ekin)
    Nν, Nω = 2 * nF, 2 * nB + 1
    perm = lambda a: np.asfortranarray(np.transpose(np.asarray(a), (1, 2, 0)))
    Γ_m, Γ_d = perm(-f["Γsp"]), perm(-f["Γch"])
    χDMFT_m, χDMFT_d = perm(f["χDMFTsp"]), perm(f["χDMFTch"])
    if Γ_m.shape != (Nν, Nν, Nω):
        raise ValueError(f"vertex shape {Γ_m.shape} does not match grid_nFermi={nF}, grid_nBose={nB}")
    gImp_in = np.asarray(f["gImp"], dtype=complex)
    Σ_loc = 1.0 / np.asarray(f["g0"], dtype=complex) - 1.0 / gImp_in
    last = sP.fft_range.stop - 1
    if len(Σ_loc) < last + 1:
        raise ValueError(f"the file holds {len(Σ_loc)} Matsubara frequencies; fft_range needs {last + 1}")

    # G tables over fft_range (full grid)
    νf = np.arange(sP.fft_range.start, sP.fft_range.stop)
    Σf = _Σ_at(Σ_loc, νf)
    iνf = 1j * (2 * νf + 1) * math.pi / β
    axes = tuple(range(kG.D))
    G = 1.0 / (iνf + μ - kG.eps_full[..., None] - Σf)
    G_fft = np.asfortranarray(np.fft.fftn(G, axes=axes))
    G_rfft = np.asfortranarray(np.fft.fftn(G[(slice(None, None, -1),) * kG.D], axes=axes))
    gLoc = np.asfortranarray(G[tuple(kG.wedge.T)])
    del G

    # local chain: gImp over -rm .. rm-1, stored over fft_range (index ν - first(fft_range))
    rm = max(abs(sP.fft_range.start), abs(last))
    gImp = np.zeros(len(νf), dtype=complex)
    for i, ν in enumerate(νf):
        if -rm <= ν < 0:
            gImp[i] = np.conj(gImp_in[-ν - 1])
        elif 0 <= ν < rm:
            gImp[i] = gImp_in[ν]
    f0 = sP.fft_range.start
    ωn = np.arange(-nB, nB + 1)
    sh = shift * np.trunc(ωn / 2).astype(int)
    νn = np.arange(Nν)[:, None] - nF - sh[None, :]
    gg = gImp[νn - f0] * gImp[νn + ωn[None, :] - f0]          # g(ν) g(ν+ω), (Nν, Nω)
    χ0_loc = -β * gg
    F_m = -(χDMFT_m + np.eye(Nν)[:, :, None] * (β * gg)[:, None, :]) / (gg[:, None, :] * gg[None, :, :])
    F_m = np.asfortranarray(F_m)
    A = np.transpose(Γ_m, (2, 0, 1)).copy()
    idx = np.arange(Nν)
    A[:, idx, idx] += (1.0 / χ0_loc).T
    χ_m_loc = np.real(np.linalg.inv(A).sum(axis=(1, 2))) / β ** 2
    iω = 2j * math.pi * ωn / β
    with np.errstate(divide="ignore", invalid="ignore"):
        tail = np.where(ωn != 0, np.real(Ekin / iω ** 2), 0.0)
    χloc_m_sum = float((np.sum(χ_m_loc - tail) + Ekin * (-β ** 2 / 12)) / β)

    # impurity density correlator (plain mode: the χ_DMFT sums)
    χLoc_m = np.real(χDMFT_m.sum(axis=(0, 1))) / β ** 2
    χLoc_d = np.real(χDMFT_d.sum(axis=(0, 1))) / β ** 2
    um = find_usable_χ_interval(χLoc_m, reduce_range_prct=sP.usable_prct_reduction)
    ud = find_usable_χ_interval(χLoc_d, reduce_range_prct=sP.usable_prct_reduction)
    loc = range(max(um.start, ud.start), min(um.stop, ud.stop))
    χupup = 0.5 * (χLoc_m + χLoc_d)[loc.start:loc.stop]
    imp_density = float(np.real(np.sum(_subtract_tail(χupup, Ekin, iω[loc.start:loc.stop], 2))) / β
                        - Ekin * β / 12)

    cfg = LadderConfig(f"ED {path} on {kgrid} Nk={Nk}", kG.D, kG.Ns, kG.t, U, β, n, nF, nB, 0.0, 0, shift)
    inp = LadderInputs(cfg, kG, mP, sP, Σ_loc, gImp, G_fft, G_rfft, Γ_m, Γ_d, χDMFT_m, F_m, χloc_m_sum, gLoc)
    return inp, imp_density, χDMFT_d
