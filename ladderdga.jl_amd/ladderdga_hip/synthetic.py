"""Seeded synthetic inputs for the ladder path (the role of setup_LDGA).

This is the input side of the path (src/runHelpers.jl:85-201), not the hot
path: it builds, on the host, what setup_LDGA would read from a DMFT file
and tabulate:

* Σ_loc(ν) = nU/2 + U^2/(4 iν_n) (atomic-limit self-energy with the Hartree
  shift included), G(k,ν) = 1/(iν_n + μ - ε_k - Σ_loc(ν)) over fft_range and
  its tables G_fft = fft(G), G_rfft = fft(reverse(G)) (src/runHelpers.jl:113-130);
* μ = U/2 at half filling, otherwise the root of filling_pos(G) = n
  (src/GFTools.jl:255-268);
* local vertices Γ_d = U + δ_d, Γ_m = -U + δ_m, with δ(ν,ν',ω) the symmetrised
  complex Gaussian noise scaled c·U/(1+|ν-ν'|);
* the local chain at Nk = 1: χ₀_loc = -β g g, χ_DMFT = inv(Γ + diag(1/χ₀_loc)),
  F_m = F_from_χ(χ_DMFT_m, g) (src/BSETools.jl:40-53) and
  χloc_m_sum = sum_ω(χ_m_loc) (src/runHelpers.jl:134-143);
* Ekin_DMFT (the χ tail coefficient) from calc_E's kinetic part on G, Σ_loc
  (src/thermodynamics.jl:142-168).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy.special import polygamma

from .kgrid import KGrid
from .types import ModelParameters, SimulationParameters


@dataclass
class LadderConfig:
    name: str
    D: int
    Ns: int
    t: float
    U: float
    β: float
    n: float
    n_iν: int
    n_iω: int
    noise: float
    seed: int
    shift: int = 1


# The five workloads of the benchmark description (plain sums, shell 0).
CONFIGS = {
    "2d16": LadderConfig("2D square 16x16, U=1, beta=10, n=1", 2, 16, 1.0, 1.0, 10.0, 1.0, 20, 20, 0.10, 1000),
    "2d64": LadderConfig("2D square 64x64, U=2, beta=30, n=0.9", 2, 64, 1.0, 2.0, 30.0, 0.9, 60, 60, 0.15, 1001),
    "3d24": LadderConfig("3D cubic 24^3, U=1.5, beta=20, n=1", 3, 24, 1 / math.sqrt(6), 1.5, 20.0, 1.0, 40, 40, 0.15,
                         1002),
    "2d48": LadderConfig("2D square 48x48, U=2.5, beta=15, n=0.85", 2, 48, 1.0, 2.5, 15.0, 0.85, 50, 50, 0.10, 1003),
    "3d32": LadderConfig("3D cubic 32^3, U=2, beta=25, n=1", 3, 32, 1 / math.sqrt(6), 2.0, 25.0, 1.0, 80, 80, 0.10,
                         1004),
}


def tiny_config(D=2, Ns=4, n_iν=4, n_iω=3, U=1.0, β=5.0, n=1.0, noise=0.1, seed=7, t=1.0, shift=1):
    return LadderConfig(f"tiny {D}D Ns={Ns}", D, Ns, t, U, β, n, n_iν, n_iω, noise, seed, shift)


@dataclass
class LadderInputs:
    cfg: LadderConfig
    kG: KGrid
    mP: ModelParameters
    sP: SimulationParameters
    Σ_loc: np.ndarray       # ν = 0 .. last(fft_range)
    gImp: np.ndarray        # over fft_range
    G_fft: np.ndarray       # (Ns,..,Ns, Nfft) Fortran order
    G_rfft: np.ndarray
    Γ_m: np.ndarray         # (Nν, Nν, Nω) Fortran order
    Γ_d: np.ndarray
    χDMFT_m: np.ndarray
    F_m: np.ndarray
    χloc_m_sum: float
    gLoc: np.ndarray        # wedge G (Nq, Nfft)


def G_shell_sum(N, β):
    """src/GFTools.jl:281-283."""
    return float((polygamma(1, N + 0.5) - polygamma(1, 0.5 - N)) * β / (4 * math.pi ** 2) + β / 4)


def _Σ_loc(U, n, β, nmax):
    ν = (2 * np.arange(nmax + 1) + 1) * math.pi / β
    return n * U / 2 + U * U / (4j * ν)


def _Σ_at(Σ_loc, νn):
    νn = np.asarray(νn)
    out = np.where(νn >= 0, Σ_loc[np.clip(νn, 0, None)], np.conj(Σ_loc[np.clip(-νn - 1, 0, None)]))
    return out


def _filling_pos(Gk_pos, kMult, U, μ, β):
    """filling_pos(G::Matrix, kG, U, μ, β; improved_sum=true), src/GFTools.jl:255-273."""
    g = (kMult[:, None] * Gk_pos).sum(axis=0) / kMult.sum()
    sG = g.sum()
    shell = G_shell_sum(len(g) + 1, β)
    return 2 * (np.real(sG + np.conj(sG)) / β + 0.5 + μ * shell) / (1 + U * shell)


def build_inputs(cfg: LadderConfig) -> LadderInputs:
    kG = KGrid(cfg.D, cfg.Ns, cfg.t)
    sP = SimulationParameters(cfg.n_iω, cfg.n_iν, 0, cfg.shift)
    last = sP.fft_range.stop - 1
    Σ_loc = _Σ_loc(cfg.U, cfg.n, cfg.β, last)
    νpos = np.arange(last + 1)
    iν = 1j * (2 * νpos + 1) * math.pi / cfg.β

    def fill(μ):
        G = 1.0 / (iν[None, :] + μ - kG.eps_wedge[:, None] - Σ_loc[None, :])
        return _filling_pos(G, kG.kMult, cfg.U, μ, cfg.β)

    if cfg.n == 1.0:
        μ = cfg.U / 2
    else:
        lo, hi = -10.0, 10.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if fill(mid) < cfg.n:
                lo = mid
            else:
                hi = mid
        μ = 0.5 * (lo + hi)
    # Ekin from the kinetic part of calc_E on (G, Σ_loc), nu = 0..2 n_iν - 1
    νE = np.arange(2 * cfg.n_iν)
    iνE = 1j * (2 * νE + 1) * math.pi / cfg.β
    Σh = cfg.n * cfg.U / 2
    GE = 1.0 / (iνE[None, :] + μ - kG.eps_wedge[:, None] - Σ_loc[None, νE])
    ek = kG.eps_wedge
    Ekt = ek + Σh - μ
    E_kin_full = ek[:, None] * np.real(GE - Ekt[:, None] / iνE[None, :] ** 2)
    E_kin_tail_inv = (cfg.β / 2) * ek * (1 - cfg.β * Ekt)
    Ekin = float(np.dot(kG.kMult, 4 * E_kin_full.sum(axis=1) + E_kin_tail_inv) / kG.kMult.sum() / cfg.β)
    mP = ModelParameters(cfg.U, μ, cfg.β, cfg.n, 0.0, 0.0, Ekin)

    # G tables over fft_range (full grid), src/runHelpers.jl:120-129
    νf = np.arange(sP.fft_range.start, sP.fft_range.stop)
    Σf = _Σ_at(Σ_loc, νf)
    iνf = 1j * (2 * νf + 1) * math.pi / cfg.β
    axes = tuple(range(cfg.D))
    G = 1.0 / (iνf + μ - kG.eps_full[..., None] - Σf)
    G_fft = np.asfortranarray(np.fft.fftn(G, axes=axes))
    G_rfft = np.asfortranarray(np.fft.fftn(G[(slice(None, None, -1),) * cfg.D], axes=axes))
    gLoc = np.asfortranarray(G[tuple(kG.wedge.T)])
    gImp = 1.0 / (iνf + μ - Σf)
    del G

    # local vertices
    Nν, Nω = 2 * cfg.n_iν, 2 * cfg.n_iω + 1
    rng = np.random.default_rng(cfg.seed)
    dist = np.abs(np.arange(Nν)[:, None] - np.arange(Nν)[None, :])
    scale = cfg.noise * cfg.U / (1.0 + dist)

    def noise():
        X = (rng.standard_normal((Nω, Nν, Nν)) + 1j * rng.standard_normal((Nω, Nν, Nν))) / math.sqrt(2)
        X = 0.5 * (X + np.transpose(X, (0, 2, 1)))
        return np.transpose(X * scale[None], (1, 2, 0))

    Γ_d = np.asfortranarray(cfg.U + noise())
    Γ_m = np.asfortranarray(-cfg.U + noise())

    # local chain (Nk = 1)
    f0 = sP.fft_range.start
    ωn = np.arange(-cfg.n_iω, cfg.n_iω + 1)
    sh = cfg.shift * np.trunc(ωn / 2).astype(int)
    νn = np.arange(Nν)[:, None] - cfg.n_iν - sh[None, :]       # (Nν, Nω) Matsubara index
    gg = gImp[νn - f0] * gImp[νn + ωn[None, :] - f0]          # g(ν) g(ν+ω)
    χ0_loc = -cfg.β * gg
    A = np.transpose(Γ_m, (2, 0, 1)).copy()
    idx = np.arange(Nν)
    A[:, idx, idx] += (1.0 / χ0_loc).T
    χDMFT_m = np.asfortranarray(np.transpose(np.linalg.inv(A), (1, 2, 0)))
    F_m = -(χDMFT_m + np.eye(Nν)[:, :, None] * (cfg.β * gg)[:, None, :]) / (gg[:, None, :] * gg[None, :, :])
    F_m = np.asfortranarray(F_m)
    χ_m_loc = np.real(χDMFT_m.sum(axis=(0, 1))) / cfg.β ** 2
    iω = 2j * math.pi * ωn / cfg.β
    with np.errstate(divide="ignore", invalid="ignore"):
        tail = np.where(ωn != 0, np.real(Ekin / iω ** 2), 0.0)
    χloc_m_sum = float((np.sum(χ_m_loc - tail) + Ekin * (-cfg.β ** 2 / 12)) / cfg.β)
    return LadderInputs(cfg, kG, mP, sP, Σ_loc, gImp, G_fft, G_rfft, Γ_m, Γ_d, χDMFT_m, F_m, χloc_m_sum, gLoc)
