"""Host mirror of the reference functions on the ladder hot path.

Same names, argument meaning and error behaviour as LadderDGA.jl; every
array computation runs in libldga_hip.so on the GPU (host-array entry
points).  Julia's in-place ``χ_λ!`` is spelled ``χ_λ_inplace``.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
import warnings

import numpy as np

from . import _lib
from ._lib import Freq, Model, BseReport, default_context, ptr
from .kgrid import KGrid
from .types import ModelParameters, SimulationParameters, χ0T, χT, γT

c128, f64 = np.complex128, np.float64


def _freq(sP: SimulationParameters) -> Freq:
    return Freq(sP.n_iω, sP.n_iν, sP.n_iν_shell, int(sP.shift), sP.fft_range.start, len(sP.fft_range))


def _tail_c2(χ, default: float) -> float:
    """c₂ of χ's 1/ω² tail (χ.tail_c[3] in Julia's numbering); 0 for an empty tail, as sum_kω takes it."""
    if χ is None:
        return default
    return float(χ.tail_c[2]) if len(χ.tail_c) == 3 else 0.0


def _model(mP: ModelParameters, χm=None, χd=None) -> Model:
    """The model parameters and the tail coefficients of the χ sums: χm.tail_c[3] and χd.tail_c[3]
    (Ekin_DMFT when a χ is not given, the value calc_χγ sets)."""
    return Model(mP.U, mP.μ, mP.β, mP.n, mP.Ekin_DMFT, _tail_c2(χm, mP.Ekin_DMFT), _tail_c2(χd, mP.Ekin_DMFT))


def _ctx(ctx, kG):
    ctx = ctx or default_context()
    ctx.set_kgrid(kG)
    return ctx


# ---------------------------------------------------------------------------
# G tables (src/helpers.jl:153-183)
# ---------------------------------------------------------------------------
def G_fft(G, kG: KGrid, sP: SimulationParameters, ctx=None):
    """(G_fft, G_rfft) = (fft(expandKArr(G)), fft(reverse(expandKArr(G)))) for every ν of
    sP.fft_range, from G on the wedge (Nq x length(fft_range)), as the reference's
    G_fft(G, kG, sP); the result feeds calc_bubble and calc_Σ."""
    ctx = _ctx(ctx, kG)
    nf = len(sP.fft_range)
    if G.shape != (kG.Nq, nf):
        raise ValueError(f"G must be Nq x length(fft_range) = {kG.Nq} x {nf}")
    shape = kG.gridshape() + (nf,)
    out = np.empty(shape, dtype=c128, order="F")
    out_r = np.empty(shape, dtype=c128, order="F")
    ctx.check(_lib.lib().ldga_g_tables(ctx.handle, C.byref(_freq(sP)), ptr(G, c128, "G"), out.ctypes.data,
                                       out_r.ctypes.data))
    return out, out_r


# ---------------------------------------------------------------------------
# bubble (src/BSETools.jl:157-172)
# ---------------------------------------------------------------------------
def calc_bubble(Gνω, Gνω_r, kG: KGrid, mP: ModelParameters, sP: SimulationParameters, local_tail=False, ctx=None):
    ctx = _ctx(ctx, kG)
    Nν, Nω = 2 * (sP.n_iν + sP.n_iν_shell), 2 * sP.n_iω + 1
    if Gνω.shape != kG.gridshape() + (len(sP.fft_range),) or Gνω_r.shape != Gνω.shape:
        raise ValueError("G tables must have shape gridshape(kG) x length(fft_range)")
    out = np.empty((kG.Nq, Nν, Nω), dtype=c128, order="F")
    ctx.check(_lib.lib().ldga_bubble(ctx.handle, C.byref(_freq(sP)), mP.β, ptr(Gνω, c128, "Gνω"),
                                     ptr(Gνω_r, c128, "Gνω_r"), out.ctypes.data))
    return χ0T(out, mP.β, ν_shell_size=sP.n_iν_shell)


# ---------------------------------------------------------------------------
# Bethe-Salpeter (src/BSETools.jl:192-248)
# ---------------------------------------------------------------------------
def calc_χγ(type_: str, Γr, χ0: χ0T, kG: KGrid, mP: ModelParameters, sP: SimulationParameters,
            precision: str = "fp64", ctx=None, report: dict | None = None):
    if type_ == "d" or type_ == ":d":
        s = -1
    elif type_ == "m" or type_ == ":m":
        s = 1
    else:
        raise ValueError("Unkown type")
    ctx = _ctx(ctx, kG)
    Nν, Nω = 2 * sP.n_iν, 2 * sP.n_iω + 1
    if Γr.shape != (Nν, Nν, Nω):
        raise ValueError(f"Γr must be {Nν}x{Nν}x{Nω}")
    χ = np.empty((kG.Nq, Nω), dtype=f64, order="F")
    γ = np.empty((kG.Nq, Nν, Nω), dtype=c128, order="F")
    rep = BseReport()
    prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
    rc = _lib.lib().ldga_chi_gamma(ctx.handle, s, prec, C.byref(_freq(sP)), mP.U, mP.β, ptr(Γr, c128, "Γr"),
                                   ptr(χ0.data, c128, "χ0.data"), χ.ctypes.data, γ.ctypes.data, C.byref(rep))
    ctx.check(rc, allow=(_lib.LDGA_ERR_ZERO_PIVOT,))
    if rc == _lib.LDGA_ERR_ZERO_PIVOT:
        # the reference ignores LAPACK's info (src/LapackWrapper.jl:31); keep going with NaN
        warnings.warn(f"{rep.n_zero_pivot} BSE matrices are exactly singular; their χ, γ are NaN")
    if precision == "mixed" and rep.n_fallback > rep.n_fp64_resolved:
        # refinement failures are re-solved in FP64 where an FP64 kernel exists (Nν <= 128)
        warnings.warn(f"{rep.n_fallback - rep.n_fp64_resolved} BSE matrices did not reach FP64 accuracy by "
                      f"iterative refinement (largest final correction {rep.max_rel_resid:.3g} of |y|)")
    if report is not None:
        report.update(n_matrices=rep.n_matrices, n_zero_pivot=rep.n_zero_pivot, ir_steps_max=rep.ir_steps_max,
                      n_fallback=rep.n_fallback, max_rel_resid=rep.max_rel_resid, ir_steps_mean=rep.ir_steps_mean,
                      n_fp64_resolved=rep.n_fp64_resolved)
    return χT(χ, mP.β, tail_c=[0.0, 0.0, mP.Ekin_DMFT]), γT(γ)


def calc_gen_χ(Γr, χ0: χ0T, kG: KGrid, ctx=None):
    """χ[ν,ν',q,ω] = (Γr[:,:,ω] + diag(1/χ₀[q, shell+1:end-shell, ω]))⁻¹ (src/BSETools.jl:259-275), in
    FP64 on the GPU (global-memory Gauss-Jordan; an analysis function, not the hot path).  Exactly
    singular matrices give NaN/Inf with a warning, as inv's output would."""
    ctx = _ctx(ctx, kG)
    Nν, _, Nω = Γr.shape
    if Γr.shape != (Nν, Nν, Nω) or Nν % 2 or Nω % 2 == 0:
        raise ValueError("Γr must be Nν x Nν x Nω with Nν even and Nω odd")
    sh = χ0.ν_shell_size
    if χ0.data.shape != (kG.Nq, Nν + 2 * sh, Nω):
        raise ValueError(f"χ₀ must be {kG.Nq} x {Nν + 2 * sh} x {Nω}")
    sP = SimulationParameters((Nω - 1) // 2, Nν // 2, sh, 0)
    out = np.empty((Nν, Nν, kG.Nq, Nω), dtype=c128, order="F")
    rc = _lib.lib().ldga_gen_chi(ctx.handle, C.byref(_freq(sP)), ptr(Γr, c128, "Γr"), ptr(χ0.data, c128, "χ0.data"),
                                 out.ctypes.data)
    ctx.check(rc, allow=(_lib.LDGA_ERR_ZERO_PIVOT,))
    if rc == _lib.LDGA_ERR_ZERO_PIVOT:
        warnings.warn("some BSE matrices are exactly singular; their inverses hold NaN/Inf")
    return out


# ---------------------------------------------------------------------------
# local full vertex (src/BSETools.jl:40-53)
# ---------------------------------------------------------------------------
def F_from_χ(χ, G, sP: SimulationParameters, β: float, diag_term: bool = True, ctx=None):
    """F[ν,ν',ω] = -(χ[ν,ν',ω] + δ_νν' β G(ν)G(ν+ω)) / (G(ν)G(ν+ω)G(ν')G(ν'+ω)) from the local χ
    (2n_iν x 2n_iν x 2n_iω+1) and the local G over sP.fft_range (e.g. gImp)."""
    ctx = ctx or default_context()
    Nν, Nω = 2 * sP.n_iν, 2 * sP.n_iω + 1
    if χ.shape != (Nν, Nν, Nω):
        raise ValueError(f"χ must be {Nν}x{Nν}x{Nω}")
    G = np.ascontiguousarray(G, dtype=c128)
    if G.shape != (len(sP.fft_range),):
        raise ValueError(f"G must hold the {len(sP.fft_range)} frequencies of fft_range")
    out = np.empty((Nν, Nν, Nω), dtype=c128, order="F")
    ctx.check(_lib.lib().ldga_F_from_chi(ctx.handle, C.byref(_freq(sP)), β, int(diag_term), ptr(χ, c128, "χ"),
                                         G.ctypes.data, out.ctypes.data))
    return out


# ---------------------------------------------------------------------------
# λ0, plain sums (src/BSETools.jl:118-144)
# ---------------------------------------------------------------------------
def calc_λ0(χ0: χ0T, Fr, mP: ModelParameters, sP: SimulationParameters, kG: KGrid, ctx=None):
    ctx = _ctx(ctx, kG)
    warnings.warn("Using plain summation for λ0, check Σ_ladder tails!")
    Nν, Nω = 2 * sP.n_iν, 2 * sP.n_iω + 1
    out = np.empty((kG.Nq, Nν, Nω), dtype=c128, order="F")
    ctx.check(_lib.lib().ldga_lambda0(ctx.handle, C.byref(_freq(sP)), mP.β, ptr(Fr, c128, "Fr"),
                                      ptr(χ0.data, c128, "χ0.data"), out.ctypes.data))
    return out


# ---------------------------------------------------------------------------
# χ_λ and the bosonic sums (LambdaCorrection/helpers.jl:18-70, DataTypes.jl:215-249)
# ---------------------------------------------------------------------------
def χ_λ(χ, λ: float):
    """Elementwise χ/(λχ+1) on scalars/arrays, or a new λ-corrected χT."""
    if isinstance(χ, χT):
        out = χT(np.asfortranarray(χ.data / (λ * χ.data + 1)) if λ != 0 else χ.data.copy(order="F"), χ.β,
                 tail_c=list(χ.tail_c), usable_ω=χ.usable_ω)
        out.λ = χ.λ + λ
        return out
    return χ / (λ * χ + 1)


def dχ_λ(χ, λ: float):
    return -χ_λ(χ, λ) ** 2


def χ_λ_inplace(χ: χT, λ: float):
    """χ_λ!(χ, λ): maps χ.data in place and records λ (undone by reset)."""
    if λ == 0.0:
        return
    if not math.isfinite(λ):
        print(f"WARNING. SKIPPING λ correction because {λ} is not finite!", file=sys.stderr)
        return
    χ.data[...] = χ.data / (λ * χ.data + 1)
    χ.λ = χ.λ + λ


def reset(χ: χT):
    """reset!(χ): undo χ_λ! (LambdaCorrection/helpers.jl:64-70)."""
    if χ.λ != 0:
        lam = -χ.λ
        χ.data[...] = χ.data / (lam * χ.data + 1)
        χ.λ = 0.0


def _usable(χ: χT):
    u = list(χ.usable_ω)
    lo, hi = u[0], u[-1]
    if u != list(range(lo, hi + 1)):
        raise ValueError("usable_ω must be a contiguous range")
    return lo, hi


def sum_kω(kG: KGrid, χ: χT, λ: float = math.nan, force_full_range=False, ctx=None, with_derivative=False):
    """Σ_q Σ_ω χ(q, ω) with the 1/ω² tail (src/DataTypes.jl:215-249)."""
    if χ.λ != 0 and not math.isnan(λ):
        raise RuntimeError("χ already λ-corrected, but external λ provided!")
    if len(χ.tail_c) == 3 and (χ.tail_c[0] != 0 or χ.tail_c[1] != 0):
        raise RuntimeError("sum_kω only implemented for ω^2 tail!")
    ctx = _ctx(ctx, kG)
    lo, hi = (0, χ.data.shape[1] - 1) if force_full_range else _usable(χ)
    if (lo, hi) != (0, χ.data.shape[1] - 1):
        # the reference builds the 1/ω² tail over all ω but sums a usable_ω slice of χ:
        # it only runs for the full range (src/DataTypes.jl:220-226)
        raise IndexError("sum_kω(χT): the ω tail spans all frequencies but usable_ω is a strict subset")
    c2 = χ.tail_c[2] if len(χ.tail_c) == 3 else 0.0
    lam = 0.0 if math.isnan(λ) else λ
    out = np.zeros(3)
    ctx.check(_lib.lib().ldga_sum_kw(ctx.handle, ptr(χ.data, f64, "χ.data"), χ.data.shape[1], lo, hi, c2, χ.β, lam,
                                     out.ctypes.data))
    return (out[0], out[2]) if with_derivative else float(out[0])


def _sum_w_call(ctx, χ: χT, lo: int, hi: int, lam: float, want_q: bool, want_k: bool):
    c = χ.tail_c
    if len(c) != 3:
        raise ValueError("sum_ω needs the three tail coefficients χ.tail_c = [c₀, c₁, c₂]")
    if not math.isclose(sum(c), c[2], rel_tol=1.4901161193847656e-08):
        raise ValueError("sum_ω only implemented for 1/ω^2, but other tail coefficients are non-zero!")
    Nq, Nω = χ.data.shape
    out_q = np.empty(Nq) if want_q else None
    out_k = np.zeros(1) if want_k else None
    ctx.check(_lib.lib().ldga_sum_w(ctx.handle, ptr(χ.data, f64, "χ.data"), Nq, Nω, lo, hi, float(c[2]), χ.β,
                                    lam, out_q.ctypes.data if want_q else None,
                                    out_k.ctypes.data if want_k else None))
    return out_q, out_k


def sum_ω(χ: χT, force_full_range=False, ctx=None) -> np.ndarray:
    """sum_ω(χ) (src/DataTypes.jl:277-308): per q, Σ_ω over the usable range with the c₂/ω² tail
    subtracted and its exact sum -c₂β²/12 added, divided by β.  As in the reference, element n of the
    usable slice is paired with the n-th frequency of the whole ω grid for the tail term
    (DataTypes.jl:293-301); on the full range that is its own frequency."""
    ctx = ctx or default_context()
    lo, hi = (0, χ.data.shape[1] - 1) if force_full_range else _usable(χ)
    return _sum_w_call(ctx, χ, lo, hi, 0.0, True, False)[0]


def sum_ωk(kG: KGrid, χ: χT, λ: float = math.nan, force_full_range=False, ctx=None) -> float:
    """sum_ωk (src/DataTypes.jl:257-264): kintegrate(sum_ω(χ_λ(χ, λ))), the ω sum per q first, on the
    usable range (a strict subset included, see sum_ω).  χ is not modified."""
    λ_check = not math.isnan(λ) and λ != 0
    if χ.λ != 0 and λ_check:
        raise RuntimeError("χ already λ-corrected, but external λ provided!")
    ctx = _ctx(ctx, kG)
    if χ.data.shape[0] != kG.Nq:
        raise ValueError(f"χ must have the Nq = {kG.Nq} rows of the k grid")
    lo, hi = (0, χ.data.shape[1] - 1) if force_full_range else _usable(χ)
    return float(_sum_w_call(ctx, χ, lo, hi, λ if λ_check else 0.0, False, True)[1][0])


def Σ_loc_correction(Σ_ladder, Σ_ladderLoc, Σ_loc):
    """Σ_ladder[q, ν] - Σ_ladderLoc[0, ν] + Σ_loc[ν] (src/ladderDGATools.jl:249-257), on the host as
    in the reference: O(Nq νmax) additions on arrays the caller holds."""
    Σ_ladder = np.asarray(Σ_ladder)
    Σ_ladderLoc = np.asarray(Σ_ladderLoc)
    νmax = Σ_ladder.shape[1]
    if Σ_ladderLoc.ndim != 2 or Σ_ladderLoc.shape[1] < νmax or len(Σ_loc) < νmax:
        raise ValueError("Σ_ladderLoc (1 x ≥νmax) and Σ_loc (≥νmax) must cover ν = 0..νmax-1")
    return Σ_ladder - Σ_ladderLoc[0, :νmax][None, :] + np.asarray(Σ_loc)[:νmax][None, :]


def update_tail(χ: χT, new_tail_c, ωnGrid=None) -> None:
    """update_tail!(χ, new_tail_c, ωnGrid) (src/DataTypes.jl:315-345): replaces the c₂/ω² tail of χ,
    χ[:, ω] += (c₂_new - c₂_old)/(iω_n)² for ω ≠ 0, with χ's λ taken off and put back around it.  On
    the host, as in the reference: O(Nq Nω) additions on the caller's array."""
    new_tail_c = [float(c) for c in new_tail_c]
    Nω = χ.data.shape[1]
    if ωnGrid is None:
        ωnGrid = 2j * np.pi * np.asarray(χ.indices_ω) / χ.β
    ωnGrid = np.asarray(ωnGrid)
    if len(new_tail_c) != len(χ.tail_c):
        raise ValueError(f"length of new tail coefficient array ({len(new_tail_c)}) must be the same as the old "
                         f"one ({len(χ.tail_c)})!")
    if len(ωnGrid) != Nω:
        raise ValueError(f"length of ωnGrid ({len(ωnGrid)}) must be the same as for χ ({Nω})!")
    if not all(math.isfinite(c) for c in new_tail_c):
        raise ValueError("One or more tail coefficients are not finite!")
    if sum(new_tail_c) != new_tail_c[2]:
        raise ValueError("update tail only implemented for c_2 coefficient!")
    λ_bak = χ.λ
    if λ_bak != 0:
        reset(χ)
    if new_tail_c[2] != χ.tail_c[2]:
        ω = np.abs(ωnGrid) >= 1e-10
        χ.data[:, ω] += np.real((new_tail_c[2] - χ.tail_c[2]) / ωnGrid[ω] ** 2)[None, :]
        χ.tail_c[2] = new_tail_c[2]
    if λ_bak != 0:
        χ_λ_inplace(χ, λ_bak)


def get_λ_min(χr: np.ndarray) -> float:
    """-min_q 1/χ(q, ω₀) (LambdaCorrection/helpers.jl:217-220)."""
    nh = math.ceil(χr.shape[1] / 2) - 1
    return float(-np.min(1.0 / χr[:, nh]))


# ---------------------------------------------------------------------------
# EoM (src/ladderDGATools_singleCore.jl:108-189)
# ---------------------------------------------------------------------------
def _sigma_call(ctx, sP, mP, νmax, λm, λd, tc, kind, χ_m_sum, χm, γm, χd, γd, λ0, Gνω_r, Nq):
    out = np.empty((Nq, νmax), dtype=c128, order="F")
    ctx.check(_lib.lib().ldga_sigma(ctx.handle, C.byref(_freq(sP)), C.byref(_model(mP, χm, χd)), νmax, λm, λd, int(tc),
                                    kind, χ_m_sum, ptr(χm.data, f64, "χm"), ptr(γm.data, c128, "γm"),
                                    ptr(χd.data, f64, "χd"), ptr(γd.data, c128, "γd"), ptr(λ0, c128, "λ0"),
                                    ptr(Gνω_r, c128, "Gνω_r"), out.ctypes.data))
    return out


def calc_Σ(χm: χT, γm: γT, χd: χT, γd: γT, χ_m_sum: float, λ0, Gνω_r, kG: KGrid, mP: ModelParameters,
           sP: SimulationParameters, νmax: int | None = None, λm: float = 0.0, λd: float = 0.0, tc: bool = True,
           ctx=None):
    """Σ_ladder[q, ν] for ν = 0..νmax-1 (the reference's OffsetMatrix with origin 0)."""
    if χm.λ != 0 and λm != 0:
        raise RuntimeError(f"Stopping self energy calculation: λm = {λm} AND χm.λ = {χm.λ}")
    if χd.λ != 0 and λd != 0:
        raise RuntimeError(f"Stopping self energy calculation: λd = {λd} AND χd.λ = {χd.λ}")
    νmax = sP.n_iν if νmax is None else νmax
    if χm.λ != 0 or χd.λ != 0:
        # χ was mapped in place by χ_λ_inplace: the data already carry λ
        pass
    ctx = _ctx(ctx, kG)
    return _sigma_call(ctx, sP, mP, νmax, λm, λd, tc, 0, χ_m_sum, χm, γm, χd, γd, λ0, Gνω_r, kG.Nq)


def calc_Σ_ω(eomf: str, χm: χT, γm: γT, χd: χT, γd: γT, Gνω_r, λ0, U: float, kG: KGrid,
             sP: SimulationParameters, νmax: int | None = None, ctx=None):
    """Σ_ladder[q, ν] = Σ_ω conv_q(K, G_rfft(ν + ω)), K = eomf(U, γm, γd, χm, χd, λ₀), without 1/β, tail or
    Hartree term (calc_Σ_ω!, src/ladderDGATools_singleCore.jl:49-69).  eomf names one of the reference's
    real-χ EoM functions (src/ladderDGATools.jl:260-271): "full" (eom), "χ_m", "γ_m", "χ_d", "γ_d",
    "rest_01", "rest", "sp_01", "sp_02", "sp", "ch" (eom_χ_m, ..., eom_ch).  χ is used as given (apply λ to
    it beforehand, as the reference's caller does)."""
    if eomf not in _lib.EOM_KINDS:
        raise ValueError(f"unknown EoM function {eomf!r}; one of {sorted(_lib.EOM_KINDS)}")
    ctx = _ctx(ctx, kG)
    νmax = sP.n_iν if νmax is None else νmax
    # the kernel's scale is 1/β and it adds n U/2 only for "full" and "rest": β = 1 and n = 0 give the bare sum
    bare = ModelParameters(U=U, μ=0.0, β=1.0, n=0.0)
    return _sigma_call(ctx, sP, bare, νmax, 0.0, 0.0, False, _lib.EOM_KINDS[eomf], 0.0, χm, γm, χd, γd, λ0, Gνω_r,
                       kG.Nq)


def calc_Σ_parts(χm: χT, γm: γT, χd: χT, γd: γT, χ_m_sum: float, λ0, Gνω_r, kG: KGrid, mP: ModelParameters,
                 sP: SimulationParameters, tc: bool = True, λm: float = 0.0, λd: float = 0.0, ctx=None):
    """The 7 contributions χ_m, γ_m, χ_d, γ_d, U, rest(+Hartree), tail (src/ladderDGATools_singleCore.jl:153-189)."""
    ctx = _ctx(ctx, kG)
    νmax = sP.n_iν
    out = np.empty((kG.Nq, νmax, 7), dtype=c128, order="F")
    for i, kind in enumerate(["χ_m", "γ_m", "χ_d", "γ_d", "rest_01", "rest"]):
        out[:, :, i] = _sigma_call(ctx, sP, mP, νmax, λm, λd, False, _lib.EOM_KINDS[kind], χ_m_sum, χm, γm, χd,
                                   γd, λ0, Gνω_r, kG.Nq)
    if tc:
        s = sum_kω(kG, χm, λ=λm if λm != 0 else math.nan, ctx=ctx)
        tail = -mP.U * (s - χ_m_sum)
    else:
        tail = 0.0
    iν = 1j * (2 * np.arange(νmax) + 1) * math.pi / mP.β
    out[:, :, 6] = (tail / iν)[None, :]
    return out
