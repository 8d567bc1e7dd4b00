"""λ corrections (LambdaCorrection/ of the reference) over libldga_hip.

Host logic only: the root finders (newton_right, LambdaCorrection/helpers.jl:
133-208) and the bookkeeping of λm_correction / λdm_correction
(conditions_new.jl:144-324).  Every evaluation of a sum or a residual runs on
the GPU; for λdm the ladder arrays are uploaded once and stay resident.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import LambdaDmOut, default_context
from .api import _freq, _model, get_λ_min, sum_kω, sum_ωk
from .kgrid import KGrid
from .types import ModelParameters, SimulationParameters, χT, γT


@dataclass
class λ_result:
    """conditions_new.jl:13-46."""
    λm: float
    λd: float
    type: str
    converged: bool
    sc_converged: bool = True
    EKin: float = math.nan
    EPot_p1: float = math.nan
    EPot_p2: float = math.nan
    PP_p1: float = math.nan
    PP_p2: float = math.nan
    trace: list = field(default_factory=list)
    G_ladder: np.ndarray = None
    Σ_ladder: np.ndarray = None
    μ: float = math.nan
    n: float = math.nan
    newton_steps: int = 0
    residual_evals: int = 0


# ---------------------------------------------------------------------------
# root finders (LambdaCorrection/helpers.jl:133-208)
# ---------------------------------------------------------------------------
def newton_right_1d(f, df, start: float, min_: float, nsteps: int = 100, atol: float = 1e-13) -> float:
    δ = 1e-4
    x0 = start
    xi = x0
    i = 1
    while True:
        fi = f(xi)
        dfii = 1.0 / df(xi)
        xi = x0 - dfii * fi
        if abs(fi) < atol:
            break
        x0 = (min_ + δ + abs(min_ - x0) / 2) if xi < min_ else xi
        if i >= nsteps:
            break
        i += 1
    return xi


def finite_difference_jacobian(f, x, fx):
    """Forward differences, relative step sqrt(eps) (FiniteDiff.jl's default)."""
    x = np.asarray(x, dtype=float)
    J = np.empty((len(fx), len(x)))
    rel = math.sqrt(np.finfo(float).eps)
    for j in range(len(x)):
        h = rel * max(abs(x[j]), 1.0)
        xp = x.copy()
        xp[j] += h
        J[:, j] = (np.asarray(f(xp)) - fx) / h
    return J


def newton_right(f, start, min_, nsteps: int = 500, atol: float = 1e-8, max_reset: int = 5,
                 reset_backoff: float = 1.0, fj=None, verbose: bool = False):
    """2D Newton that stays right of `min_` (helpers.jl:165-208).

    `fj(x) -> (f(x), J(x))` supplies an analytic Jacobian; without it J is the
    forward finite-difference Jacobian of the reference.  Returns (x, steps).
    """
    start = np.asarray(start, dtype=float)
    min_ = np.asarray(min_, dtype=float)
    xi = start.copy()
    i = 1
    reset_i = 1
    bisection_fail = False
    while True:
        if fj is None:
            fi = np.asarray(f(xi), dtype=float)
            J = finite_difference_jacobian(f, xi, fi)
        else:
            fi, J = fj(xi)
            fi = np.asarray(fi, dtype=float)
        try:
            dfii = np.linalg.inv(J)
        except np.linalg.LinAlgError:
            dfii = np.full((len(xi), len(xi)), np.nan)
        xi_last = xi.copy()
        xi = xi - dfii @ fi
        if verbose:
            print(f"i={i} : xi = {xi}, fi = {fi}. {np.linalg.norm(fi)} ?<? {atol}", file=sys.stderr)
        for l in range(len(xi)):
            if xi[l] < min_[l]:
                xi[l] = min_[l] + abs(xi_last[l] - min_[l]) / 2
                if abs(xi_last[l] - min_[l]) / 2 < 1e-8:
                    print("WARNING: Bisection step is smaller than 1e-8. This indicates faulty data or no solution.", file=sys.stderr)
                    bisection_fail = True
        if bisection_fail or not np.all(np.isfinite(fi)):
            bisection_fail = False
            if reset_i >= max_reset:
                raise RuntimeError(f"λ root-finding returned NaN after {i} iterations and {max_reset} resets. "
                                   f"min values are {min_}")
            reset_i += 1
            xi = start + reset_i * reset_backoff * np.abs(min_)
        if abs(np.linalg.norm(fi)) < atol:
            break
        if i >= nsteps:
            break
        i += 1
    return xi, i


# ---------------------------------------------------------------------------
# λm (conditions_new.jl:144-172)
# ---------------------------------------------------------------------------
def λm_correction(χm: χT, rhs: float, kG: KGrid, mP: ModelParameters, sP: SimulationParameters,
                  validate_threshold: float = 1e-8, ctx=None):
    """λm from Σ_{q,ω} χm^λ = rhs by 1D Newton with the reference's derivative."""
    ctx = ctx or default_context()
    ctx.set_kgrid(kG)
    λm_min = get_λ_min(χm.data)
    u = list(χm.usable_ω)
    lo, hi = u[0], u[-1]
    L = _lib.lib()
    Nω = χm.data.shape[1]
    data = np.asfortranarray(χm.data, dtype=np.float64)
    out = np.zeros(3)

    def sums(l):
        ctx.check(L.ldga_sum_kw(ctx.handle, data.ctypes.data, Nω, lo, hi, χm.tail_c[2], χm.β, l, out.ctypes.data))
        return out.copy()

    f = lambda l: sums(l)[0] - rhs
    df = lambda l: sums(l)[1]         # the reference's df_c1 (transform = dχ_λ, tails included)
    λm = newton_right_1d(f, df, 0.0, λm_min)
    check = sums(λm)[0]
    return λm, bool(abs(rhs - check) <= validate_threshold)


def λm_rhs(imp_density: float, χm: χT, χd: χT, kG: KGrid, mP: ModelParameters, sP: SimulationParameters,
           λd: float = math.nan, λ_rhs: str = "native", ctx=None) -> float:
    """Right-hand side of the λm condition (LambdaCorrection/helpers.jl:234-255).

    "fixed": n (1 - n/2) - Σ χd; "native" without a χ helper (plain mode):
    2 imp_density - Σ χd, where imp_density is the impurity's equal-time
    same-spin density correlator.  Σ χd is taken at λd when it is given.
    """
    if χd.λ != 0 and not math.isnan(λd):
        raise RuntimeError(f"λd = {λd} given while χd.λ = {χd.λ} is set")
    if λ_rhs not in ("native", "fixed"):
        raise ValueError(f"λ_rhs {λ_rhs!r}: supported are 'native' and 'fixed'")
    if λ_rhs == "native" and not math.isfinite(imp_density):
        raise ValueError("imp_density argument is not finite! Cannot use DMFT error compensation method")
    χd_sum = sum_kω(kG, χd, λ=λd, ctx=ctx)
    if λ_rhs == "fixed":
        return mP.n * (1 - mP.n / 2) - χd_sum
    return 2 * imp_density - χd_sum


# ---------------------------------------------------------------------------
# λdm (conditions_new.jl:219-324)
# ---------------------------------------------------------------------------
def gen_νω_indices(χm: χT, χd: χT, sP: SimulationParameters):
    """LambdaCorrection/helpers.jl:280-289 with usable_ωindices (GFFit.jl:81-84)."""
    if sP.dbg_full_eom_omega:
        ω = list(range(χm.data.shape[1]))
    else:
        ω = sorted(set(χm.usable_ω) & set(χd.usable_ω))
    return ω, min(sP.n_iν, (3 * len(ω)) // 8)


class LambdaDmProblem:
    """The λdm residual with the ladder arrays resident on the device."""

    def __init__(self, χm: χT, γm: γT, χd: χT, γd: γT, Σ_loc, gLoc_rfft, χloc_m_sum, λ0, kG: KGrid,
                 mP: ModelParameters, sP: SimulationParameters, νmax: int, tc: bool = True, μ: float | None = None,
                 ctx=None):
        self.ctx = ctx or default_context()
        self.ctx.set_kgrid(kG)
        self.kG, self.mP, self.sP, self.νmax, self.tc = kG, mP, sP, νmax, tc
        self.χloc_m_sum = float(χloc_m_sum)
        up = self.ctx.to_device
        self.d = dict(χm=up(np.asarray(χm.data, np.float64)), χd=up(np.asarray(χd.data, np.float64)),
                      γm=up(np.asarray(γm.data, np.complex128)), γd=up(np.asarray(γd.data, np.complex128)),
                      λ0=up(np.asarray(λ0, np.complex128)), G=up(np.asarray(gLoc_rfft, np.complex128)),
                      Σloc=up(np.ascontiguousarray(Σ_loc[:sP.fft_range.stop], np.complex128)))
        self.fr, self.mo = _freq(sP), _model(mP, χm, χd)
        self.μ = mP.μ if μ is None else float(μ)      # start of the first μ fit
        self.evals = 0

    def eval(self, λ, want_jac=False, fit_μ=True, arrays=False):
        d = self.d
        out = LambdaDmOut()
        Σ = G = None
        if arrays:
            Σ = self.ctx.empty((self.kG.Nq, self.νmax), np.complex128)
            G = self.ctx.empty((self.kG.Nq, len(self.sP.fft_range)), np.complex128)
        rc = _lib.lib().ldga_lambdadm_residual_dev(
            self.ctx.handle, C.byref(self.fr), C.byref(self.mo), self.νmax, float(λ[0]), float(λ[1]), int(self.tc),
            self.χloc_m_sum, int(fit_μ), self.μ, d["χm"].ptr, d["γm"].ptr, d["χd"].ptr, d["γd"].ptr, d["λ0"].ptr,
            d["G"].ptr, d["Σloc"].ptr, int(want_jac), C.byref(out), Σ.ptr if arrays else None,
            G.ptr if arrays else None, None)
        self.ctx.check(rc)
        self.evals += 1
        if fit_μ and math.isfinite(out.mu):
            self.μ = out.mu          # the reference's closure keeps the last μ as the next start
        res = {k: getattr(out, k) for k in ("n", "mu", "Ekin", "Epot1", "Epot2", "lhs_c1", "res_pp", "res_epot",
                                            "mu_iters", "mu_converged", "n_full")}
        res["residual"] = np.array([out.res_pp, out.res_epot])
        res["jac"] = np.array(out.jac[:]).reshape(2, 2) if want_jac else None
        if arrays:
            res["Σ_ladder"], res["G_ladder"] = Σ.to_host(), G.to_host()
            Σ.free()
            G.free()
        return res

    def free(self):
        for a in self.d.values():
            a.free()


def λdm_correction(χm: χT, γm: γT, χd: χT, γd: γT, Σ_loc, gLoc_rfft, χloc_m_sum, λ0, kG: KGrid,
                   mP: ModelParameters, sP: SimulationParameters, νmax: int = -1, λ_min_δ: float = 1e-4,
                   λ_val_only: bool = False, fit_μ: bool = True, λinit=(0.0, 0.0), validate_threshold: float = 1e-8,
                   tc: bool = True, jacobian: str = "fd", verbose: bool = False, μ: float | None = None, ctx=None):
    """λm, λd such that the Pauli sum rule and E_pot consistency hold.

    μ (default mP.μ) starts the first μ fit; later fits start from the last one.

    jacobian = "fd" reproduces the reference (forward finite differences of the
    residual, helpers.jl:177-180); "adjoint" uses the analytic Jacobian.
    """
    if χm.λ != 0 or χd.λ != 0:
        raise RuntimeError("λ parameter already set. Aborting λdm calculation")
    _, νmax_d = gen_νω_indices(χm, χd, sP)
    νmax = νmax_d if νmax < 1 else min(νmax_d, νmax)
    prob = LambdaDmProblem(χm, γm, χd, γd, Σ_loc, gLoc_rfft, χloc_m_sum, λ0, kG, mP, sP, νmax, tc=tc, μ=μ, ctx=ctx)
    rhs_c1 = mP.n / 2 * (1 - mP.n / 2)
    f = lambda λ: prob.eval(λ, fit_μ=fit_μ)["residual"]

    def fj(λ):
        r = prob.eval(λ, want_jac=True, fit_μ=fit_μ)
        return r["residual"], r["jac"]

    λm_min, λd_min = get_λ_min(χm.data), get_λ_min(χd.data)
    min_λ = np.array([λm_min + λ_min_δ * abs(λm_min), λd_min + λ_min_δ * abs(λd_min)])
    try:
        root, steps = newton_right(f, list(λinit), min_λ, verbose=verbose,
                                   fj=fj if jacobian == "adjoint" else None)
    except RuntimeError as e:
        print(f"Error: {e}", file=sys.stderr)
        root, steps = np.array([math.nan, math.nan]), 0
    if np.any(np.isnan(root)):
        print("WARNING: No λ root was found!", file=sys.stderr)
    elif np.any(root < min_λ):
        print(f"WARNING: λ = {root} outside region ({min_λ})!", file=sys.stderr)
    if np.all(np.isfinite(root)):
        r = prob.eval(root, fit_μ=fit_μ, arrays=True)
        converged = abs(rhs_c1 - r["lhs_c1"]) <= validate_threshold and abs(r["Epot1"] - r["Epot2"]) <= validate_threshold
        prob.free()
        if λ_val_only:
            return root[0], root[1], converged
        Gl = r["G_ladder"]
        # the reference's constructor stores its 4th/5th arguments swapped (conditions_new.jl:42):
        # `converged` is the sum-rule validation, `sc_converged` is true
        return λ_result(root[0], root[1], "dm", converged=converged, sc_converged=True, EKin=r["Ekin"],
                        EPot_p1=r["Epot1"], EPot_p2=r["Epot2"], PP_p1=rhs_c1, PP_p2=r["lhs_c1"], trace=[],
                        G_ladder=Gl, Σ_ladder=r["Σ_ladder"], μ=r["mu"], n=r["n"], newton_steps=steps,
                        residual_evals=prob.evals)
    prob.free()
    if λ_val_only:
        return root[0], root[1], False
    return λ_result(root[0], root[1], "dm", converged=False, sc_converged=True, newton_steps=steps,
                    residual_evals=prob.evals)


# ---------------------------------------------------------------------------
# self-consistency loop (LambdaCorrection/conditions_new.jl:342-401, run_sc! :405-560)
# ---------------------------------------------------------------------------
def run_sc(χm: χT, γm: γT, χd: χT, γd: γT, λ0, μ: float, Σ_loc, gLoc_rfft, χloc_m_sum, kG: KGrid,
           mP: ModelParameters, sP: SimulationParameters, type: str = "fix", νmax: int = -1, maxit: int = 100,
           mixing: float = 0.2, conv_abs: float = 1e-8, fit_μ: bool = True, tc: bool = True, λm: float = 0.0,
           λd: float = 0.0, λ_min_δ: float = 1e-4, update_χ_tail: bool = False, trace: bool = False,
           imp_density: float = math.nan, ctx=None) -> λ_result:
    """Self-consistent lDΓA with the loop on the GPU.

    Every iteration runs, with the arrays resident on the device: the EoM with
    the current G_rfft (calc_Σ!, :479), mixing with the previous Σ (:484),
    G_from_Σladder! with the μ fit and calc_E (:485-490), and the rebuild of
    G_rfft from G_ladder (G_rfft!, :491).  type "fix" keeps (λm, λd); "dm"
    starts from them and re-solves λdm_correction with the current G_rfft in
    every later iteration (:437-449); "pre_dm" first solves λdm_correction
    (:370-372).  "pre_m" first solves λm_correction with λm_rhs (plain mode:
    2 imp_density - Σ χd) and sets λd = 0 (:365-369); the reference's later
    iterations then keep λ fixed for "m" and "pre_m" (its `type == [:pre_m,
    :m]` test at :456 never holds), so "m" runs as "fix", except that the
    result of "fix" carries λm = λd = 0 as the reference's does (the λ used
    are in the trace).  The loop stops when both potential energies change by less
    than conv_abs, or after maxit iterations.  χm and χd are left unchanged.
    """
    import warnings
    if type not in ("fix", "m", "pre_m", "dm", "pre_dm"):
        raise ValueError(f"run_sc type {type!r}: supported are 'fix', 'm', 'pre_m', 'dm' and 'pre_dm'")
    if update_χ_tail:
        raise NotImplementedError("update_χ_tail (the χ tail self-consistency) is not supported")
    if not fit_μ:
        warnings.warn("Not fitting μ can lead to unphysical results!")
    if χm.λ != 0 or χd.λ != 0:
        raise RuntimeError("λ parameter already set; run_sc applies λm, λd itself")
    _, νmax_d = gen_νω_indices(χm, χd, sP)
    νmax = νmax_d if νmax < 1 else min(νmax_d, νmax)
    validation = True
    if type == "pre_m":
        rhs = λm_rhs(imp_density, χm, χd, kG, mP, sP, ctx=ctx)
        λm, validation = λm_correction(χm, rhs, kG, mP, sP, ctx=ctx)
        λd = 0.0
    if type == "pre_dm":
        λm, λd, validation = λdm_correction(χm, γm, χd, γd, Σ_loc, gLoc_rfft, χloc_m_sum, λ0, kG, mP, sP,
                                            νmax=νmax, λ_min_δ=λ_min_δ, λ_val_only=True, fit_μ=fit_μ,
                                            validate_threshold=conv_abs, tc=tc, μ=μ, ctx=ctx)
    if not validation or not (math.isfinite(λm) and math.isfinite(λd)):
        nan = math.nan      # conditions_new.jl:391-393: everything NaN when the pre-solve failed
        return λ_result(nan, nan, type, converged=False, sc_converged=False, EKin=nan, EPot_p1=nan, EPot_p2=nan,
                        PP_p1=nan, PP_p2=nan, μ=nan, n=nan)

    ctx = ctx or default_context()
    ctx.set_kgrid(kG)
    L = _lib.lib()
    up = ctx.to_device
    nf = len(sP.fft_range)
    d = dict(χm=up(np.asarray(χm.data, np.float64)), χd=up(np.asarray(χd.data, np.float64)),
             γm=up(np.asarray(γm.data, np.complex128)), γd=up(np.asarray(γd.data, np.complex128)),
             λ0=up(np.asarray(λ0, np.complex128)), Gr=up(np.asarray(gLoc_rfft, np.complex128)),
             Σloc=up(np.ascontiguousarray(Σ_loc[:sP.fft_range.stop], np.complex128)))
    d["Gf"] = ctx.empty(d["Gr"].shape, np.complex128)                  # G_fft of the rebuild (unused)
    d["G"] = ctx.empty((kG.Nq, nf), np.complex128)
    cur, nxt = ctx.empty((kG.Nq, νmax), np.complex128), ctx.empty((kG.Nq, νmax), np.complex128)
    fr, mo = _freq(sP), _model(mP, χm, χd)
    out = LambdaDmOut()
    # the sums at (λm, λd) before the loop (run_sc! :417-421)
    χm_sum = sum_kω(kG, χm, λ=λm, ctx=ctx)
    χd_sum = sum_kω(kG, χd, λ=λd, ctx=ctx)
    lhs_c1 = (χd_sum + χm_sum) / 2
    rhs_c1 = mP.n / 2 * (1 - mP.n / 2)
    E_pot_2 = (mP.U / 2) * (χd_sum - χm_sum) + mP.U * (mP.n / 2 * mP.n / 2)
    E_pot_1 = E_kin = math.inf
    rows = []
    it, converged, n = 1, False, math.nan
    try:
        while True:
            if type in ("dm", "pre_dm") and it > 1:
                r = λdm_correction(χm, γm, χd, γd, Σ_loc, d["Gr"].to_host(), χloc_m_sum, λ0, kG, mP, sP,
                                   νmax=νmax, λ_min_δ=λ_min_δ, fit_μ=fit_μ, λinit=(λm, λd), tc=tc, μ=μ, ctx=ctx)
                λm, λd, validation = r.λm, r.λd, r.converged
                if math.isfinite(r.μ):
                    μ = r.μ
            if not (math.isfinite(λm) and math.isfinite(λd)) or not validation:
                print(f"ERROR: λm = {λm} or λd = {λd} not finite, OR internal λ correction did not find root!", file=sys.stderr)
                break
            ctx.check(L.ldga_sigma_dev(ctx.handle, C.byref(fr), C.byref(mo), νmax, λm, λd, int(tc), 0,
                                       float(χloc_m_sum), d["χm"].ptr, d["γm"].ptr, d["χd"].ptr, d["γd"].ptr,
                                       d["λ0"].ptr, d["Gr"].ptr, nxt.ptr, None))
            prev = cur.ptr if (mixing != 0 and it > 1) else None
            ctx.check(L.ldga_g_from_sigma_ladder_dev(ctx.handle, C.byref(fr), C.byref(mo), νmax, λm, λd,
                                                     int(fit_μ), μ, d["χm"].ptr, d["χd"].ptr, nxt.ptr, prev,
                                                     float(mixing), d["Σloc"].ptr, C.byref(out), d["G"].ptr, None))
            cur, nxt = nxt, cur
            if math.isnan(out.mu):
                break
            μ, n = out.mu, out.n
            E_pot_1_old, E_pot_2_old = E_pot_1, E_pot_2
            E_kin, E_pot_1 = out.Ekin, out.Epot1
            ctx.check(L.ldga_g_tables_dev(ctx.handle, C.byref(fr), d["G"].ptr, d["Gf"].ptr, d["Gr"].ptr, None))
            if type in ("m", "pre_m", "dm", "pre_dm"):
                lhs_c1, E_pot_2 = out.lhs_c1, out.Epot2
            if trace:
                # the reference's trace row (conditions_new.jl:507-515): χ sums in both orders, the
                # filling of the whole G_ladder and the checksums |Σ Σ_ladder|, |Σ G_ladder|
                cs_m, cs_d = sum_kω(kG, χm, λ=λm, ctx=ctx), sum_kω(kG, χd, λ=λd, ctx=ctx)
                cs_m2, cs_d2 = sum_ωk(kG, χm, λ=λm, ctx=ctx), sum_ωk(kG, χd, λ=λd, ctx=ctx)
                rows.append(dict(it=it, λm=λm, λd=λd, μ=μ, n=out.n_full, EKin=E_kin, EPot=E_pot_1,
                                 lhs_c1=(cs_d + cs_m) / 2, EPot_c2=E_pot_2, cs_m=cs_m, cs_m2=cs_m2, cs_d=cs_d,
                                 cs_d2=cs_d2, cs_Σ=float(abs(cur.to_host().sum())),
                                 cs_G=float(abs(d["G"].to_host().sum()))))
            if abs(E_pot_1 - E_pot_1_old) < conv_abs and abs(E_pot_2 - E_pot_2_old) < conv_abs:
                converged = True
                break
            if it >= maxit:
                break
            it += 1
        Σ_ladder, G_ladder = cur.to_host(), d["G"].to_host()
    finally:
        for a in list(d.values()) + [cur, nxt]:
            a.free()
    ok = (converged and math.isfinite(lhs_c1) and math.isfinite(E_pot_2) and abs(rhs_c1 - lhs_c1) <= conv_abs
          and abs(E_pot_1 - E_pot_2) <= conv_abs)
    # the reference's constructor stores its 4th/5th arguments swapped (conditions_new.jl:42):
    # `converged` is the sum-rule validation, `sc_converged` the loop's convergence
    if type == "fix":
        # run_sc! returns its loop-local λ, which stay 0 for :fix (the χ carry the λ there,
        # conditions_new.jl:424-425, 461); the trace rows log the λ used
        λm, λd = 0.0, 0.0
    return λ_result(λm, λd, type, converged=ok, sc_converged=converged, EKin=E_kin, EPot_p1=E_pot_1,
                    EPot_p2=E_pot_2, PP_p1=rhs_c1, PP_p2=lhs_c1, trace=rows, G_ladder=G_ladder, Σ_ladder=Σ_ladder,
                    μ=μ, n=n)


# ---------------------------------------------------------------------------
# the helper and the λ_correction entry point (conditions_new.jl:60-85, 174-190)
# ---------------------------------------------------------------------------
@dataclass
class lDΓAHelper:
    """The fields of the reference's lDΓAHelper (src/runHelpers.jl:28-49) that the λ corrections read:
    grid, parameters, local self-energy (ν = 0 .. last(fft_range)), the G_rfft table, the local
    Σ_ω χm, and the impurity's equal-time same-spin density correlator for the native λm rhs."""
    kG: KGrid
    mP: ModelParameters
    sP: SimulationParameters
    Σ_loc: np.ndarray
    gLoc_rfft: np.ndarray
    χloc_m_sum: float
    imp_density: float = math.nan


def λm_correction_full(χm: χT, γm: γT, χd: χT, γd: γT, λ0, h: lDΓAHelper, νmax: int = -1,
                       λ_min_δ: float = 1e-4, λ_val_only: bool = False, verbose: bool = False, fit_μ: bool = True,
                       validate_threshold: float = 1e-8, ctx=None) -> λ_result:
    """λm from the native rhs, then Σ_ladder at (λm, χd.λ) over ν < νmax (default n_iν), its μ fit,
    energies and sums (conditions_new.jl:174-190).  As the reference stores it, the result's
    `converged` is true and `sc_converged` carries the λm validation."""
    νmax = γm.data.shape[1] // 2 if νmax < 0 else νmax
    rhs = λm_rhs(h.imp_density, χm, χd, h.kG, h.mP, h.sP, λ_rhs="native", ctx=ctx)
    λm, validation = λm_correction(χm, rhs, h.kG, h.mP, h.sP, validate_threshold=validate_threshold, ctx=ctx)
    prob = LambdaDmProblem(χm, γm, χd, γd, h.Σ_loc, h.gLoc_rfft, h.χloc_m_sum, λ0, h.kG, h.mP, h.sP, νmax, ctx=ctx)
    try:
        r = prob.eval((λm, χd.λ), fit_μ=fit_μ, arrays=True)
    finally:
        prob.free()
    rhs_c1 = h.mP.n / 2 * (1 - h.mP.n / 2)
    return λ_result(λm, χd.λ, "m", converged=True, sc_converged=validation, EKin=r["Ekin"], EPot_p1=r["Epot1"],
                    EPot_p2=r["Epot2"], PP_p1=rhs_c1, PP_p2=r["lhs_c1"], G_ladder=r["G_ladder"],
                    Σ_ladder=r["Σ_ladder"], μ=r["mu"], n=h.mP.n)


def λ_correction(type: str, χm: χT, γm: γT, χd: χT, γd: γT, λ0, h: lDΓAHelper, λm_rhs_type: str = "native",
                 νmax: int = -1, λ_min_δ: float = 1e-4, maxit: int = 100, mixing: float = 0.2,
                 conv_abs: float = 1e-8, trace: bool = False, par: bool = False, λ_val_only: bool = False,
                 verbose: bool = False, validate_threshold: float = 1e-8, tc: bool = True, ctx=None):
    """The reference's dispatcher (conditions_new.jl:60-85): "m" (λm_correction_full), "dm"
    (λdm_correction), "sc" (run_sc, type :fix at λ = 0) and "sc_m" (run_sc, type :m).  As there,
    λm_rhs_type is accepted and not used, and `par` selects nothing on one GPU."""
    t = type.lstrip(":")
    if t == "m":
        return λm_correction_full(χm, γm, χd, γd, λ0, h, νmax=νmax, λ_min_δ=λ_min_δ, verbose=verbose,
                                  validate_threshold=validate_threshold, ctx=ctx)
    if t == "dm":
        return λdm_correction(χm, γm, χd, γd, h.Σ_loc, h.gLoc_rfft, h.χloc_m_sum, λ0, h.kG, h.mP, h.sP,
                              νmax=νmax, λ_min_δ=λ_min_δ, validate_threshold=validate_threshold, verbose=verbose,
                              tc=tc, λ_val_only=λ_val_only, ctx=ctx)
    if t in ("sc", "sc_m"):
        return run_sc(χm, γm, χd, γd, λ0, h.mP.μ, h.Σ_loc, h.gLoc_rfft, h.χloc_m_sum, h.kG, h.mP, h.sP,
                      type="fix" if t == "sc" else "m", maxit=maxit, mixing=mixing, conv_abs=conv_abs, trace=trace,
                      imp_density=h.imp_density, ctx=ctx)
    raise ValueError(f"λ-correction type '{type}' not recognized!")


def G_from_Σladder(Σ_ladder, Σ_loc, kG: KGrid, mP: ModelParameters, sP: SimulationParameters, fix_n: bool = False,
                   μ: float | None = None, ctx=None):
    """(μ, G_ladder) from Σ_ladder (Nq x νmax, ν = 0..νmax-1) over sP.fft_range (GFTools.jl:154-186):
    Σ_ladder below νmax, Σ_loc (ν = 0..last(fft_range)) above, negative ν by conjugation; with fix_n,
    μ (starting from μ, default mP.μ) is refit so that filling_pos(G[:, 0:νmax-1]) = n.  μ is NaN
    when the fit fails, as in the reference."""
    ctx = ctx or default_context()
    ctx.set_kgrid(kG)
    Σ_ladder = np.asarray(Σ_ladder, np.complex128)
    if Σ_ladder.ndim != 2 or Σ_ladder.shape[0] != kG.Nq:
        raise ValueError(f"Σ_ladder must be Nq x νmax with Nq = {kG.Nq}")
    νmax = Σ_ladder.shape[1]
    μ0 = mP.μ if μ is None else float(μ)
    up = ctx.to_device
    Σd = up(Σ_ladder)
    Σl = up(np.ascontiguousarray(np.asarray(Σ_loc)[:sP.fft_range.stop], np.complex128))
    G = ctx.empty((kG.Nq, len(sP.fft_range)), np.complex128)
    out = LambdaDmOut()
    try:
        ctx.check(_lib.lib().ldga_g_from_sigma_ladder_dev(ctx.handle, C.byref(_freq(sP)), C.byref(_model(mP)), νmax,
                                                          0.0, 0.0, int(fix_n), μ0, None, None, Σd.ptr, None, 0.0,
                                                          Σl.ptr, C.byref(out), G.ptr, None))
        G_new = G.to_host()
    finally:
        for a in (Σd, Σl, G):
            a.free()
    return (out.mu if fix_n else μ0), G_new
