"""The reference's worker-pool API on one GPU (src/ladderDGATools.jl:40-247,
src/parallelization_helpers.jl:132-420, the cache filled by setup_LDGA at
src/runHelpers.jl:189-197).

LadderDGA.jl splits ω (bubble, BSE) or ν (EoM) over Distributed workers; each
keeps its part in a worker cache (`wcache`) and `collect_*` gathers the parts.
Here the one GPU is the only worker and `wcache` holds device buffers: the G
tables, χ₀, χ and γ of both channels, λ₀ and the EoM output stay resident
between calls, so the chain calc_bubble_par -> calc_χγ_par -> initialize_EoM
-> calc_Σ_par uploads its inputs once and downloads only what is collected.
Names and argument meaning follow the reference; Julia's `calc_Σ_par!`,
`update_wcache!`, `update_wcaches_G_rfft!` and `clear_wcache!` drop the `!`
(`calc_Σ_par!` becomes `calc_Σ_par_inplace`), and `collect_χ₀` is spelled
`collect_χ0` (a subscript digit is no Python identifier character).

Differences kept on purpose (by default; set_reference_semantics(True)
switches both to the reference's worker code as written):
- calc_χγ_par solves the plain BSE, as calc_χγ does.  The reference's worker
  kernel (ladderDGATools.jl:92-95) requires the asymptotic helper of BSE_SC
  and throws without it; that helper is not part of this library.
- calc_Σ_par with tc=False divides the ω sum by β, as calc_Σ does.  The
  reference's worker EoM (ladderDGATools.jl:189) applies 1/β only together
  with the tail term, so its tc=false result is β times larger (the Hartree
  term, added when collecting (ladderDGATools.jl:201-215), is not scaled).
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _lib
from ._lib import BseReport, default_context
from .api import _freq, _model
from .kgrid import KGrid
from .types import ModelParameters, SimulationParameters, χ0T, χT, γT

c128, f64 = np.complex128, np.float64


class WorkerCache:
    """Device-resident state of the single worker (the reference's WorkerCache,
    parallelization_helpers.jl:97-159)."""

    def __init__(self):
        self._bufs = {}
        self.clear()

    def clear(self):
        for a in self._bufs.values():
            a.free()
        self._bufs = {}
        self.initialized = {}
        self.ctx = None
        self.kG = self.mP = self.sP = None
        self.χ_meta = {}                 # channel -> host χT of the resident χ (tail, usable_ω)
        self.χloc_m_sum = 0.0
        self.EoMVars_initialized = False
        self.EoM_νGrid = range(0, 0)
        self.reference_semantics = getattr(self, "reference_semantics", False)

    def put(self, name: str, host: np.ndarray | None = None, shape=None, dtype=None):
        """Store an array on the device under `name` (upload `host`, or allocate `shape`)."""
        ctx = self.context()
        self.drop(name)
        self._bufs[name] = ctx.to_device(host) if host is not None else ctx.empty(shape, dtype)
        self.initialized[name] = True
        return self._bufs[name]

    def drop(self, name: str):
        a = self._bufs.pop(name, None)
        if a is not None:
            a.free()
        self.initialized.pop(name, None)

    def get(self, name: str):
        if name not in self._bufs:
            raise RuntimeError(f"worker cache holds no {name}: compute or update it first")
        return self._bufs[name]

    def has(self, name: str) -> bool:
        return name in self._bufs

    def context(self):
        if self.ctx is None:
            self.ctx = default_context()
        if self.kG is not None:
            self.ctx.set_kgrid(self.kG)
        return self.ctx


wcache = WorkerCache()


def update_wcache(name: str, val, override: bool = True, ctx=None):
    """update_wcache!(name, val; override) (parallelization_helpers.jl:161-171): kG, mP, sP,
    or the G tables G_fft / G_fft_reverse (gridshape(kG) x length(fft_range)), which go to
    the device."""
    if ctx is not None:
        wcache.ctx = ctx
    if wcache.initialized.get(name, False) and not override:
        warnings.warn(f"Value of {name} already set.")
        return
    if name in ("kG", "mP", "sP"):
        setattr(wcache, name, val)
        wcache.initialized[name] = True
        if name == "kG":
            wcache.context()
    elif name in ("G_fft", "G_fft_reverse"):
        wcache.put(name, np.asarray(val, dtype=c128))
    else:
        raise ValueError(f"update_wcache: unsupported field {name!r}")


def update_wcaches_G_rfft(G):
    """update_wcaches_G_rfft!(G) (parallelization_helpers.jl:173-180)."""
    update_wcache("G_fft_reverse", G)


def set_reference_semantics(on: bool = True):
    """Follow the reference's worker code where it differs from the serial functions: calc_χγ_par raises
    as bse_inv does without the BSE_SC helper (ladderDGATools.jl:92), and calc_Σ_par(tc=False) returns
    the ω sum without the 1/β (ladderDGATools.jl:189).  Off by default; kept across clear_wcache."""
    wcache.reference_semantics = bool(on)


def clear_wcache():
    """clear_wcache!() (parallelization_helpers.jl:313-324): frees every device buffer."""
    wcache.clear()


def _need(*names):
    for n in names:
        if not wcache.initialized.get(n, False):
            raise RuntimeError("Add workers and run lDGA_setup before calling parallel functions!"
                               f" (worker cache has no {n})")


# ---------------------------------------------------------------------------
# bubble (ladderDGATools.jl:40-90)
# ---------------------------------------------------------------------------
def calc_bubble_par(kG: KGrid, mP: ModelParameters, sP: SimulationParameters, collect_data: bool = True):
    """χ₀ from the cached G tables into the worker cache; returned when collect_data."""
    _need("G_fft", "G_fft_reverse")
    ctx = wcache.context()
    Gf, Gr = wcache.get("G_fft"), wcache.get("G_fft_reverse")
    if Gf.shape != kG.gridshape() + (len(sP.fft_range),):
        raise ValueError("cached G tables do not match gridshape(kG) x length(fft_range)")
    Nν, Nω = 2 * (sP.n_iν + sP.n_iν_shell), 2 * sP.n_iω + 1
    χ0 = wcache.put("χ₀", shape=(kG.Nq, Nν, Nω), dtype=c128)
    ctx.check(_lib.lib().ldga_bubble_dev(ctx.handle, C.byref(_freq(sP)), mP.β, Gf.ptr, Gr.ptr, χ0.ptr, None))
    return collect_χ0(kG, mP, sP) if collect_data else None


def collect_χ0(kG: KGrid, mP: ModelParameters, sP: SimulationParameters) -> χ0T:
    """collect_χ0 (ladderDGATools.jl:67-90)."""
    return χ0T(wcache.get("χ₀").to_host(), mP.β, ν_shell_size=sP.n_iν_shell)


# ---------------------------------------------------------------------------
# χ, γ (ladderDGATools.jl:130-145, parallelization_helpers.jl:365-420)
# ---------------------------------------------------------------------------
def calc_χγ_par(type_: str, Γr, kG: KGrid, mP: ModelParameters, sP: SimulationParameters,
                collect_data: bool = True, precision: str = "fp64"):
    """χ and γ of channel type_ ("m" or "d") from the cached χ₀; returned when collect_data."""
    t = type_.lstrip(":")
    s = {"m": 1, "d": -1}.get(t)
    if s is None:
        raise ValueError("Unkown type")
    if wcache.reference_semantics:
        raise RuntimeError("Current version ONLY supports BSE_Asym_Helper!")
    ctx = wcache.context()
    χ0 = wcache.get("χ₀")
    Nν, Nω = 2 * sP.n_iν, 2 * sP.n_iω + 1
    if Γr.shape != (Nν, Nν, Nω):
        raise ValueError(f"Γr must be {Nν}x{Nν}x{Nω}")
    Γ = wcache.put("Γ", np.asarray(Γr, dtype=c128))
    χ = wcache.put("χ" + t, shape=(kG.Nq, Nω), dtype=f64)
    γ = wcache.put("γ" + t, shape=(kG.Nq, Nν, Nω), dtype=c128)
    rep = BseReport()
    prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
    rc = _lib.lib().ldga_chi_gamma_dev(ctx.handle, s, prec, C.byref(_freq(sP)), mP.U, mP.β, Γ.ptr, χ0.ptr,
                                       χ.ptr, γ.ptr, C.byref(rep), None)
    wcache.drop("Γ")
    ctx.check(rc, allow=(_lib.LDGA_ERR_ZERO_PIVOT,))
    if rc == _lib.LDGA_ERR_ZERO_PIVOT:
        warnings.warn(f"{rep.n_zero_pivot} BSE matrices are exactly singular; their χ, γ are NaN")
    χh = χT(χ.to_host(), mP.β, tail_c=[0.0, 0.0, mP.Ekin_DMFT])
    wcache.χ_meta[t] = χh
    wcache.EoMVars_initialized = False
    return (χh, collect_γ(t, kG, mP, sP)) if collect_data else None


def collect_χ(type_: str, kG: KGrid, mP: ModelParameters, sP: SimulationParameters) -> χT:
    """collect_χ (parallelization_helpers.jl:365-399)."""
    t = type_.lstrip(":")
    return χT(wcache.get("χ" + t).to_host(), mP.β, tail_c=[0.0, 0.0, mP.Ekin_DMFT])


def collect_γ(type_: str, kG: KGrid, mP: ModelParameters, sP: SimulationParameters) -> γT:
    """collect_γ (parallelization_helpers.jl:401-420)."""
    if wcache.EoMVars_initialized:
        raise RuntimeError("Cannot collect γ after initialize_EoM has been called!")
    return γT(wcache.get("γ" + type_.lstrip(":")).to_host())


# ---------------------------------------------------------------------------
# EoM (parallelization_helpers.jl:207-279, ladderDGATools.jl:168-247)
# ---------------------------------------------------------------------------
def initialize_EoM(G_fft_reverse, χloc_m_sum: float, λ0, νGrid, kG: KGrid, mP: ModelParameters,
                   sP: SimulationParameters, force_reinit: bool = False, χ_m: χT | None = None,
                   γ_m: γT | None = None, χ_d: χT | None = None, γ_d: γT | None = None):
    """Makes G_fft_reverse, λ₀, χ and γ of both channels resident for calc_Σ_par.  χ/γ not
    given are the ones calc_χγ_par left in the cache.  νGrid must be 0:νmax-1."""
    νGrid = range(νGrid.start, νGrid.stop) if isinstance(νGrid, range) else range(νGrid[0], νGrid[-1] + 1)
    if νGrid.start != 0 or len(νGrid) < 1 or len(νGrid) > sP.n_iν:
        raise ValueError("νGrid must be 0:νmax-1 with νmax <= n_iν")
    force_reinit = force_reinit or wcache.EoM_νGrid != νGrid or not wcache.EoMVars_initialized
    if not force_reinit:
        return
    update_wcache("kG", kG)
    update_wcache("mP", mP)
    update_wcache("sP", sP)
    for t, χx, γx in (("m", χ_m, γ_m), ("d", χ_d, γ_d)):
        if χx is not None:
            if χx.λ != 0:
                raise RuntimeError("λ parameter already set; initialize_EoM takes uncorrected χ")
            wcache.put("χ" + t, np.asarray(χx.data, dtype=f64))
            wcache.χ_meta[t] = χx
        elif t not in wcache.χ_meta:
            raise RuntimeError(f"no χ{t} in the worker cache: pass χ_{t} or run calc_χγ_par first")
        if γx is not None:
            wcache.put("γ" + t, np.asarray(γx.data, dtype=c128))
        else:
            wcache.get("γ" + t)
    χm, χd = wcache.χ_meta["m"], wcache.χ_meta["d"]
    Nω = χm.data.shape[1]
    if not (list(χm.usable_ω) == list(χd.usable_ω) == list(range(Nω))):
        raise RuntimeError("ERROR! usable ranges of susceptibilities are not over the complete ω-inteval. "
                           "Use non-parallel version")
    update_wcache("G_fft_reverse", G_fft_reverse)
    wcache.put("λ₀", np.asarray(λ0, dtype=c128))
    wcache.put("Σ_ladder", shape=(kG.Nq, len(νGrid)), dtype=c128)
    wcache.χloc_m_sum = float(np.real(χloc_m_sum))
    wcache.EoM_νGrid = νGrid
    wcache.EoMVars_initialized = True


def calc_Σ_par(λm: float = 0.0, λd: float = 0.0, collect_data: bool = True, tc: bool = True):
    """Σ_ladder[q, ν] (ν in the νGrid of initialize_EoM) at (λm, λd) from the resident arrays;
    None when not collect_data (the result stays in the cache)."""
    if not wcache.EoMVars_initialized:
        raise RuntimeError("Call initialize_EoM before calc_Σ_par in order to initialize worker caches.")
    return calc_Σ_par_inplace(None, λm=λm, λd=λd, collect_data=collect_data, tc=tc)


def calc_Σ_par_inplace(Σ_ladder, λm: float = 0.0, λd: float = 0.0, collect_data: bool = True, tc: bool = True):
    """calc_Σ_par! (ladderDGATools.jl:231-247): as calc_Σ_par, written into Σ_ladder
    (complex, Nq x νmax) when given."""
    if not wcache.EoMVars_initialized:
        raise RuntimeError("Call initialize_EoM before calc_Σ_par in order to initialize worker caches.")
    ctx = wcache.context()
    kG, mP, sP = wcache.kG, wcache.mP, wcache.sP
    νmax = len(wcache.EoM_νGrid)
    g = wcache.get
    mo = _model(mP, wcache.χ_meta["m"], wcache.χ_meta["d"])
    ctx.check(_lib.lib().ldga_sigma_dev(ctx.handle, C.byref(_freq(sP)), C.byref(mo), νmax, λm, λd,
                                        int(tc), 0, wcache.χloc_m_sum, g("χm").ptr, g("γm").ptr, g("χd").ptr,
                                        g("γd").ptr, g("λ₀").ptr, g("G_fft_reverse").ptr, g("Σ_ladder").ptr,
                                        None))
    if not collect_data:
        return None
    Σ = g("Σ_ladder").to_host()
    if wcache.reference_semantics and not tc:
        hartree = mP.n * mP.U / 2.0
        Σ = (Σ - hartree) * mP.β + hartree
    if Σ_ladder is None:
        return Σ
    if Σ_ladder.shape != Σ.shape:
        raise ValueError(f"Σ_ladder must be {Σ.shape}")
    Σ_ladder[...] = Σ
    return Σ_ladder


__all__ = ["wcache", "WorkerCache", "update_wcache", "update_wcaches_G_rfft", "clear_wcache", "set_reference_semantics",
           "calc_bubble_par",
           "collect_χ0", "calc_χγ_par", "collect_χ", "collect_γ", "initialize_EoM", "calc_Σ_par",
           "calc_Σ_par_inplace"]
