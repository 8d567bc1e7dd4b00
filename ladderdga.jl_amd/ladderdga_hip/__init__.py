"""ladderdga_hip -- MI355X (gfx950) ladder hot path of LadderDGA.jl.

Python mirror of the reference API over the C ABI of libldga_hip.so.
"""
from ._lib import Context, LdgaError, default_context, lib  # noqa: F401
from .kgrid import KGrid, gen_kGrid  # noqa: F401
from .types import ModelParameters, SimulationParameters, χ0T, χT, γT, find_usable_χ_interval  # noqa: F401
from .api import (G_fft, calc_bubble, calc_χγ, calc_λ0, calc_Σ, calc_Σ_ω, calc_Σ_parts, χ_λ, dχ_λ, χ_λ_inplace, reset,  # noqa: F401
                  sum_kω, sum_ωk, sum_ω, Σ_loc_correction, get_λ_min, F_from_χ, calc_gen_χ, update_tail)

from .lambda_correction import (λ_result, newton_right, newton_right_1d, λm_rhs, λm_correction, λdm_correction, run_sc,  # noqa: F401
                                λm_correction_full, λ_correction, lDΓAHelper, G_from_Σladder,
                                LambdaDmProblem, gen_νω_indices)
from . import par  # noqa: F401
from .par import (calc_bubble_par, calc_χγ_par, initialize_EoM, calc_Σ_par, calc_Σ_par_inplace,  # noqa: F401
                  clear_wcache, collect_χ0, collect_χ, collect_γ, update_wcache, update_wcaches_G_rfft,
                  set_reference_semantics)

__all__ = ["set_reference_semantics", "λ_result", "newton_right", "newton_right_1d", "λm_rhs", "λm_correction", "λm_correction_full", "λ_correction", "lDΓAHelper", "G_from_Σladder", "λdm_correction", "run_sc", "LambdaDmProblem",
           "gen_νω_indices", "Context", "LdgaError", "KGrid", "gen_kGrid", "ModelParameters", "SimulationParameters", "χ0T", "χT", "find_usable_χ_interval",
           "γT", "G_fft", "calc_bubble", "calc_χγ", "calc_λ0", "calc_Σ", "calc_Σ_ω", "calc_Σ_parts", "χ_λ", "dχ_λ", "χ_λ_inplace",
           "reset", "sum_kω", "sum_ωk", "sum_ω", "Σ_loc_correction", "get_λ_min", "F_from_χ", "calc_gen_χ", "update_tail", "par", "calc_bubble_par", "calc_χγ_par", "initialize_EoM", "calc_Σ_par",
           "calc_Σ_par_inplace", "clear_wcache", "collect_χ0", "collect_χ", "collect_γ", "update_wcache",
           "update_wcaches_G_rfft"]
