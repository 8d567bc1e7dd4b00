"""Parameter and Matsubara-function containers of the reference path.

ModelParameters / SimulationParameters follow src/Config.jl:39-47, :66-75;
χ0T (χ₀T), χT and γT follow src/DataTypes.jl:52-69, :177-194, :358-364.
Arrays keep the reference's index order ([q, ν, ω]) and column-major
(Fortran) memory layout, so they can be handed to libldga_hip unchanged.
"""
from __future__ import annotations

import math
from dataclasses import InitVar, dataclass, field

import numpy as np


@dataclass
class ModelParameters:
    U: float
    μ: float
    β: float
    n: float
    sVk: float = 0.0
    Epot_DMFT: float = 0.0
    Ekin_DMFT: float = 0.0

    @property
    def mu(self):
        return self.μ

    @property
    def beta(self):
        return self.β


@dataclass
class SimulationParameters:
    n_iω: int
    n_iν: int
    n_iν_shell: int = 0
    shift: int = 1
    χ_helper: object = None
    fft_range: range = None
    usable_prct_reduction: float = 0.0
    dbg_full_eom_omega: bool = False

    def __post_init__(self):
        if self.fft_range is None:       # src/IO.jl:89-93
            fr = 2 * (self.n_iν + self.n_iω)
            self.fft_range = range(-fr, fr + 1)
        if self.χ_helper is not None:
            raise NotImplementedError("only plain sums (χ_helper = nothing) are implemented")


def F(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


@dataclass
class χ0T:
    """Bubble χ₀(q, ν, ω) (src/DataTypes.jl:52-69); ``asym`` is not formed (plain sums)."""
    data: np.ndarray
    β: float
    ν_shell_size: int = 0
    axis_types: dict = field(default_factory=lambda: {"q": 0, "ν": 1, "ω": 2})

    def core(self):
        s = self.ν_shell_size
        return self.data[:, s:self.data.shape[1] - s, :]


def find_usable_χ_interval(χ, sum_type="common", reduce_range_prct: float = 0.1) -> range:
    """Usable ω indices (0-based range) of χ(ω), or of χ(q, ω) intersected over q (src/GFFit.jl:31-73):
    the largest interval around ω = 0 where χ stays positive and falls off monotonically on both
    sides, shrunk by reduce_range_prct.  sum_type "full" gives every index, a pair (a, b) the 1-based
    a:b.  Over q the reference fixes reduce_range_prct at 0.1 (GFFit.jl:72), and so does this."""
    χ = np.asarray(χ, dtype=np.float64)
    if χ.ndim == 2:
        rs = [find_usable_χ_interval(row, sum_type="common", reduce_range_prct=0.1) for row in χ]
        lo, hi = max(r.start for r in rs), min(r.stop for r in rs)
        return range(lo, max(lo, hi))
    n = len(χ)
    if n % 2 == 0:
        raise ValueError("Finding a usable interval is only implemented for uneven size of χ_ω!")
    mid = math.ceil(n / 2) - 1                # 0-based centre
    if sum_type == "full":
        return range(n)
    if isinstance(sum_type, tuple):
        return range(sum_type[0] - 1, sum_type[1])
    if χ[mid] < 0.0:
        return range(mid, mid + 1)
    d = np.diff(χ)
    r1 = 0                                    # positivity
    while r1 < mid and χ[mid - r1 - 1] > 0 and χ[mid + r1 + 1] > 0:
        r1 += 1
    r2 = 0                                    # rising before the centre, falling after it
    while r2 < mid and d[mid - r2 - 1] > 0 and d[mid + r2] < 0:
        r2 += 1
    r = math.ceil(min(r1, r2) * (1 - reduce_range_prct))
    return range(mid - r, mid + r + 1)


@dataclass
class χT:
    """Physical susceptibility χ(q, ω) (src/DataTypes.jl:177-194).  As there, the usable ω range is
    the whole axis unless full_range=False, which finds it with find_usable_χ_interval."""
    data: np.ndarray
    β: float
    tail_c: list = field(default_factory=list)
    λ: float = 0.0
    usable_ω: range = None
    indices_ω: np.ndarray = None
    full_range: InitVar[bool] = True
    reduce_range_prct: InitVar[float] = 0.0

    def __post_init__(self, full_range=True, reduce_range_prct=0.0):
        Nω = self.data.shape[1]
        if self.usable_ω is None:
            self.usable_ω = range(Nω) if full_range else find_usable_χ_interval(
                self.data, reduce_range_prct=reduce_range_prct)
        if self.indices_ω is None:
            self.indices_ω = np.arange(1, Nω + 1) - math.ceil(Nω / 2)

    @property
    def shape(self):
        return self.data.shape


@dataclass
class γT:
    """Triangular vertex γ(q, ν, ω) (src/DataTypes.jl:358-364)."""
    data: np.ndarray
    axis_types: dict = field(default_factory=lambda: {"q": 0, "ν": 1, "ω": 2})

    @property
    def shape(self):
        return self.data.shape
