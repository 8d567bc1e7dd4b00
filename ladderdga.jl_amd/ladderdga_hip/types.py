"""Parameter and Matsubara-function containers of the reference path.

ModelParameters / SimulationParameters follow src/Config.jl:39-47, :66-75;
χ0T (χ₀T), χT and γT follow src/DataTypes.jl:52-69, :177-194, :358-364.
Arrays keep the reference's index order ([q, ν, ω]) and column-major
(Fortran) memory layout, so they can be handed to libldga_hip unchanged.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

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


@dataclass
class χT:
    """Physical susceptibility χ(q, ω) (src/DataTypes.jl:177-194)."""
    data: np.ndarray
    β: float
    tail_c: list = field(default_factory=list)
    λ: float = 0.0
    usable_ω: range = None
    indices_ω: np.ndarray = None

    def __post_init__(self):
        Nω = self.data.shape[1]
        if self.usable_ω is None:
            self.usable_ω = range(Nω)
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
