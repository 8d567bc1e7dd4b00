"""k grid of the simple-cubic lattices (the part of Dispersions.jl the path uses).

Mirrors ``gen_kGrid("2Dsc-t" | "3Dsc-t", Ns)`` as called at
src/runHelpers.jl:88: the full Ns^D grid k_d = 2 pi n_d / Ns, the dispersion
eps_k = -2 t sum_d cos k_d and the irreducible wedge (n_0 >= n_1 >= ...,
n_d <= Ns/2) with multiplicities kMult.  The wedge itself is produced by the
native library (ldga_default_wedge).
"""
from __future__ import annotations

import numpy as np

from . import _lib


class KGrid:
    def __init__(self, D: int, Ns: int, t: float):
        if D not in (2, 3):
            raise ValueError("only 2D and 3D simple-cubic grids are supported")
        self.D, self.Ns, self.t = D, Ns, float(t)
        self.Nk = Ns ** D
        w, m = _lib.default_wedge(D, Ns)
        self.wedge = w                          # (Nq, D) int32, descending coordinates
        self.kMult = m                          # (Nq,)
        self.Nq = len(m)
        self.eps_wedge = self.dispersion_at(w)  # ϵkGrid
        full = np.stack(np.meshgrid(*([np.arange(Ns)] * D), indexing="ij"), -1)
        self.eps_full = self.dispersion_at(full)
        fold = np.minimum(full, Ns - full)
        canon = -np.sort(-fold, axis=-1)
        lut = {tuple(c): i for i, c in enumerate(w.tolist())}
        self.full_to_wedge = np.vectorize(lambda *c: lut[c])(*np.moveaxis(canon, -1, 0)).astype(np.int64)

    @property
    def ϵkGrid(self):
        return self.eps_wedge

    def gridshape(self):
        return (self.Ns,) * self.D

    def dispersion_at(self, n):
        k = 2.0 * np.pi * np.asarray(n, dtype=float) / self.Ns
        return -2.0 * self.t * np.cos(k).sum(axis=-1)

    def expandKArr(self, f_wedge):
        return np.asarray(f_wedge)[self.full_to_wedge]

    def reduceKArr(self, f_full):
        return np.asarray(f_full)[tuple(self.wedge.T)]


def gen_kGrid(name: str, Ns: int) -> KGrid:
    """``gen_kGrid("2Dsc-1.0", 64)`` -> 2D square lattice with t = 1.0."""
    s = name.strip().lower()
    if not (s[1:4] == "dsc" and s[0] in "23"):
        raise ValueError(f"unsupported grid {name!r} (2Dsc-t / 3Dsc-t)")
    return KGrid(int(s[0]), Ns, float(s.split("-", 1)[1]))
