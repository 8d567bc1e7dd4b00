"""CPU FP64 oracle for the ladder hot path of LadderDGA.jl (v0.15.1).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or, for ``cpu_baseline``, as the timed CPU restatement).  The product
path (``ladderdga.jl_amd``) never imports it.

Every function restates the reference loop nest it cites, with the same
frequency bookkeeping, in numpy/scipy:

* LAPACK ``zgetrf``/``zgetri`` via ``scipy.linalg.lapack`` -- the same routine
  pair the reference calls through ``LapackWrapper.inv!``.
* FFTs via ``numpy.fft`` on the full k grid -- the reference's FFTW path.

Conventions that live in the un-vendored ``Dispersions.jl`` (k grid, irreducible
wedge, ``conv_post`` shift) are restated here as:

* full grid k_d = 2*pi*n_d/Ns, n_d in [0, Ns); eps_k = -2 t sum_d cos k_d;
* wedge = points with n_0 >= n_1 >= ... and n_d <= Ns/2 (the caller may give
  its own wedge ordering as an (Nq, D) coordinate array);
* conv(a, b)[q] = (1/Nk) sum_k a(k) b(k+q), obtained from the reference's
  tables ``ifft(fft(a) .* fft(reverse(b)))`` read at index (-q-1) mod Ns and
  divided by Nk.

These three are "parity unpinned" against Dispersions itself; they are pinned
by brute-force sums and by the filling golden value of the reference test
suite (see tests/test_oracle_golden.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.linalg import lapack as _lapack
from scipy.special import polygamma as _polygamma

# ----------------------------------------------------------------------------
# parameters (src/Config.jl:39-47, :66-75)
# ----------------------------------------------------------------------------


@dataclass
class ModelParameters:
    """src/Config.jl:39-47 (U, mu, beta, n, sVk, Epot_DMFT, Ekin_DMFT)."""
    U: float
    mu: float
    beta: float
    n: float
    sVk: float = 0.0
    Epot_DMFT: float = 0.0
    Ekin_DMFT: float = 0.0


@dataclass
class SimulationParameters:
    """src/Config.jl:66-75.  fft_range follows src/IO.jl:89-93."""
    n_iw: int
    n_inu: int
    n_inu_shell: int = 0
    shift: int = 1
    fft_range: range = None
    usable_prct_reduction: float = 0.0
    dbg_full_eom_omega: bool = False

    def __post_init__(self):
        if self.fft_range is None:
            fr = 2 * (self.n_inu + self.n_iw)
            self.fft_range = range(-fr, fr + 1)


def trunc_div2(x: int) -> int:
    """Julia ``trunc(Int, x/2)`` (rounds toward zero)."""
    return int(x / 2)


# ----------------------------------------------------------------------------
# k grid (Dispersions.jl restatement, used via src/BSETools.jl:165 etc.)
# ----------------------------------------------------------------------------


class KGrid:
    """Simple-cubic k grid with its irreducible wedge.

    ``gen_kGrid("2Dsc-t", Ns)`` / ``("3Dsc-t", Ns)`` of Dispersions.jl.
    """

    def __init__(self, D: int, Ns: int, t: float, wedge_coords=None):
        self.D, self.Ns, self.t = D, Ns, t
        self.Nk = Ns ** D
        half = Ns // 2
        full = np.stack(np.meshgrid(*([np.arange(Ns)] * D), indexing="ij"), -1).reshape(-1, D)
        # the fold of each full-grid point into the wedge (point symmetry)
        fold = np.minimum(full, Ns - full)
        canon = -np.sort(-fold, axis=1)          # descending coordinates
        if wedge_coords is None:
            keys = sorted({tuple(c) for c in canon})
            wedge_coords = np.array(keys, dtype=np.int64).reshape(-1, D)
        self.wedge = np.asarray(wedge_coords, dtype=np.int64).reshape(-1, D)
        lut = {tuple(c): i for i, c in enumerate(self.wedge)}
        self.full_to_wedge = np.array([lut[tuple(c)] for c in canon], dtype=np.int64)
        # full array shape (Ns,)*D in Julia (column-major) axis order: axis 0 = first coord
        self.full_to_wedge = self.full_to_wedge.reshape((Ns,) * D)
        self.kMult = np.bincount(self.full_to_wedge.ravel(), minlength=len(self.wedge)).astype(float)
        self.Nq = len(self.wedge)
        assert np.all(self.wedge <= half)
        self.eps_full = self._eps(np.stack(np.meshgrid(*([np.arange(Ns)] * D), indexing="ij"), -1))
        self.eps_wedge = self._eps(self.wedge)
        # representative full index for reduceKArr
        self.rep = tuple(self.wedge.T)

    def _eps(self, n):
        k = 2.0 * np.pi * np.asarray(n, dtype=float) / self.Ns
        return -2.0 * self.t * np.cos(k).sum(axis=-1)

    # Dispersions.expandKArr / reduceKArr / kintegrate
    def expand(self, f_wedge):
        return np.asarray(f_wedge)[self.full_to_wedge]

    def reduce(self, f_full):
        return np.asarray(f_full)[self.rep]

    def kintegrate(self, f_wedge, axis=0):
        f = np.moveaxis(np.asarray(f_wedge), axis, 0)
        return np.tensordot(self.kMult, f, axes=(0, 0)) / self.kMult.sum()


def gen_kgrid(name: str, Ns: int) -> KGrid:
    """Parse ``"2Dsc-t"`` / ``"3Dsc-t"`` (case-insensitive), src/runHelpers.jl:88."""
    s = name.lower()
    D = int(s[0])
    t = float(s.split("-", 1)[1])
    return KGrid(D, Ns, t)


def conv_post_index(kG: KGrid):
    """Full-grid index (-q-1) mod Ns of every wedge point q (tuple of arrays)."""
    return tuple(((-kG.wedge - 1) % kG.Ns).T)


def conv_fft(kG: KGrid, a_fft, b_rfft):
    """Dispersions ``conv_fft!``: wedge values of (1/Nk) sum_k a(k) b(k+q).

    ifft(fft(a) .* fft(reverse(b))) is the circular convolution a * rev(b);
    conv_post reads it at (-q-1) mod Ns and divides by Nk.
    """
    tmp = np.fft.ifftn(a_fft * b_rfft)
    return tmp[conv_post_index(kG)] / kG.Nk


def conv_brute(kG: KGrid, a_full, b_full):
    """O(Nk^2) definition of the convolution, wedge values.  Used to pin conv_fft."""
    Ns, D = kG.Ns, kG.D
    out = np.empty(kG.Nq, dtype=complex)
    for qi, q in enumerate(kG.wedge):
        shifted = b_full
        for d in range(D):
            shifted = np.roll(shifted, -int(q[d]), axis=d)
        out[qi] = np.mean(a_full * shifted)
    return out


# ----------------------------------------------------------------------------
# Matsubara helpers (src/GFTools.jl:31,46,85-86; src/helpers.jl:20,46-58)
# ----------------------------------------------------------------------------


def iν_array(beta, grid):
    """src/GFTools.jl:31."""
    return np.array([1j * ((2.0 * n + 1) * math.pi / beta) for n in grid])


def iω_array(beta, grid):
    """src/GFTools.jl:46."""
    return np.array([1j * ((2.0 * n) * math.pi / beta) for n in grid])


def G_from_Σ(ind, beta, mu, eps, sigma):
    """src/GFTools.jl:85-86."""
    return 1.0 / (1j * (2 * ind + 1) * math.pi / beta + mu - eps - sigma)


def νnGrid(ωn, sP: SimulationParameters):
    """src/helpers.jl:20."""
    lo, hi = -sP.n_inu - sP.n_inu_shell, sP.n_inu + sP.n_inu_shell - 1
    return [v - sP.shift * trunc_div2(ωn) for v in range(lo, hi + 1)]


def OneToIndex_to_Freq(ωi, νi, sP):
    """src/helpers.jl:46-50 (1-based indices in, Matsubara numbers out)."""
    ωn = ωi - sP.n_iw - 1
    νn = (νi - sP.n_inu - 1) - sP.shift * trunc_div2(ωn)
    return ωn, νn


def ν0Index_of_ωIndex(ωi, sP):
    """src/helpers.jl:58 (1-based)."""
    return sP.n_inu + sP.shift * trunc_div2(ωi - sP.n_iw - 1) + 1


def G_shell_sum(N, beta):
    """src/GFTools.jl:281-283."""
    return float((_polygamma(1, N + 0.5) - _polygamma(1, 0.5 - N)) * beta / (4 * math.pi ** 2) + beta / 4)


def filling(G, U, mu, beta, shell=None):
    """src/GFTools.jl:234-246 (vector forms)."""
    G = np.asarray(G)
    if shell is None:
        N = len(G) // 2
        shell = G_shell_sum(N, beta)
    return 2 * (np.real(G.sum()) / beta + 0.5 + mu * shell) / (1 + U * shell)


def filling_kG(G_wedge_nu, kG: KGrid, U, mu, beta):
    """src/GFTools.jl:244-246: filling(G::Matrix, kG, U, mu, beta)."""
    return filling(kG.kintegrate(G_wedge_nu, axis=0), U, mu, beta)


def filling_pos(G, U, mu, beta, improved_sum=True):
    """src/GFTools.jl:255-268 (vector forms)."""
    G = np.asarray(G)
    sG = G.sum()
    if not improved_sum:
        return 2 * np.real(sG + np.conj(sG)) / beta + 1
    shell = G_shell_sum(len(G) + 1, beta)
    return 2 * (np.real(sG + np.conj(sG)) / beta + 0.5 + mu * shell) / (1 + U * shell)


def filling_pos_kG(G_wedge_nu, kG, U, mu, beta, improved_sum=True):
    """src/GFTools.jl:271-273."""
    return filling_pos(kG.kintegrate(G_wedge_nu, axis=0), U, mu, beta, improved_sum)


# ----------------------------------------------------------------------------
# G tables (src/runHelpers.jl:113-130; src/helpers.jl:153-183)
# ----------------------------------------------------------------------------


def Σ_loc_at(Σ_loc, νn):
    """Negative frequencies by conjugation, src/runHelpers.jl:124."""
    return np.conj(Σ_loc[-νn - 1]) if νn < 0 else Σ_loc[νn]


def G_tables(kG: KGrid, mP: ModelParameters, sP: SimulationParameters, Σ_loc):
    """gLoc (wedge), gLoc_fft and gLoc_rfft over sP.fft_range.

    Returns arrays with the fft_range index last, offset by first(fft_range).
    """
    shape = (kG.Ns,) * kG.D
    nf = len(sP.fft_range)
    gLoc = np.empty((kG.Nq, nf), dtype=complex)
    g_fft = np.empty(shape + (nf,), dtype=complex)
    g_rfft = np.empty(shape + (nf,), dtype=complex)
    for i, νn in enumerate(sP.fft_range):
        s = Σ_loc_at(Σ_loc, νn)
        g = G_from_Σ(νn, mP.beta, mP.mu, kG.eps_full, s)
        gLoc[:, i] = kG.reduce(g)
        g_fft[..., i] = np.fft.fftn(g)
        g_rfft[..., i] = np.fft.fftn(g[(slice(None, None, -1),) * kG.D])   # Julia reverse(A)
    return gLoc, g_fft, g_rfft


def G_fft_from_wedge(G_wedge, kG: KGrid):
    """G_fft!/G_rfft! (src/helpers.jl:165-183): for every column ν of G_wedge
    (Nq x n), expandKArr to the full grid, then fft, and fft of reverse.

    Returns (G_fft, G_rfft), each gridshape x n.
    """
    G_wedge = np.asarray(G_wedge)
    shape = (kG.Ns,) * kG.D
    n = G_wedge.shape[1]
    g_fft = np.empty(shape + (n,), dtype=complex)
    g_rfft = np.empty(shape + (n,), dtype=complex)
    for i in range(n):
        g = kG.expand(G_wedge[:, i])
        g_fft[..., i] = np.fft.fftn(g)
        g_rfft[..., i] = np.fft.fftn(g[(slice(None, None, -1),) * kG.D])   # Julia reverse(A)
    return g_fft, g_rfft


# ----------------------------------------------------------------------------
# bubble (src/BSETools.jl:157-172)
# ----------------------------------------------------------------------------


def calc_bubble(g_fft, g_rfft, kG: KGrid, mP, sP, ωsel=None):
    """chi0[q, nu, omega] = -beta conv(G(nu), G(nu+omega)).  Local case Nk=1 is a product.
    ωsel (optional): only these omega indices are computed, the others stay 0 -- the same loop
    on a subset, for splitting the pass over processes."""
    f0 = sP.fft_range.start
    Nν = 2 * (sP.n_inu + sP.n_inu_shell)
    Nω = 2 * sP.n_iw + 1
    data = np.empty((kG.Nq, Nν, Nω), dtype=complex) if ωsel is None else np.zeros((kG.Nq, Nν, Nω), dtype=complex)
    for ωi, ωn in enumerate(range(-sP.n_iw, sP.n_iw + 1)):
        if ωsel is not None and ωi not in ωsel:
            continue
        νrange = [v - trunc_div2(sP.shift * ωn) for v in
                  range(-(sP.n_inu + sP.n_inu_shell), sP.n_inu + sP.n_inu_shell)]
        for νi, νn in enumerate(νrange):
            if kG.Nk == 1:
                data[:, νi, ωi] = g_fft[..., νn - f0].ravel() * g_rfft[..., νn + ωn - f0].ravel()
            else:
                data[:, νi, ωi] = conv_fft(kG, g_fft[..., νn - f0], g_rfft[..., νn + ωn - f0])
            data[:, νi, ωi] *= -mP.beta
    return data


def calc_bubble_local(gImp, mP, sP):
    """Nk = 1 bubble of the impurity G (src/runHelpers.jl:137)."""
    f0 = sP.fft_range.start
    Nν = 2 * (sP.n_inu + sP.n_inu_shell)
    Nω = 2 * sP.n_iw + 1
    data = np.empty((1, Nν, Nω), dtype=complex)
    for ωi, ωn in enumerate(range(-sP.n_iw, sP.n_iw + 1)):
        νrange = [v - trunc_div2(sP.shift * ωn) for v in
                  range(-(sP.n_inu + sP.n_inu_shell), sP.n_inu + sP.n_inu_shell)]
        for νi, νn in enumerate(νrange):
            data[0, νi, ωi] = -mP.beta * gImp[νn - f0] * gImp[νn + ωn - f0]
    return data


# ----------------------------------------------------------------------------
# BSE (src/BSETools.jl:192-248, src/LapackWrapper.jl:64-72)
# ----------------------------------------------------------------------------


def inv_lapack(A):
    """LapackWrapper.inv!: zgetrf + zgetri (src/LapackWrapper.jl:64-72)."""
    lu, piv, info = _lapack.zgetrf(A)
    inv, info2 = _lapack.zgetri(lu, piv)
    return inv, info, info2


def calc_χγ(type_: str, Γr, χ0, kG: KGrid, mP, sP):
    """Plain branch (chi_helper = nothing).  Returns (chi[q,w] real, gamma[q,nu,w])."""
    if type_ == "d":
        s = -1
    elif type_ == "m":
        s = 1
    else:
        raise ValueError("Unkown type")
    Nν = 2 * sP.n_inu
    Nq = χ0.shape[0]
    Nω = χ0.shape[2]
    sh = sP.n_inu_shell
    χ = np.empty((Nq, Nω))
    γ = np.empty((Nq, Nν, Nω), dtype=complex)
    for ωi in range(Nω):
        for qi in range(Nq):
            A = np.array(Γr[:, :, ωi], dtype=complex, order="F")
            A[np.arange(Nν), np.arange(Nν)] += 1.0 / χ0[qi, sh:sh + Nν, ωi]
            Ainv, _, _ = inv_lapack(A)
            χ[qi, ωi] = np.real(Ainv.sum()) / mP.beta ** 2
            γ[qi, :, ωi] = Ainv.sum(axis=0) / (χ0[qi, :Nν, ωi] * (1.0 + s * mP.U * χ[qi, ωi]))
    return χ, γ


def calc_gen_χ(Γr, χ0, sh=0):
    """src/BSETools.jl:259-275: inv(Gamma + diag(1/chi0)) per (q, w)."""
    Nν = Γr.shape[0]
    Nq, _, Nω = χ0.shape
    out = np.empty((Nν, Nν, Nq, Nω), dtype=complex)
    for ωi in range(Nω):
        for qi in range(Nq):
            A = Γr[:, :, ωi] + np.diag(1.0 / χ0[qi, sh:sh + Nν, ωi])
            out[:, :, qi, ωi] = np.linalg.inv(A)
    return out


# ----------------------------------------------------------------------------
# F, lambda0 (src/BSETools.jl:40-53, :118-144)
# ----------------------------------------------------------------------------


def F_from_χ(χ, G, sP, beta, diag_term=True):
    """src/BSETools.jl:40-53.  G is indexed by Matsubara number via G[n - first]."""
    Gf, first = G
    F = np.empty_like(χ)
    Nν, _, Nω = χ.shape
    for ωi in range(1, Nω + 1):
        for νpi in range(1, Nν + 1):
            ωn, νpn = OneToIndex_to_Freq(ωi, νpi, sP)
            for νi in range(1, Nν + 1):
                _, νn = OneToIndex_to_Freq(ωi, νi, sP)
                g1, g2 = Gf[νn - first], Gf[ωn + νn - first]
                g3, g4 = Gf[νpn - first], Gf[ωn + νpn - first]
                F[νi - 1, νpi - 1, ωi - 1] = -(χ[νi - 1, νpi - 1, ωi - 1]
                                               + (νn == νpn and diag_term) * beta * g1 * g2) / (g1 * g2 * g3 * g4)
    return F


def calc_λ0(χ0, F, mP, sP):
    """Plain sum branch of src/BSETools.jl:128-141."""
    Nq, _, Nω = χ0.shape
    Niν = F.shape[0]
    sh = sP.n_inu_shell
    λ0 = np.zeros((Nq, Niν, Nω), dtype=complex)
    for ωi in range(Nω):
        for νi in range(Niν):
            v1 = F[νi, :, ωi]
            λ0[:, :, ωi] += v1[None, :] * χ0[:, sh:χ0.shape[1] - sh, ωi] / mP.beta ** 2
    return λ0


# ----------------------------------------------------------------------------
# chi-lambda and the bosonic sums (LambdaCorrection/helpers.jl:18-70; DataTypes.jl:201-308)
# ----------------------------------------------------------------------------


def χ_λ(χ, λ):
    """LambdaCorrection/helpers.jl:18."""
    return χ / (λ * χ + 1)


def dχ_λ(χ, λ):
    """LambdaCorrection/helpers.jl:61."""
    return -χ_λ(χ, λ) ** 2


@dataclass
class ChiT:
    """src/DataTypes.jl:177-194 (data[q, w], tail_c, lambda, beta, usable_w)."""
    data: np.ndarray
    beta: float
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


def ωn_grid(χ: ChiT):
    """src/DataTypes.jl:201."""
    return 2j * math.pi * χ.indices_ω / χ.beta


def sum_kω(kG: KGrid, χ: ChiT, λ=None, transform=None, force_full_range=False):
    """src/DataTypes.jl:215-249 (tail treated only in 1/w^2)."""
    ω_slice = np.arange(χ.data.shape[1])
    if not force_full_range:
        ω_slice = ω_slice[list(χ.usable_ω)]
    with np.errstate(divide="ignore", invalid="ignore"):
        ωn2_tail = np.real(χ.tail_c[2] / ωn_grid(χ) ** 2)
    ωn2_tail[~np.isfinite(ωn2_tail)] = 0.0
    data = χ.data[:, ω_slice]
    if λ is not None:
        transform = lambda x: χ_λ(x, λ)
    return sum_kω_raw(kG, data, χ.beta, χ.tail_c[2], ωn2_tail, transform)


def sum_kω_raw(kG, χ, beta, c2, ωn2_tail, transform=None):
    """src/DataTypes.jl:230-249."""
    res = 0.0
    norm = kG.kMult.sum()
    for i, ωt in enumerate(ωn2_tail):
        col = χ[:, i] if transform is None else transform(χ[:, i])
        res += float(np.dot(kG.kMult, col)) / norm - ωt
    return (res + (-beta ** 2 / 12) * c2) / beta


def sum_ω(χ: ChiT, force_full_range=False):
    """src/DataTypes.jl:277-308, as written.  sum_ω! (:285-291) hands the inner
    sum_ω (:293-305) the whole ωn_arr together with the usable slice of χ, and
    the inner loop pairs slice element n with ωn_arr[n]: the n-th frequency of
    the whole grid, counted from -n_iω, not the slice's own frequency.  On the
    full range the two coincide."""
    ω_slice = list(range(χ.data.shape[1]))
    if not force_full_range:
        ω_slice = [ω_slice[i] for i in χ.usable_ω]
    ωn_arr = ωn_grid(χ)
    c = χ.tail_c
    if not math.isclose(sum(c), c[2], rel_tol=1.4901161193847656e-08):
        raise ValueError("sum_ω only implemented for 1/ω^2, but other tail coefficients are non-zero!")
    out = np.empty(χ.data.shape[0])
    for qi in range(χ.data.shape[0]):
        res = 0.0
        for n, wi in enumerate(ω_slice):
            if ωn_arr[n] != 0:
                res += χ.data[qi, wi] - np.real(c[2] / ωn_arr[n] ** 2)
            else:
                res += χ.data[qi, wi]
        out[qi] = (res + c[2] * (-χ.beta ** 2 / 12)) / χ.beta
    return out


def sum_ωk(kG, χ: ChiT, λ=None, force_full_range=False):
    """src/DataTypes.jl:257-264: kintegrate(sum_ω(χ_λ(χ, λ)))."""
    if λ is not None and λ != 0:
        χ2 = ChiT(χ_λ(χ.data, λ), χ.beta, tail_c=χ.tail_c, usable_ω=χ.usable_ω)
        return float(kG.kintegrate(sum_ω(χ2, force_full_range=force_full_range)))
    return float(kG.kintegrate(sum_ω(χ, force_full_range=force_full_range)))


def get_λ_min(χr):
    """LambdaCorrection/helpers.jl:217-220."""
    nh = math.ceil(χr.shape[1] / 2) - 1
    return -np.min(1.0 / χr[:, nh])


# ----------------------------------------------------------------------------
# EoM (src/ladderDGATools_singleCore.jl:49-81; src/ladderDGATools.jl:147-160, :260-284)
# ----------------------------------------------------------------------------

EOM = {
    "full": lambda U, gm, gd, cm, cd, l0: U * (gm * 1.5 * (1 + U * cm) - gd * 0.5 * (1 - U * cd) - 1.5 + 0.5 + l0),
    "χ_m": lambda U, gm, gd, cm, cd, l0: U * (gm * 1.5 * (U * cm)),
    "γ_m": lambda U, gm, gd, cm, cd, l0: U * (gm * 1.5) + 0 * gm,
    "χ_d": lambda U, gm, gd, cm, cd, l0: -U * (gd * 0.5 * (-U * cd)),
    "γ_d": lambda U, gm, gd, cm, cd, l0: -U * (gd * 0.5) + 0 * gd,
    "rest_01": lambda U, gm, gd, cm, cd, l0: (-U * 1.0 + 0.0j) + 0 * gm,
    "rest": lambda U, gm, gd, cm, cd, l0: U * l0,
    # src/ladderDGATools.jl:267-270
    "sp_01": lambda U, gm, gd, cm, cd, l0: U * (gm * 0.5 * (1 + U * cm) - 0.5),
    "sp_02": lambda U, gm, gd, cm, cd, l0: U * (gm * 1.0 * (1 + U * cm) - 1.0),
    "sp": lambda U, gm, gd, cm, cd, l0: U * (gm * 1.5 * (1 + U * cm) - 1.5),
    "ch": lambda U, gm, gd, cm, cd, l0: -U * (gd * 0.5 * (1 - U * cd) - 0.5),
}


def conv_eom(kG: KGrid, K_wedge, g_rfft_slice):
    """conv_tmp_add!: fft(expand(K)) .* G_rfft -> ifft -> conv_post (src/ladderDGATools.jl:147-160)."""
    if kG.Nk == 1:
        return K_wedge * g_rfft_slice.ravel()
    Kf = np.fft.fftn(kG.expand(K_wedge))
    return conv_fft(kG, Kf, g_rfft_slice)


def calc_Σ_ω(eomf, χm, γm, χd, γd, g_rfft, λ0, U, kG, sP, νmax, per_ω=False, νsel=None, ωsel=None):
    """src/ladderDGATools_singleCore.jl:28-69: the omega/nu/q loop nest (chi already lambda-mapped).
    νsel / ωsel (optional): only these ν columns (0-based ν_n) / omega indices are summed, the
    others stay 0 -- the same loop on a subset, for grids where the whole nest takes too long
    and for splitting it over processes."""
    Nq, Nω = χm.shape
    f0 = sP.fft_range.start
    Σ = np.zeros((Nq, νmax, Nω) if per_ω else (Nq, νmax), dtype=complex)
    for ωi in range(Nω):
        if ωsel is not None and ωi not in ωsel:
            continue
        ωn = ωi - sP.n_iw
        νZero = ν0Index_of_ωIndex(ωi + 1, sP)
        νlist = list(range(νZero, 2 * sP.n_inu + 1))[:νmax]
        for νii, νi in enumerate(νlist):
            if νsel is not None and νii not in νsel:
                continue
            K = eomf(U, γm[:, νi - 1, ωi], γd[:, νi - 1, ωi], χm[:, ωi], χd[:, ωi], λ0[:, νi - 1, ωi])
            c = conv_eom(kG, K, g_rfft[..., νii + ωn - f0])
            if per_ω:
                Σ[:, νii, ωi] += c
            else:
                Σ[:, νii] += c
    return Σ


def calc_Σ(χm: ChiT, γm, χd: ChiT, γd, χ_m_sum, λ0, g_rfft, kG, mP, sP,
           νmax=None, λm=0.0, λd=0.0, tc=True, νsel=None):
    """src/ladderDGATools_singleCore.jl:114-136 with calc_Σ! (:71-81).  Returns Sigma[q, nu] (nu = 0..numax-1)."""
    if νmax is None:
        νmax = sP.n_inu
    cm = χ_λ(χm.data, λm) if λm != 0 else χm.data
    cd = χ_λ(χd.data, λd) if λd != 0 else χd.data
    Σ = calc_Σ_ω(EOM["full"], cm, γm, cd, γd, g_rfft, λ0, mP.U, kG, sP, νmax, νsel=νsel)
    χm_l = ChiT(cm, χm.beta, tail_c=χm.tail_c, usable_ω=χm.usable_ω)
    tail = (-mP.U * (sum_kω(kG, χm_l) - χ_m_sum)) if tc else 0.0
    tail_c = tail / iν_array(mP.beta, range(νmax))
    return Σ / mP.beta + tail_c[None, :] + mP.n * mP.U / 2.0


def calc_Σ_parts(χm, γm, χd, γd, χ_m_sum, λ0, g_rfft, kG, mP, sP, λm=0.0, λd=0.0, tc=True):
    """src/ladderDGATools_singleCore.jl:153-189 (7 parts, nu = 0..n_inu-1)."""
    νmax = sP.n_inu
    cm = χ_λ(χm.data, λm) if λm != 0 else χm.data
    cd = χ_λ(χd.data, λd) if λd != 0 else χd.data
    χm_l = ChiT(cm, χm.beta, tail_c=χm.tail_c, usable_ω=χm.usable_ω)
    tail = (-mP.U * (sum_kω(kG, χm_l) - χ_m_sum)) if tc else 0.0
    tail_c = tail / iν_array(mP.beta, range(νmax))
    out = np.empty((χm.data.shape[0], νmax, 7), dtype=complex)
    for i, name in enumerate(["χ_m", "γ_m", "χ_d", "γ_d", "rest_01", "rest"]):
        out[:, :, i] = calc_Σ_ω(EOM[name], cm, γm, cd, γd, g_rfft, λ0, mP.U, kG, sP, νmax) / mP.beta
    out[:, :, 5] += mP.n * mP.U / 2.0
    out[:, :, 6] = tail_c[None, :]
    return out


# ----------------------------------------------------------------------------
# G ladder, energies, lambda_dm (GFTools.jl:162-186, thermodynamics.jl:142-168,
# LambdaCorrection/conditions_new.jl:219-324, helpers.jl:133-208)
# ----------------------------------------------------------------------------


def G_from_Σladder(Σ_ladder, Σ_loc, kG, mP, sP, fix_n=True, μ=None, secant_tol=1e-13):
    """src/GFTools.jl:162-186.  Returns (mu, G[q, nu in fft_range]).

    The mu root is found with a secant iteration to |f| < secant_tol instead
    of Roots.find_zero(Order0, atol=1e-8); both converge to the same root.  A
    fit that ends with |f| > 1e-8 gives mu = NaN, as the reference's does.
    """
    if μ is None:
        μ = mP.mu
    νmaxΣ = Σ_ladder.shape[1]
    last = sP.fft_range.stop - 1
    νR = np.arange(0, last + 1)
    Σfull = np.empty((kG.Nq, len(νR)), dtype=complex)
    Σfull[:, :νmaxΣ] = Σ_ladder
    for j in range(νmaxΣ, len(νR)):
        Σfull[:, j] = Σ_loc[νR[j]]
    iν = 1j * (2 * νR + 1) * math.pi / mP.beta

    def Gpos(m):
        return 1.0 / (iν[None, :] + m - kG.eps_wedge[:, None] - Σfull)

    def fμ(m):
        G = Gpos(m)
        return filling_pos_kG(G[:, :νmaxΣ], kG, mP.U, m, mP.beta) - mP.n

    if fix_n:
        μ = secant(fμ, μ, tol=secant_tol)
        # find_zero (atol 1e-8) throwing makes G_from_Σladder! return NaN (GFTools.jl:171-176)
        if not abs(fμ(μ)) <= 1e-8:
            μ = math.nan
    Gp = Gpos(μ)
    nf = len(sP.fft_range)
    G = np.empty((kG.Nq, nf), dtype=complex)
    f0 = sP.fft_range.start
    G[:, -f0:] = Gp
    # negative frequencies by conjugation: G(-n-1) = conj(G(n))  (GFTools.jl:183-184)
    G[:, :-f0] = np.conj(Gp[:, last - 1::-1])
    return μ, G


def secant(f, x0, tol=1e-13, maxit=100):
    """Root of the filling condition: a secant from x0; if it does not reach
    |f| < tol, the fallback of a bracketing solver (the filling grows with μ):
    step out from x0 until the sign changes, then Illinois false position."""
    xs = x0
    x1 = x0 + 1e-4 if x0 == 0 else x0 * (1 + 1e-4)
    f0, f1 = f(x0), f(x1)
    for _ in range(maxit):
        if abs(f1) < tol:
            return x1
        if f1 == f0:
            break
        x2 = x1 - f1 * (x1 - x0) / (f1 - f0)
        x0, f0, x1 = x1, f1, x2
        f1 = f(x1)
    a = xs
    fa = f(a)
    d = 1.0 if fa < 0 else -1.0
    h = 1.0
    b = a + d * h
    fb = f(b)
    k = 0
    while not (fa * fb <= 0) and k < 60:
        a, fa = b, fb
        h *= 2.0
        b = a + d * h
        fb = f(b)
        k += 1
    if not (fa * fb <= 0):
        return x1
    side, c = 0, x1
    for _ in range(200):
        c = (a * fb - b * fa) / (fb - fa)
        fc = f(c)
        if abs(fc) < tol or abs(b - a) <= 4e-16 * max(abs(a), abs(b)):
            return c
        if fc * fb > 0:
            b, fb = c, fc
            if side == -1:
                fa *= 0.5
            side = -1
        else:
            a, fa = c, fc
            if side == 1:
                fb *= 0.5
            side = 1
    return c


def calc_E_ED(ϵk, Vk, GImp, U, n, μ, β):
    """src/thermodynamics.jl:46-64: kinetic and potential energy of the Anderson impurity from its
    bath (ϵk, Vk) and G_imp(iν_n), n = 0..len-1, with the 1/(iν)^2 tails summed analytically."""
    ϵk, Vk, GImp = np.asarray(ϵk, float), np.asarray(Vk, float), np.asarray(GImp, complex)
    vk = float(np.sum(Vk ** 2))
    Σ_h = n * U / 2
    E_pot_tail = (U ** 2) / 2 * n * (1 - n / 2) - Σ_h * (Σ_h - μ)
    E_kin_tail = vk
    iν = iν_array(β, range(len(GImp)))
    E_kin = E_pot = 0.0
    for m in range(len(GImp)):
        Δ = np.sum((Vk * np.conj(Vk)) / (iν[m] - ϵk))
        Σ_n = iν[m] - Δ - 1.0 / GImp[m] + μ
        E_kin += 2 * np.real(GImp[m] * Δ - E_kin_tail / iν[m] ** 2)
        E_pot += 2 * np.real(GImp[m] * Σ_n - E_pot_tail / iν[m] ** 2)
    E_kin = E_kin * (2 / β) - (β / 2) * E_kin_tail
    E_pot = E_pot * (1 / β) + 0.5 * Σ_h - (β / 4) * E_pot_tail
    return float(E_kin), float(E_pot)


def calc_E(G, Σ, μ, kG, mP, sP):
    """src/thermodynamics.jl:142-168 (trace=false).  G over fft_range, Sigma over 0..numax-1."""
    νmax = min(sP.fft_range.stop - 1, Σ.shape[1] - 1)
    νGrid = np.arange(0, νmax + 1)
    iν_n = 1j * (2 * νGrid + 1) * math.pi / mP.beta
    Σ_h = mP.n * mP.U / 2
    ek = kG.eps_wedge
    E_kin_tail_c = ek + Σ_h - μ
    E_pot_tail_c = mP.U ** 2 * 0.5 * mP.n * (1 - 0.5 * mP.n) + Σ_h * (ek + Σ_h - μ)
    tail = 1.0 / iν_n ** 2
    E_pot_tail = E_pot_tail_c[:, None] * tail[None, :]
    E_kin_tail = E_kin_tail_c[:, None] * tail[None, :]
    E_pot_tail_inv = (mP.beta / 2) * Σ_h + (mP.beta / 2) * (-mP.beta / 2) * E_pot_tail_c
    E_kin_tail_inv = (mP.beta / 2) * ek * (1 + -mP.beta * E_kin_tail_c)
    f0 = sP.fft_range.start
    Gs = G[:, νGrid - f0]
    E_pot_full = np.real(Gs * Σ[:, νGrid] - E_pot_tail)
    E_kin_full = ek[:, None] * np.real(Gs - E_kin_tail)
    E_kin = kG.kintegrate(4 * E_kin_full.sum(axis=1) + E_kin_tail_inv) / mP.beta
    E_pot = kG.kintegrate(2 * E_pot_full.sum(axis=1) + E_pot_tail_inv) / mP.beta
    return float(np.real(E_kin)), float(np.real(E_pot))


def gen_νω_indices(χm: ChiT, χd: ChiT, mP, sP):
    """LambdaCorrection/helpers.jl:280-289 (+ GFFit.jl:81-84)."""
    if sP.dbg_full_eom_omega:
        ωind = list(range(χm.data.shape[1]))
    else:
        ωind = sorted(set(χm.usable_ω) & set(χd.usable_ω))
    νmax = min(sP.n_inu, (3 * len(ωind)) // 8)
    return ωind, νmax


def λdm_residual_vals(λ, χm, γm, χd, γd, Σ_loc, g_rfft, χloc_m_sum, λ0, kG, mP, sP, νmax, μ0, tc=True):
    """residual_vals of LambdaCorrection/conditions_new.jl:250-270."""
    if λ[0] > 1e5 or λ[1] > 1e5:
        return (np.nan,) * 6 + (None, None)
    Σ = calc_Σ(χm, γm, χd, γd, χloc_m_sum, λ0, g_rfft, kG, mP, sP, νmax=νmax, λm=λ[0], λd=λ[1], tc=tc)
    μ, G = G_from_Σladder(Σ, Σ_loc, kG, mP, sP, fix_n=True, μ=μ0)
    E_kin, E_pot_1 = calc_E(G, Σ, μ, kG, mP, sP)
    f0 = sP.fft_range.start
    n = filling_pos_kG(G[:, -f0:-f0 + νmax], kG, mP.U, μ, mP.beta)
    χm_sum = sum_kω(kG, ChiT(χ_λ(χm.data, λ[0]), χm.beta, tail_c=χm.tail_c, usable_ω=χm.usable_ω))
    χd_sum = sum_kω(kG, ChiT(χ_λ(χd.data, λ[1]), χd.beta, tail_c=χd.tail_c, usable_ω=χd.usable_ω))
    lhs_c1 = (χd_sum + χm_sum) / 2
    E_pot_2 = (mP.U / 2) * (χd_sum - χm_sum) + mP.U * (mP.n / 2 * mP.n / 2)
    return n, μ, E_kin, E_pot_1, E_pot_2, lhs_c1, Σ, G


def run_sc_fix(χm, γm, χd, γd, λ0, μ, Σ_loc, g_rfft, χloc_m_sum, kG, mP, sP, νmax, λm, λd,
               maxit=100, mixing=0.2, conv_abs=1e-8, fit_μ=True, tc=True):
    """run_sc! (LambdaCorrection/conditions_new.jl:405-560) for type :fix: λ fixed, no χ-tail update.

    Each iteration: EoM with the current G_rfft (:479), mixing with the previous Σ (:484),
    G_from_Σladder! with the μ fit (:485), calc_E (:490) and the G_rfft! rebuild (:491);
    converged when both potential energies change by less than conv_abs (:503).
    """
    f0 = sP.fft_range.start
    χm_l = ChiT(χ_λ(χm.data, λm), χm.beta, tail_c=χm.tail_c, usable_ω=χm.usable_ω)
    χd_l = ChiT(χ_λ(χd.data, λd), χd.beta, tail_c=χd.tail_c, usable_ω=χd.usable_ω)
    χm_sum, χd_sum = sum_kω(kG, χm_l), sum_kω(kG, χd_l)
    lhs_c1 = (χd_sum + χm_sum) / 2
    rhs_c1 = mP.n / 2 * (1 - mP.n / 2)
    E_pot_2 = (mP.U / 2) * (χd_sum - χm_sum) + mP.U * (mP.n / 2 * mP.n / 2)
    E_pot_1, E_kin = math.inf, math.inf
    g_rfft = np.array(g_rfft, copy=True)
    Σ, G, it, converged = None, None, 1, False
    while True:
        Σ_work = Σ
        Σ = calc_Σ(χm, γm, χd, γd, χloc_m_sum, λ0, g_rfft, kG, mP, sP, νmax=νmax, λm=λm, λd=λd, tc=tc)
        if mixing != 0 and it > 1:
            Σ = (1 - mixing) * Σ + mixing * Σ_work
        μ, G = G_from_Σladder(Σ, Σ_loc, kG, mP, sP, fix_n=fit_μ, μ=μ)
        if math.isnan(μ):
            break
        E_pot_1_old, E_pot_2_old = E_pot_1, E_pot_2
        E_kin, E_pot_1 = calc_E(G, Σ, μ, kG, mP, sP)
        _, g_rfft = G_fft_from_wedge(G, kG)
        if abs(E_pot_1 - E_pot_1_old) < conv_abs and abs(E_pot_2 - E_pot_2_old) < conv_abs:
            converged = True
            break
        if it >= maxit:
            break
        it += 1
    n = filling_pos_kG(G[:, -f0:-f0 + νmax], kG, mP.U, μ, mP.beta)
    return dict(μ=μ, n=n, E_kin=E_kin, E_pot_1=E_pot_1, E_pot_2=E_pot_2, lhs_c1=lhs_c1, rhs_c1=rhs_c1,
                sc_converged=converged, it=it, Σ=Σ, G=G)


def newton_right_1d(f, df, start, min_, nsteps=100, atol=1e-13):
    """LambdaCorrection/helpers.jl:133-153."""
    δ = 1e-4
    x0 = start
    xi = x0
    i = 1
    while True:
        fi = f(xi)
        dfii = 1 / df(xi)
        xi = x0 - dfii * fi
        if abs(fi) < atol:
            break
        if xi < min_:
            x0 = min_ + δ + abs(min_ - x0) / 2
        else:
            x0 = xi
        if i >= nsteps:
            break
        i += 1
    return xi


def finite_difference_jacobian(f, x, fx=None):
    """Forward differences with relstep sqrt(eps) (FiniteDiff.jl default)."""
    x = np.asarray(x, dtype=float)
    if fx is None:
        fx = np.asarray(f(x))
    J = np.empty((len(fx), len(x)))
    rel = math.sqrt(np.finfo(float).eps)
    for j in range(len(x)):
        h = rel * max(abs(x[j]), 1.0)
        xp = x.copy()
        xp[j] += h
        J[:, j] = (np.asarray(f(xp)) - fx) / h
    return J


def newton_right_nd(f, start, min_, nsteps=500, atol=1e-8, max_reset=5, reset_backoff=1.0, jac=None):
    """LambdaCorrection/helpers.jl:165-208 (jac=None -> finite differences)."""
    start = np.asarray(start, dtype=float)
    min_ = np.asarray(min_, dtype=float)
    xi = start.copy()
    i = 1
    reset_i = 1
    bisection_fail = False
    while True:
        fi = np.asarray(f(xi), dtype=float)
        J = jac(xi) if jac is not None else finite_difference_jacobian(f, xi, fi)
        dfii = np.linalg.inv(J)
        xi_last = xi.copy()
        xi = xi - dfii @ fi
        for l in range(len(xi)):
            if xi[l] < min_[l]:
                xi[l] = min_[l] + abs(xi_last[l] - min_[l]) / 2
                if abs(xi_last[l] - min_[l]) / 2 < 1e-8:
                    bisection_fail = True
        if bisection_fail or not np.all(np.isfinite(fi)):
            bisection_fail = False
            if reset_i >= max_reset:
                raise RuntimeError("lambda root-finding returned NaN")
            reset_i += 1
            xi = start + reset_i * reset_backoff * np.abs(min_)
        if abs(np.linalg.norm(fi)) < atol:
            break
        if i >= nsteps:
            break
        i += 1
    return xi


def λm_correction(χm: ChiT, rhs, kG, mP, sP):
    """LambdaCorrection/conditions_new.jl:144-172 (returns lambda_m)."""
    λm_min = get_λ_min(χm.data)
    χr = χm.data[:, list(χm.usable_ω)]
    idx = np.arange(-sP.n_iw, sP.n_iw + 1)[list(χm.usable_ω)]
    iωn = 2j * idx * math.pi / mP.beta
    with np.errstate(divide="ignore", invalid="ignore"):
        ωn2_tail = np.real(χm.tail_c[2] / iωn ** 2)
    ωn2_tail[np.abs(iωn) < 1e-10] = 0.0
    f = lambda l: sum_kω_raw(kG, χr, χm.beta, χm.tail_c[2], ωn2_tail, lambda x: χ_λ(x, l)) - rhs
    df = lambda l: sum_kω_raw(kG, χr, χm.beta, χm.tail_c[2], ωn2_tail, lambda x: dχ_λ(x, l))
    return newton_right_1d(f, df, 0.0, λm_min)


def λdm_correction(χm, γm, χd, γd, Σ_loc, g_rfft, χloc_m_sum, λ0, kG, mP, sP,
                   νmax=-1, λ_min_δ=1e-4, tc=True, jac=None, λinit=(0.0, 0.0), μ0=None):
    """LambdaCorrection/conditions_new.jl:219-324 (lambda values only).  Every residual's μ fit
    starts from μ0 (default mP.mu)."""
    μ0 = mP.mu if μ0 is None else μ0
    _, νmax_d = gen_νω_indices(χm, χd, mP, sP)
    νmax = νmax_d if νmax < 1 else min(νmax_d, νmax)
    rhs_c1 = mP.n / 2 * (1 - mP.n / 2)

    def res(l):
        n, μ, _, ep1, ep2, lhs, _, _ = λdm_residual_vals(l, χm, γm, χd, γd, Σ_loc, g_rfft, χloc_m_sum, λ0,
                                                         kG, mP, sP, νmax, μ0, tc)
        return np.array([lhs - rhs_c1, ep1 - ep2])

    lm_min = get_λ_min(χm.data)
    ld_min = get_λ_min(χd.data)
    min_λ = np.array([lm_min + λ_min_δ * abs(lm_min), ld_min + λ_min_δ * abs(ld_min)])
    return newton_right_nd(res, list(λinit), min_λ, jac=jac), νmax


def run_sc_dm(χm, γm, χd, γd, λ0, μ, Σ_loc, g_rfft, χloc_m_sum, kG, mP, sP, νmax, λm, λd,
              maxit=100, mixing=0.2, conv_abs=1e-8, tc=True):
    """run_sc! (LambdaCorrection/conditions_new.jl:405-560) for types :dm / :pre_dm, from (λm, λd).

    The first iteration runs at the given λ (:431-435); every later one first re-solves λdm with the
    current G_rfft, starting from the last λ and μ (:437-449), and stops the loop when no root
    validates (:465-471).  After each rebuild the sums are taken at the current λ (:494-499).
    Returns the run_sc_fix fields plus the last λ.
    """
    f0 = sP.fft_range.start
    rhs_c1 = mP.n / 2 * (1 - mP.n / 2)

    def sums(lm, ld):
        s_m = sum_kω(kG, ChiT(χ_λ(χm.data, lm), χm.beta, tail_c=χm.tail_c, usable_ω=χm.usable_ω))
        s_d = sum_kω(kG, ChiT(χ_λ(χd.data, ld), χd.beta, tail_c=χd.tail_c, usable_ω=χd.usable_ω))
        return (s_d + s_m) / 2, (mP.U / 2) * (s_d - s_m) + mP.U * (mP.n / 2 * mP.n / 2)

    lhs_c1, E_pot_2 = sums(λm, λd)
    E_pot_1, E_kin = math.inf, math.inf
    g_rfft = np.array(g_rfft, copy=True)
    Σ, G, it, converged = None, None, 1, False
    while True:
        if it > 1:
            root, _ = λdm_correction(χm, γm, χd, γd, Σ_loc, g_rfft, χloc_m_sum, λ0, kG, mP, sP, νmax=νmax, tc=tc,
                                     λinit=(λm, λd), μ0=μ)
            if not np.all(np.isfinite(root)):
                λm, λd = root
                break
            n_r, μ_r, _, ep1, ep2, lhs, _, _ = λdm_residual_vals(root, χm, γm, χd, γd, Σ_loc, g_rfft, χloc_m_sum,
                                                                  λ0, kG, mP, sP, νmax, μ, tc)
            λm, λd = root
            if not (abs(rhs_c1 - lhs) <= 1e-8 and abs(ep1 - ep2) <= 1e-8):
                break
            μ = μ_r
        Σ_work = Σ
        Σ = calc_Σ(χm, γm, χd, γd, χloc_m_sum, λ0, g_rfft, kG, mP, sP, νmax=νmax, λm=λm, λd=λd, tc=tc)
        if mixing != 0 and it > 1:
            Σ = (1 - mixing) * Σ + mixing * Σ_work
        μ, G = G_from_Σladder(Σ, Σ_loc, kG, mP, sP, fix_n=True, μ=μ)
        if math.isnan(μ):
            break
        E_pot_1_old, E_pot_2_old = E_pot_1, E_pot_2
        E_kin, E_pot_1 = calc_E(G, Σ, μ, kG, mP, sP)
        _, g_rfft = G_fft_from_wedge(G, kG)
        lhs_c1, E_pot_2 = sums(λm, λd)
        if abs(E_pot_1 - E_pot_1_old) < conv_abs and abs(E_pot_2 - E_pot_2_old) < conv_abs:
            converged = True
            break
        if it >= maxit:
            break
        it += 1
    n = filling_pos_kG(G[:, -f0:-f0 + νmax], kG, mP.U, μ, mP.beta) if G is not None else math.nan
    return dict(μ=μ, n=n, E_kin=E_kin, E_pot_1=E_pot_1, E_pot_2=E_pot_2, lhs_c1=lhs_c1, rhs_c1=rhs_c1,
                sc_converged=converged, it=it, Σ=Σ, G=G, λm=λm, λd=λd)


# ----------------------------------------------------------------------------
# small numeric helpers pinned by the reference test-suite
# ----------------------------------------------------------------------------


def reduce_range(rng, red_prct):
    """src/helpers.jl:141-146 on a 1-based range."""
    first, last = rng[0], rng[-1]
    L = len(rng)
    sub = math.floor(L / 2 * red_prct)
    lst = max(last - sub, math.ceil(L / 2 + (L % 2 == 0)))
    fst = min(first + sub, math.ceil(L / 2))
    return range(fst, lst + 1)
