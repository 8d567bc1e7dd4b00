"""The allocation-free BSE entry point (ldga_chi_gamma_noalloc_dev) against ldga_chi_gamma_dev through the
C ABI: two calls in a row on one GPU, the second on a smaller input, in FP64 and in mixed precision with an
ill-conditioned Γ so that the FP64 re-solve list in the static scratch array is used.  The two entry points
run the same kernels, so χ, γ and the report agree bit for bit; both also meet the 1e-10 bar against the
oracle's LAPACK solve."""
import ctypes as C

import numpy as np
import pytest

from oracle import ldga_oracle as O

pytestmark = pytest.mark.gpu

CALLS = [(dict(D=3, Ns=4, n_iν=60, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863), [(9, 2), (8, 0)]),
         (dict(D=2, Ns=4, n_iν=4, n_iω=1), [(0, 0)])]


def _run(ctx, fn, inp, Γ, χ0, prec):
    from ladderdga_hip import _lib
    from ladderdga_hip.api import _freq
    Nq, Nν, Nω = inp.kG.Nq, 2 * inp.sP.n_iν, 2 * inp.sP.n_iω + 1
    dΓ, dχ0 = ctx.to_device(np.asfortranarray(Γ)), ctx.to_device(np.asfortranarray(χ0))
    dχ, dγ = ctx.empty((Nq, Nω), np.float64), ctx.empty((Nq, Nν, Nω), np.complex128)
    rep = _lib.BseReport()
    ctx.check(getattr(_lib.lib(), fn)(ctx.handle, 1, prec, C.byref(_freq(inp.sP)), inp.mP.U, inp.mP.β, dΓ.ptr,
                                      dχ0.ptr, dχ.ptr, dγ.ptr, C.byref(rep), None))
    out = dχ.to_host(), dγ.to_host(), {k: getattr(rep, k) for k, _ in rep._fields_}
    for a in (dΓ, dχ0, dχ, dγ):
        a.free()
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_bse_matches_existing_entry_point(ctx, prec):
    from test_gpu_parity import _make_ill_conditioned, setup
    sizes = []
    for kw, targets in CALLS:
        inp, oK, omP, osP = setup("tiny", **kw)
        ctx.set_kgrid(inp.kG)
        χ0 = O.calc_bubble(inp.G_fft, inp.G_rfft, oK, omP, osP)
        Γ = _make_ill_conditioned(inp.Γ_m, χ0, targets) if prec == 1 else inp.Γ_m
        sizes.append((inp.kG.Nq * (2 * inp.sP.n_iω + 1), 2 * inp.sP.n_iν))
        print(f"no-alloc BSE precision={prec} Nq={inp.kG.Nq} Nν={2 * inp.sP.n_iν} Nω={2 * inp.sP.n_iω + 1}",
              flush=True)
        χa, γa, ra = _run(ctx, "ldga_chi_gamma_noalloc_dev", inp, Γ, χ0, prec)
        χb, γb, rb = _run(ctx, "ldga_chi_gamma_dev", inp, Γ, χ0, prec)
        assert np.array_equal(χa, χb) and np.array_equal(γa, γb)
        assert ra == rb
        if prec == 1:
            assert ra["n_fallback"] >= 1 and ra["n_fp64_resolved"] == ra["n_fallback"]
        rχ, rγ = O.calc_χγ("m", Γ, χ0, oK, omP, osP)
        # the ill-conditioned matrices (κ = 1e9: the targets and the q sharing their χ₀) are off the exact
        # solution by ~κ 2^-53 in any FP64 solve; every other (q, ω) meets the bar
        κ = np.array([[np.linalg.cond(Γ[:, :, w] + np.diag(1.0 / χ0[q, :, w])) for w in range(Γ.shape[2])]
                      for q in range(χ0.shape[0])])
        keep = κ <= 1e8
        assert np.max(np.abs(χa[keep] - rχ[keep])) <= 1e-10 * np.max(np.abs(rχ[keep]))
    assert sizes[1][0] < sizes[0][0] and sizes[1][1] < sizes[0][1]


def test_noalloc_bse_rejects_sizes_beyond_its_limit(ctx):
    """Nν above 128 would need the global-memory FP64 kernel's workspace: a clear LDGA_ERR_SHAPE, no launch."""
    from ladderdga_hip import _lib
    from ladderdga_hip.api import _freq
    from test_gpu_parity import setup
    inp, oK, omP, osP = setup("tiny", D=3, Ns=4, n_iν=65, n_iω=1, U=1.5, β=12.0, n=1.0, t=0.408248290463863)
    ctx.set_kgrid(inp.kG)
    rc = _lib.lib().ldga_chi_gamma_noalloc_dev(ctx.handle, 1, 1, C.byref(_freq(inp.sP)), inp.mP.U, inp.mP.β,
                                               C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), None, None)
    assert rc == _lib.LDGA_ERR_SHAPE
    assert "Nnu <= 128" in _lib.lib().ldga_last_error(ctx.handle).decode()
