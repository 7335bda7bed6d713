"""fj_rhofDF_noalloc (the entry point that allocates no device memory) against
fj_rhofDF through the C ABI, bit for bit, on two calls in a row (the second
with fewer points, so the static scratch array holds the first call's
contents); free device memory is unchanged by the calls; sizes beyond the
maximum are refused."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def free_device_memory(lib):
    """hipMemGetInfo of the HIP runtime the library itself is bound to (looked
    up through the library's handle, which searches its dependencies), not of
    another copy of the runtime that the process may also have loaded."""
    free, total = C.c_size_t(), C.c_size_t()
    assert lib.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_rhofDF(prec):
    import fjmodels_amd as fj
    p = O.guess_model(40, 0.01, 100.0, 5, 20)[0]
    ctx = fj.Context(0)
    try:
        ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
        ctx.set_quad(fj.FJ_QUAD_LOG22)
        ctx.set_precision(prec)
        ctx.set_potential(*p.tables())
        ctx.set_delta(*p.delta())
        ctx.set_df(fj.df_params("Hernquist"))
        calls = [(np.array([0.9, 3.0, 0.02, 40.0]), np.array([0.3, 1.0, 0.01, 25.0])),
                 (np.array([1.7, 0.2]), np.array([0.05, 0.6]))]
        refs = []
        for R, z in calls:
            out6, mom = ctx.rhofDF(R, z)
            refs.append((out6, mom, ctx.stats()))
        before = free_device_memory(ctx.lib)
        for (R, z), (r6, rm, rst) in zip(calls, refs):
            print("noalloc call with", len(R), "points, precision", prec, flush=True)
            g6, gm = ctx.rhofDF_noalloc(R, z)
            st = ctx.stats()
            assert np.array_equal(g6, r6), np.max(np.abs(g6 - r6) / np.abs(r6))
            for a, b in zip(gm, rm):
                assert np.array_equal(a, b)
            for k in ("bound_nodes", "qk_calls", "brent_iters", "unbound_nodes", "grad_evals", "phi_evals"):
                assert st[k] == rst[k], k
        assert free_device_memory(ctx.lib) == before
        with pytest.raises(fj.FJError):
            ctx.rhofDF_noalloc(np.ones(fj.NOALLOC_MAX_POINTS + 1), np.ones(fj.NOALLOC_MAX_POINTS + 1))
    finally:
        ctx.close()
