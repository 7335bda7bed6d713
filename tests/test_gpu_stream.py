"""fj_set_stream: a context's calls on a caller's HIP stream (created here
through the HIP runtime the library itself is bound to) give the same bits
as on the context's own stream, and the context returns to its own stream."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_caller_stream_same_bits(ctx):
    import fjmodels_amd as fj
    p = O.guess_model(12, 0.05, 50.0, 3, 4)[0]
    ctx.set_grid(p.ar, p.ngauss, p.npoly, p.ngrid)
    ctx.set_quad(fj.FJ_QUAD_LOG22)
    ctx.set_potential(*p.tables())
    ctx.set_delta(*p.delta())
    ctx.set_df(fj.df_params("Hernquist"))
    ref = ctx.compute_new_phi()
    lib = ctx.lib
    s = C.c_void_p()
    assert lib.hipStreamCreate(C.byref(s)) == 0
    try:
        ctx.set_stream(s.value)
        got = ctx.compute_new_phi()
        ctx.set_stream(None)
        back = ctx.compute_new_phi()
    finally:
        ctx.set_stream(None)
        assert lib.hipStreamDestroy(s) == 0
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
        assert np.array_equal(back[k], ref[k]), k
