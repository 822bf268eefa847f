"""The allocation-free inner product (tr_dot_noalloc) against tr_dot, the
entry point it is a variant of, through the C ABI: bit-identical results on
two calls in a row, the second with a smaller input (a second context on a
smaller mesh, the same GPU's scratch), and again on the first context after
the smaller call."""
import numpy as np
import pytest

import transfer_amd as ta

pytestmark = pytest.mark.gpu


def _vectors(ctx, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((ctx.nP, ctx.aslm))
    y = np.zeros((ctx.nP, ctx.aslm))
    x[:, :ctx.slm] = rng.standard_normal((ctx.nP, ctx.slm))
    y[:, :ctx.slm] = rng.standard_normal((ctx.nP, ctx.slm))
    return ctx.vector_from(x.ravel()), ctx.vector_from(y.ravel())


def test_dot_noalloc_matches_dot():
    ctxs = []
    for n in (40, 8):
        P = ta.problems.ball(n)
        ctx = ta.HipContext(P.mesh, P.ang)
        ctxs.append(ctx)
        dx, dy = _vectors(ctx, n)
        d_res = ctx.deviceAlloc(8)
        want = ctx.dot(dx, dy)
        ctx.dot_noalloc(dx, dy, d_res)
        got = float(ctx.copyToHost(d_res, 8)[0])
        print("[no-alloc] ball%d (%d entries): tr_dot %.17g, tr_dot_noalloc %.17g" % (n, ctx.nP * ctx.slm, want, got),
              flush=True)
        assert got == want
    # the larger context again after the smaller call: partials fully rewritten
    ctx = ctxs[0]
    dx, dy = _vectors(ctx, 7)
    d_res = ctx.deviceAlloc(8)
    ctx.dot_noalloc(dx, dy, d_res)
    assert float(ctx.copyToHost(d_res, 8)[0]) == ctx.dot(dx, dy)
