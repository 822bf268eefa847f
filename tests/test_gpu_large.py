"""The operator on a mesh whose FP64 coefficient records exceed 4 GiB in
total: 15 slots x 64-byte records x nP > 2^32 bytes once nP > 4.47M.  The
buffer-load schedule (variant 3) and the matrix-core schedule (variant 4)
address each slot's records through a descriptor of their own, so they stay
exact where one descriptor for all slots would wrap its 32-bit offsets;
they must give the plain-load schedule's (variant 1, 64-bit addressing,
checked against the oracle on small meshes) A f.  Bars: 1e-12 relative to
max|A f| in FP64 (the same sums, possibly contracted differently); the FP32
matrix-core schedule with FP32 products (angular precision 0) 1e-5."""
import numpy as np
import pytest

import transfer_amd as ta

pytestmark = pytest.mark.gpu


def test_buffer_schedule_beyond_4gib_of_records():
    print("[gpu-large] building a 106 x 106 x 424 box (4.87M vertices)", flush=True)
    m = ta.Mesh.box((106, 106, 424), (0, 0, 0), (1, 1, 4))
    m.set_fields(kappa=m.field("const", [1.0]), Ip=m.field("const", [1.0]))
    ctx = ta.HipContext(m, ta.Angular(1, 0.5, 0.0, 0))
    nP, slm, aslm = ctx.nP, ctx.slm, ctx.aslm
    assert 15 * 64 * nP > 2 ** 32
    default = ctx.schedule()
    print("[gpu-large] nP %d slm %d default schedule %s" % (nP, slm, default), flush=True)
    rng = np.random.default_rng(11)
    f = np.zeros((nP, aslm))
    f[:, :slm] = rng.standard_normal((nP, slm))
    d_f, d_r = ctx.vector_from(f.reshape(-1)), ctx.allocVector()
    out = {}
    for variant in (1, 2, 3):
        ctx.set_schedule(0, J=default["J"], WPB=4, variant=variant)
        ctx.computeLhs(d_f, d_r)
        out[variant] = ctx.copyToHost(d_r)
        print("[gpu-large] variant %d done" % variant, flush=True)
    # the matrix-core schedule, FP64 and FP32 (its J is the one it takes)
    for J in (2, 4, 8):
        try:
            ctx.set_schedule(0, J=J, WPB=4, variant=4)
            break
        except Exception:
            continue
    assert ctx.schedule()["variant"] == 4, ctx.schedule()
    ctx.computeLhs(d_f, d_r)
    out[4] = ctx.copyToHost(d_r)
    print("[gpu-large] variant 4 (%s) done" % ctx.kernel_name(0), flush=True)
    ctx.set_angular_precision(0)
    ctx.set_schedule(1, J=ctx.schedule()["J"], WPB=4, variant=4)
    ctx.computeLhsPrec(1, d_f, d_r)
    out32 = ctx.copyToHost(d_r)
    print("[gpu-large] FP32 variant 4 (%s) done" % ctx.kernel_name(1), flush=True)
    ref = out[1]
    scale = np.abs(ref).max()
    for variant in (2, 3, 4):
        e = np.abs(out[variant] - ref).max() / scale
        print("[gpu-large] variant %d vs 1: rel %.2e" % (variant, e), flush=True)
        assert e <= 1e-12
    e32 = np.abs(out32 - ref).max() / scale
    print("[gpu-large] FP32 variant 4 vs FP64 variant 1: rel %.2e" % e32, flush=True)
    assert e32 <= 1e-5
    ctx.deviceFree(d_f)
    ctx.deviceFree(d_r)
    ctx.close()
