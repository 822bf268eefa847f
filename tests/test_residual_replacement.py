"""The residual replacement of the mixed solver (cg_replace_residual in
solver.hip), restated in numpy on a random SPD system with a Jacobi
preconditioner.

The mixed solver runs its inner CG on A d = r / |r| and, at a refinement
step, replaces the stopped CG's recursive residual by the new true residual
in the new scale (units 1/s of the old ones, s = |r_old| / |r_new|).  It
keeps the last direction p_k, rescaled and with its component along the new
r taken out: p = z + beta s p_k' with beta = r.z / (s^2 r_k.z_k).  With an
exact operator the true residual equals the recursive one up to rounding, so
the continued CG must reproduce the uninterrupted one: the iterates agree to
rounding, at every split point.  The check has power: a beta with the wrong
power of s, or a direction kept without its rescaling, leaves iterates
orders of magnitude further off."""
import numpy as np
import pytest


def spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, cond, n)
    a = (q * ev) @ q.T
    d = rng.uniform(0.5, 4.0, n)  # a diagonal scaling, so Jacobi does something
    return (a * d[:, None]) * d[None, :]


def pcg(a, b, dinv, iters, x=None, r=None, p=None, rz=None):
    """Plain PCG from (x, r, p, rz) or from zero; returns the state and the
    r.z of the last direction (the scalar the replacement needs)."""
    if x is None:
        x = np.zeros_like(b)
        r = b.copy()
        z = dinv * r
        p = z.copy()
        rz = r @ z
    rz_dir = rz
    for _ in range(iters):
        ap = a @ p
        alpha = rz / (p @ ap)
        x = x + alpha * p
        r = r - alpha * ap
        z = dinv * r
        rz_new = r @ z
        rz_dir = rz
        p_dir = p
        p = z + (rz_new / rz) * p
        rz = rz_new
    # the state when the solve stops: the last direction whole (the p update
    # of the stopping iteration is skipped)
    return x, r, p_dir, rz_dir


def replace(a, b, dinv, xsol, p_k, rz_dir, scale_old, variant="exact"):
    """One refinement step: the new true residual in the new scale and the
    kept direction (cg_replace_residual)."""
    r64 = b - a @ xsol
    rn = np.linalg.norm(r64)
    s = scale_old / rn
    r = r64 / rn
    z = dinv * r
    rr, rz, rp = r @ r, r @ z, r @ p_k
    if variant == "exact":
        b1 = rz / (s * rz_dir)
    elif variant == "beta_s1":  # a plausible slip: beta without the rescaled r_k.z_k
        b1 = rz / rz_dir
    else:  # the direction kept in the old units
        b1 = rz / (s * s * rz_dir)
    c = rp / rr
    p = z + b1 * (p_k - c * r)
    return r, p, rz, rn


def continued(a, b, dinv, k1, k2, variant="exact"):
    rn0 = np.linalg.norm(b)
    d, r_rec, p_k, rz_dir = pcg(a, b / rn0, dinv, k1)
    x = rn0 * d
    r, p, rz, rn1 = replace(a, b, dinv, x, p_k, rz_dir, rn0, variant)
    d2, _, _, _ = pcg(a, b / rn1, dinv, k2, x=np.zeros_like(b), r=r, p=p, rz=rz)
    return x + rn1 * d2


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("k1", [5, 17, 30])
def test_replacement_continues_the_uninterrupted_cg(seed, k1):
    # ten iterations after the split: finite-precision CG amplifies the
    # rounding difference between the true and the recursive residual with
    # every further iteration (1e-15 after 10, up to 1e-5 after 25 at this
    # conditioning)
    n, k2 = 160, 10
    a = spd(n, 1e3, seed)
    b = np.random.default_rng(seed + 10).standard_normal(n)
    dinv = 1.0 / np.diag(a)
    rn0 = np.linalg.norm(b)
    ref = rn0 * pcg(a, b / rn0, dinv, k1 + k2)[0]
    x = continued(a, b, dinv, k1, k2)
    scale = np.abs(ref).max()
    err = np.abs(x - ref).max() / scale
    assert err <= 1e-10, err
    # a restart at the same point (the previous behaviour) is measurably
    # different, so the agreement above is the continuation's
    d_restart = pcg(a, (b - a @ (rn0 * pcg(a, b / rn0, dinv, k1)[0])), dinv, k2)[0]
    x_restart = rn0 * pcg(a, b / rn0, dinv, k1)[0] + d_restart
    assert np.abs(x_restart - ref).max() / scale > 100 * max(err, 1e-14)
    for wrong in ("beta_s1", "old_units"):
        xw = continued(a, b, dinv, k1, k2, wrong)
        assert np.abs(xw - ref).max() / scale > 1e3 * max(err, 1e-14), wrong
