"""Specular reflection for the oracle (test helper): the trial/test space
restricted to harmonics even under the mirror of every reflecting plane at
the plane's vertices, derived independently of the library from point values
of the oracle's harmonics."""
import numpy as np

import oracle


def mirror_parity(maxk, ndir=64, seed=0):
    """[3][slm] of +1/-1: how each even-l harmonic (slm order) behaves under
    x_a -> -x_a, from point evaluations at random directions."""
    lmax = 2 * maxk
    idx = [l * l + l + m for l in range(0, lmax + 1, 2) for m in range(-l, l + 1)]
    d = np.random.default_rng(seed).standard_normal((ndir, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    par = np.zeros((3, len(idx)))
    for a in range(3):
        acc = np.zeros(len(idx))
        for v in d:
            w = v.copy()
            w[a] = -w[a]
            acc += oracle.sh_eval(lmax, *v)[idx] * oracle.sh_eval(lmax, *w)[idx]
        par[a] = np.sign(acc)
    return par


def reflect_mask(arrays, maxk, slm):
    """Boolean [nP][slm] (caller's numbering) of the vanishing entries."""
    nP = len(arrays["tetstart"]) - 1
    par = mirror_parity(maxk)
    mask = np.zeros((nP, slm), bool)
    bnd, kind = arrays["bnd"], arrays["face_kind"]
    for f in np.nonzero(kind == 1)[0]:
        s = np.abs(bnd[f]["s"])
        axis = int(np.argmax(s))
        mask[bnd[f]["p"][:3], :] |= (par[axis] < 0)[None, :]
    return mask


def masked_operator(op, mask):
    """A' f = P A P f with P zeroing the masked entries (padded layout)."""
    keep = np.zeros((op.nP, op.aslm), bool)
    keep[:, :op.slm] = ~mask
    keep = keep.reshape(-1)

    def apply(f):
        return op.lhs(f * keep) * keep
    return apply, keep


def masked_cg(op, mask, tol_rel=1e-10, maxit=20000):
    """CG on P A P x = P b from x = 0 (Jacobi on the diagonal of the
    angular-diagonal operator); returns (x, iterations)."""
    A, keep = masked_operator(op, mask)
    b = op.rhs() * keep
    d = op.lhs_diag(np.ones(op.N()))
    d = np.where(keep & (d > 0), d, 1.0)
    x = np.zeros_like(b)
    r = b.copy()
    z = r / d * keep
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    it = 0
    while np.linalg.norm(r) > tol_rel * bn and it < maxit:
        Ap = A(p)
        al = rz / (p @ Ap)
        x += al * p
        r -= al * Ap
        z = r / d * keep
        rz2 = r @ z
        p = z + rz2 / rz * p
        rz = rz2
        it += 1
    return x, it
