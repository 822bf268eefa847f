"""The coarse operators of the multilevel preconditioner (CPU): with the
aggregation P (x) I, the library's coarse stencil applied to g equals
P^T A P g computed from the fine stencil, for plain, scattering (HG, ten
channels) and boundary-dominated meshes; and it stays symmetric positive
definite."""
import numpy as np
import pytest

import transfer_amd as ta
from test_frontend import stencil_apply


def galerkin(st, cst, agg, g):
    nC = len(agg) - 1
    owner = np.repeat(np.arange(nC), np.diff(agg))
    f = g[owner]                                        # P g
    r = stencil_apply(st, f)                            # A P g
    out = np.zeros_like(g)
    np.add.at(out, owner, r)                            # P^T A P g
    return out


@pytest.mark.parametrize("case", ["slab", "ball-hg", "cube"])
def test_coarse_operator_is_the_galerkin_product(case, rng):
    if case == "slab":
        P = ta.problems.slab(n=3)
    elif case == "ball-hg":
        P = ta.problems.ball(8, maxk=2, kind="hg")
    else:
        P = ta.problems.cube(6, maxk=1, seed=3)
    st = ta.Stencil(P.mesh.view, P.ang.view)
    cst, agg = st.coarsen(8)
    assert cst.nP == len(agg) - 1 and agg[-1] == st.nP
    g = rng.standard_normal((cst.nP, cst.slm))
    want = galerkin(st, cst, agg, g)
    got = stencil_apply(cst, g)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # two levels down, and symmetric positive definite on a small case
    c2, agg2 = cst.coarsen(8)
    g2 = rng.standard_normal((c2.nP, c2.slm))
    assert np.abs(stencil_apply(c2, g2) - galerkin(cst, c2, agg2, g2)).max() <= 1e-12 * np.abs(
        stencil_apply(c2, g2)).max()
    if c2.nP * c2.slm <= 600:
        n = c2.nP * c2.slm
        A = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1
            A[:, j] = stencil_apply(c2, e.reshape(c2.nP, c2.slm)).ravel()
        assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
        assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
