"""Host front end: AngularData, generators and MeshData geometry, CSR
incidence, RNG, Config, and the FP64 assembly of the device operator data
(applied here in numpy) against the oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle
import transfer_amd as ta
from conftest import ROOT, ang_dict


@pytest.mark.parametrize("K,albedo,g,L", [(1, 0, 0, 0), (2, 0, 0, 0), (3, 0, 0, 0), (5, 0, 0, 0), (6, 0, 0, 0), (7, 0, 0, 0),
                                          (1, 0.5, 0, 0), (3, 0.9, 0.6, 7), (5, 0.9, 0.6, 7), (7, 0.9, 0.6, 7)])
def test_product_angular_matches_oracle(K, albedo, g, L):
    a = ta.Angular(K, albedo, g, L)
    o = oracle.angular(K, albedo, g, L)
    assert a.slm == o["slm"]
    assert np.array_equal(a.omega_pos, o["omega_pos"])
    np.testing.assert_allclose(a.omega, o["omega"], atol=2e-14)
    for k in ("Ox", "Oy", "Oz"):
        np.testing.assert_allclose(getattr(a, k), o[k], atol=2e-14)
    np.testing.assert_allclose(a.massfac, o["massfac"], atol=1e-15)


def test_generator_sizes():
    # closed forms: 6 tets per cube, (n+1)^3 lattice points, 2 triangles per boundary square
    s = ta.problems.slab()
    assert (s.mesh.nT, s.mesh.nP, s.mesh.nF) == (6 * 12 * 12 * 48, 13 * 13 * 49,
                                                 2 * 2 * (12 * 12 + 2 * 12 * 48))
    c = ta.problems.cube(8)
    assert (c.mesh.nT, c.mesh.nP, c.mesh.nF) == (6 * 512, 729, 2 * 6 * 64)
    b = ta.problems.ball(16)
    coords, tets, cen, vol = b.mesh.geometry()
    # every kept cube's centroid lies in the ball; volume = kept cubes * h^3
    assert np.all(np.linalg.norm(cen, axis=1) <= 1.0 + 2 * np.sqrt(3) / 16)
    assert abs(vol.sum() - b.mesh.nT / 6 * (2 / 16) ** 3) < 1e-12


def test_geometry_closed_forms():
    m = ta.Mesh.box((3, 4, 5), (0, 0, 0), (1, 2, 3), jitter=0.15, seed=9)
    arr = m.arrays()
    coords, tets, cen, vol = m.geometry()
    s = arr["mesh"]["s"]
    assert abs(vol.sum() - 6.0) < 1e-12
    np.testing.assert_allclose(s.sum(axis=1), 0, atol=1e-14)           # MeshData.cpp:74-76
    P = coords[tets]                                                    # [nT][4][3]
    for j in range(4):
        for mm in range(4):
            if mm == j:
                continue
            # s_j = 3 V grad e_j:  s_j . (p_j - p_m) = 3 V
            d = np.einsum("ta,ta->t", s[:, j], P[:, j] - P[:, mm])
            np.testing.assert_allclose(d, 3 * vol, rtol=1e-12)
    np.testing.assert_allclose(arr["mesh"]["kappa_volume"], vol, rtol=1e-15)
    # boundary faces: outward S n, total area of the box surface
    bnd = arr["bnd"]
    area = np.linalg.norm(bnd["s"], axis=1)
    assert abs(area.sum() - 2 * (1 * 2 + 2 * 3 + 1 * 3)) < 1e-12
    fc = coords[bnd["p"][:, :3]].mean(axis=1)
    assert np.all(np.einsum("fa,fa->f", bnd["s"], fc - np.array([0.5, 1.0, 1.5])) > 0)


def test_csr_incidence():
    m = ta.Mesh.box((2, 3, 2), (0, 0, 0), (1, 1, 1))
    arr = m.arrays()
    ts, ti, tp = arr["tetstart"], arr["tetidx"], arr["tetpos"]
    p = arr["mesh"]["p"]
    seen = set()
    for v in range(m.nP):
        for j in range(ts[v], ts[v + 1]):
            assert p[ti[j], tp[j]] == v
            seen.add((ti[j], tp[j]))
        assert list(ti[ts[v]:ts[v + 1]]) == sorted(ti[ts[v]:ts[v + 1]])
    assert len(seen) == 4 * m.nT
    fs, fi, fp = arr["facestart"], arr["faceidx"], arr["facepos"]
    for v in range(m.nP):
        for j in range(fs[v], fs[v + 1]):
            assert arr["bnd"]["p"][fi[j], fp[j]] == v
    assert fs[-1] == 3 * m.nF


def splitmix64(x):
    M = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & M
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
    return x ^ (x >> 31)


def test_rng_is_splitmix64():
    seed, stream = 42, 3
    base = splitmix64(seed ^ ((stream * 0xD1B54A32D192ED03) & ((1 << 64) - 1)))
    want = [(splitmix64((base + i) & ((1 << 64) - 1)) >> 11) / 2.0 ** 53 for i in range(16)]
    np.testing.assert_array_equal(ta.rng_uniform(seed, stream, 16), want)


def test_config_grammar():
    path = os.path.join(ROOT, "tests", "golden", "two_regions.cfg")
    c = ta.Config(path)
    assert (c.maxk, c.device, c.dump, c.prec, c.soltype, c.nreg) == (3, 0, 1, 1, 2, 2)
    assert c.kappa[1] == 11 and c.Ip[1] == 1 and c.kappa[2] == 0.25 and c.Ip[2] == 0
    assert c.meshfile == "mesh_coarse.vol" and c.outfile == "solution.vtk"
    r = oracle.ref()
    if r is not None:  # the reference's own parser, built from its sources
        ints = (ctypes.c_int * 5)()
        k, i = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        mf, of = ctypes.create_string_buffer(1024), ctypes.create_string_buffer(1024)
        assert r.ref_config(path.encode(), ints, 3, k, i, mf, of) == 0
        assert list(ints) == [c.maxk, c.device, c.dump, c.prec, c.soltype]
        assert list(k)[1:] == [c.kappa[1], c.kappa[2]] and list(i)[1:] == [c.Ip[1], c.Ip[2]]
        assert mf.value.decode() == c.meshfile


def test_config_errors(tmp_path):
    p = tmp_path / "bad.cfg"
    p.write_text("[region 1]\nkappa = 1\n")
    with pytest.raises(ta.TransferError, match="before numregs"):
        ta.Config(str(p))
    p.write_text("numregs = 1\n[region 3]\n")
    with pytest.raises(ta.TransferError, match="out of bounds"):
        ta.Config(str(p))


# ---------------------------------------------------------------------------
def stencil_apply(st, f, diag=False, ddiag=None, bdiag=None):
    """Apply the assembled data in numpy; f is [nP][slm] in the internal numbering."""
    T = np.zeros((st.NC, st.nP, st.slm))
    for s in range(st.W):
        T += st.coef[s][:, :, None] * f[st.nbr[s]][None]
    R = st.massfac[None, :] * T[0]
    if diag:
        R += T[1] * ddiag[0] + T[2] * ddiag[1] + T[3] * ddiag[2]
    else:
        for c in range(1, st.NC):
            R += T[c] @ st.omc[c - 1].T
    if st.nB:
        Tb = np.zeros((3, st.nB, st.slm))
        for s in range(st.WB):
            Tb += st.bcoef[s][:, :, None] * f[st.bnbr[s]][None]
        for n in range(3):
            R[st.bvert] += Tb[n] * bdiag[n] if diag else Tb[n] @ st.onc[n].T
    return R


@pytest.mark.parametrize("case", ["slab-iso", "ball-plain", "ball-hg", "reflective"])
def test_stencil_assembly_matches_oracle(case, rng):
    if case == "slab-iso":
        P = ta.problems.slab(n=4)
    elif case == "ball-plain":
        P = ta.problems.ball(12)
    elif case == "ball-hg":
        P = ta.problems.ball(10, maxk=2, kind="hg")
    else:
        P = ta.problems.slab(n=3, albedo=0.0)
        P.mesh.set_face_kind_plane(0, 0.0, 1)
        P.mesh.set_face_kind_plane(0, 1.0, 1)
    a = P.ang
    st = ta.Stencil(P.mesh.view, a.view)
    arr = P.mesh.arrays()
    # reflective faces (face_kind 1) carry no boundary term, in the oracle too
    op = oracle.OracleProblem(arr, ang_dict(a))
    f = np.zeros((op.nP, op.aslm))
    f[:, :op.slm] = rng.standard_normal((op.nP, op.slm))
    fi_ = f[st.perm, :op.slm]
    ddiag = a.omega.reshape(3, a.slm, a.slm)[:, np.arange(a.slm), np.arange(a.slm)]
    bdiag = np.array([np.diag(getattr(a, k).reshape(a.slm, a.slm)) for k in ("Ox", "Oy", "Oz")])
    for diag in (False, True):
        ro = (op.lhs_diag if diag else op.lhs)(f.ravel()).reshape(op.nP, op.aslm)[:, :op.slm]
        rs = np.empty_like(fi_)
        rs[st.perm] = stencil_apply(st, fi_, diag, ddiag, bdiag)
        assert np.abs(rs - ro).max() <= 1e-13 * np.abs(ro).max()
    b = op.rhs().reshape(op.nP, op.aslm)[:, 0]
    np.testing.assert_allclose(st.rhs, b[st.perm], rtol=1e-14, atol=1e-300)
    assert sorted(st.perm) == list(range(op.nP))


def test_zero_opacity_rejected_by_assembly():
    """kappa = 0 in one tet has no operator (kernels.cu:252 divides by
    9 kappa V): the host assembly fails with TR_ERR_ARG, before any device
    work."""
    m = ta.Mesh.box((3, 3, 3), (0, 0, 0), (1, 1, 1))
    ar = m.arrays()
    ar["mesh"]["kappa_volume"][7] = 0.0           # tr_mesh_set_fields refuses kappa <= 0 itself
    a = ta.Angular(1)
    with pytest.raises(ta.TransferError) as e:
        ta.Stencil(ta.mesh_view_of(ar), a.view)
    assert e.value.code == ta.TR_ERR_ARG and "kappa" in str(e.value)
    with pytest.raises(ta.TransferError):
        m.set_fields(kappa=np.where(np.arange(m.nT) == 7, 0.0, 1.0), Ip=m.field("const", [1.0]))
