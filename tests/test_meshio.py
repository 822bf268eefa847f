"""Mesh input and solution output of the reference workflow (CPU): the
Netgen .vol reader against the generator's own mesh written as .vol (with
comments, ignored sections and half of the tets in the opposite
orientation), region fields from a cfg file, error paths, and the legacy
VTK writer read back."""
import numpy as np
import pytest

import transfer_amd as ta


def write_vol(path, coords, tets, mat, flip):
    with open(path, "w") as f:
        f.write("mesh3d\ndimension\n3\ngeomtype\n0\n\n")
        f.write("# surfnr    bcnr   domin  domout      np      p1      p2      p3\nsurfaceelements\n2\n")
        f.write(" 1 1 1 0 3 1 2 3\n 1 1 1 0 3 2 3 4\n\n")
        f.write("#  matnr      np      p1      p2      p3      p4\nvolumeelements\n%d\n" % len(tets))
        for t, m, fl in zip(tets, mat, flip):
            p = [int(x) + 1 for x in t]
            if fl:
                p[2], p[3] = p[3], p[2]
            f.write("%d 4 %d %d %d %d\n" % (m, *p))
        f.write("\n# pnum             edgenr\nedgesegmentsgi2\n1\n 1 1 1 2 0 1 1 0 0 1 0 0\n\n")
        f.write("#          X             Y             Z\npoints\n%d\n" % len(coords))
        for x in coords:
            f.write("%.17g %.17g %.17g\n" % tuple(x))
        f.write("endmesh\n")


def write_cfg(path):
    with open(path, "w") as f:
        f.write("numregs = 2\nmeshfile = x.vol\nangorder = 1\n[region 1]\nkappa = 11\nIp = 1\n"
                "[region 2]\nkappa = 0.25\nIp = 0\n")


@pytest.fixture(scope="module")
def generated():
    m = ta.Mesh.box((3, 3, 3), (0, 0, 0), (1, 1, 1), jitter=0.2, seed=3)
    coords, tets, cen, vol = m.geometry()
    return m, coords, tets, cen, vol


def test_vol_roundtrip_matches_the_generated_mesh(generated, tmp_path):
    m, coords, tets, cen, vol = generated
    mat = np.where(cen[:, 0] < 0.5, 1, 2)
    flip = np.arange(len(tets)) % 2 == 1
    write_vol(tmp_path / "m.vol", coords, tets, mat, flip)
    write_cfg(tmp_path / "m.cfg")
    cfg = ta.Config(str(tmp_path / "m.cfg"))
    r = ta.Mesh.load_vol(tmp_path / "m.vol", cfg)
    c2, t2, cen2, vol2 = r.geometry()
    assert r.nP == m.nP and r.nT == m.nT and r.nF == m.nF
    assert np.array_equal(c2, coords)
    assert np.array_equal(np.sort(t2, axis=1), np.sort(tets, axis=1))
    assert np.allclose(vol2, vol, rtol=1e-13, atol=0)
    assert np.array_equal(r.regions(), mat)
    kappa = np.where(mat == 1, 11.0, 0.25)
    Ip = np.where(mat == 1, 1.0, 0.0)
    a = r.arrays()
    assert np.allclose(a["mesh"]["kappa_volume"], kappa * vol, rtol=1e-13)
    assert np.array_equal(a["mesh"]["I_p"], Ip)
    # the same operator as the generated mesh with the same fields
    m.set_fields(kappa=kappa, Ip=Ip)
    ang = ta.Angular(1)
    s1, s2 = ta.Stencil(m.view, ang.view), ta.Stencil(r.view, ang.view)
    assert np.array_equal(s1.perm, s2.perm) and np.array_equal(s1.nbr, s2.nbr)
    assert np.abs(s1.coef - s2.coef).max() <= 1e-13 * np.abs(s1.coef).max()
    assert np.abs(s1.bcoef - s2.bcoef).max() <= 1e-13 * np.abs(s1.bcoef).max()
    assert np.abs(s1.rhs - s2.rhs).max() <= 1e-13 * np.abs(s1.rhs).max()


def test_vol_without_config_and_errors(generated, tmp_path):
    m, coords, tets, cen, vol = generated
    write_vol(tmp_path / "a.vol", coords, tets, np.full(len(tets), 3), np.zeros(len(tets), bool))
    r = ta.Mesh.load_vol(tmp_path / "a.vol")
    a = r.arrays()
    assert np.allclose(a["mesh"]["kappa_volume"], vol) and np.all(a["mesh"]["I_p"] == 0)
    write_cfg(tmp_path / "m.cfg")
    with pytest.raises(ta.TransferError, match="region"):
        ta.Mesh.load_vol(tmp_path / "a.vol", ta.Config(str(tmp_path / "m.cfg")))
    bad = tmp_path / "b.vol"
    write_vol(bad, coords, np.where(tets == 0, len(coords), tets), np.ones(len(tets), int), np.zeros(len(tets), bool))
    with pytest.raises(ta.TransferError, match="missing point"):
        ta.Mesh.load_vol(bad)
    (tmp_path / "c.vol").write_text("mesh3d\npoints\n1\n0 0 0\n")
    with pytest.raises(ta.TransferError, match="no points or no volume"):
        ta.Mesh.load_vol(tmp_path / "c.vol")
    with pytest.raises(ta.TransferError, match="cannot open"):
        ta.Mesh.load_vol(tmp_path / "missing.vol")


def test_vtk_output_reads_back(generated, tmp_path):
    m, coords, tets, cen, vol = generated
    rng = np.random.default_rng(2)
    U, aU = rng.standard_normal(m.nP), rng.standard_normal(m.nP)
    T = rng.standard_normal((9, m.nP))
    W = rng.standard_normal((3, m.nT))
    m.write_vtk(tmp_path / "s.vtk", U=U, T=T, W=W, aU=aU)
    lines = (tmp_path / "s.vtk").read_text().split("\n")
    assert lines[0].startswith("# vtk DataFile") and lines[3] == "DATASET UNSTRUCTURED_GRID"
    i = lines.index("POINTS %d double" % m.nP)
    assert np.allclose(np.array([l.split() for l in lines[i + 1:i + 1 + m.nP]], float), coords, rtol=0, atol=0)
    i = lines.index("CELLS %d %d" % (m.nT, 5 * m.nT))
    cells = np.array([l.split() for l in lines[i + 1:i + 1 + m.nT]], int)
    assert np.all(cells[:, 0] == 4) and np.array_equal(cells[:, 1:], tets)
    i = lines.index("CELL_TYPES %d" % m.nT)
    assert all(l == "10" for l in lines[i + 1:i + 1 + m.nT])
    i = lines.index("VECTORS W double")
    assert np.array_equal(np.array([l.split() for l in lines[i + 1:i + 1 + m.nT]], float), W.T)
    i = lines.index("SCALARS U double 1")
    assert np.array_equal(np.array(lines[i + 2:i + 2 + m.nP], float), U)
    i = lines.index("SCALARS aU double 1")
    assert np.array_equal(np.array(lines[i + 2:i + 2 + m.nP], float), aU)
    i = lines.index("TENSORS T double")
    got = np.array([l.split() for l in lines[i + 1:i + 1 + 3 * m.nP]], float).reshape(m.nP, 3, 3)
    assert np.array_equal(got, T.reshape(3, 3, m.nP).transpose(2, 0, 1))
    assert "SCALARS aW" not in "\n".join(lines)
