"""Workload builders: the meshes, per-tet fields and angular data of the
benchmark configurations, from the library's own generators.

    slab      [0,1]x[0,1]x[0,4], 12x12x48 cubes (41,472 tets), interior nodes
              jittered by +-0.15 cell widths, kappa = 1, isotropic scattering
              albedo 0.5, source kappa*I_p = 1, K = 1 (26-point Lebedev order)
    ball64    cubes of a 64^3 grid on [-1,1]^3 with centroid r <= 1
              (824,256 tets), kappa = 50 exp(-r/0.2), I_p = (1 - 0.8 r)^4,
              no scattering, K = 3 (110-point order)
    ball96    96^3 grid, same ball, kappa = 1 + 9 U[0,1) per tet, albedo 0.9,
              Henyey-Greenstein g = 0.6 truncated after P_7, I_p = 1, K = 5
    cube32    32^3 cubes on [0,1]^3 (196,608 tets), kappa = 0.1 + 8 Gaussian
              blobs (amplitude log-U[1,100], width U[0.05,0.3]), albedo
              U[0,0.99], I_p = 1, K = 3
    batch     256 slabs n x n x 4n (n ~ U{8..32} from the seed), per-tet
              ln kappa ~ N(0, 1), albedo ~ U[0, 0.95] per problem, I_p = 1,
              K = 2 (50-point order); `count` and `nrange` shrink it for tests
    sphere    the two-region sphere of the reference's soltype-1 check
              (main.cpp:190-197; radiation.cfg.example: kappa 11 and I_p 1
              in the core r < 0.3, kappa 1 and I_p 0 outside), radius 1:
              the n^3 Kuhn cubes of [-1,1]^3 mapped onto the ball by
              x -> x |x|_inf / |x|_2 (nested cube surfaces onto nested
              spheres), so every boundary face is a chord face of the unit
              sphere with its normal off the axes; regions by tet centroid,
              no scattering
    milne     the step problem of milne.geo as a 1-D slab: [-L,L] x [-h,h]^2,
              kappa = 1, I_p = 1 for x < 0 and 0 for x > 0, vacuum at
              x = +-L, specular reflection on the four side planes (so the
              solution is the exact 1-D one, U(x) = 2 pi (E2 ... ) terms)
The Lebedev N-point orders map to the harmonic order K = (degree(N) - 2) // 4.
"""
import numpy as np

from .api import Angular, Mesh, rng_uniform


class Problem:
    def __init__(self, name, mesh, ang, albedo=0.0, g=0.0, hg_lmax=0):
        self.name, self.mesh, self.ang = name, mesh, ang
        self.albedo, self.g, self.hg_lmax = albedo, g, hg_lmax


def slab(n=12, seed=1, maxk=1, albedo=0.5, jitter=0.15):
    m = Mesh.box((n, n, 4 * n), (0, 0, 0), (1, 1, 4), jitter=jitter, seed=seed)
    m.set_fields(kappa=m.field("const", [1.0]), Ip=m.field("const", [1.0]))
    return Problem("slab%d" % n, m, Angular(maxk, albedo, 0.0, 0), albedo)


def ball(n=64, maxk=3, seed=2, kind="planck"):
    m = Mesh.box((n, n, n), (-1, -1, -1), (1, 1, 1), ball_radius=1.0, ball_center=(0, 0, 0))
    if kind == "planck":
        m.set_fields(kappa=m.field("radial_exp", [50.0, 0.2, 0, 0, 0]),
                     Ip=m.field("planck", [0.8, 0, 0, 0, 0]))
        return Problem("ball%d" % n, m, Angular(maxk))
    m.set_fields(kappa=m.field("uniform", [1.0, 9.0], seed=seed, stream=3), Ip=m.field("const", [1.0]))
    return Problem("ball%d-hg" % n, m, Angular(maxk, 0.9, 0.6, 7), 0.9, 0.6, 7)


def cube(n=32, maxk=3, seed=5, albedo=None):
    m = Mesh.box((n, n, n), (0, 0, 0), (1, 1, 1))
    k = m.field("blobs", [0.1, 8, 1.0, 100.0, 0.05, 0.3], seed=seed, stream=4)
    m.set_fields(kappa=k, Ip=m.field("const", [1.0]))
    if albedo is None:
        albedo = 0.99 * rng_uniform(seed, 6, 1)[0]
    return Problem("cube%d" % n, m, Angular(maxk, albedo, 0.0, 0), albedo)


def sphere(n=32, maxk=3, core=0.3, kappa=(11.0, 1.0), Ip=(1.0, 0.0)):
    box = Mesh.box((n, n, n), (-1, -1, -1), (1, 1, 1))
    xyz, tets, _, _ = box.geometry()
    r2 = np.linalg.norm(xyz, axis=1)
    rinf = np.abs(xyz).max(axis=1)
    xyz = xyz * np.where(r2 > 0, rinf / np.where(r2 > 0, r2, 1.0), 1.0)[:, None]
    rc = np.linalg.norm(xyz[tets.astype(np.int64)].mean(axis=1), axis=1)
    k = np.where(rc < core, kappa[0], kappa[1])
    ip = np.where(rc < core, Ip[0], Ip[1])
    return Problem("sphere%d" % n, Mesh.from_arrays(xyz, tets, k, ip), Angular(maxk))


def milne(n=32, maxk=3, L=1.0, ncross=2):
    """Step problem on a slab: 2n cells across [-L, L], ncross cells of the
    same width across each side direction."""
    h = L / n * ncross / 2.0
    m = Mesh.box((2 * n, ncross, ncross), (-L, -h, -h), (L, h, h))
    m.set_fields(kappa=m.field("const", [1.0]), Ip=m.field("step", [0, 0.0, 1.0, 0.0]))
    for axis in (1, 2):
        for value in (-h, h):
            m.set_face_kind_plane(axis, value, 1)
    return Problem("milne%d" % n, m, Angular(maxk))


class BatchProblem:
    def __init__(self, name, meshes, ang, albedo):
        self.name, self.meshes, self.ang, self.albedo = name, meshes, ang, albedo


def batch(count=256, nrange=(8, 32), maxk=2, seed=7, progress=None):
    """Independent slab problems: problem p is n x n x 4n cubes on
    [0,1]^2 x [0,4] (n from the seed), ln kappa ~ N(0, 1) per tet.
    progress(p) is called after each problem's mesh when given."""
    u = rng_uniform(seed, 20, 2 * count)
    lo, hi = nrange
    meshes, albedo = [], []
    for p in range(count):
        n = lo + min(int(u[2 * p] * (hi - lo + 1)), hi - lo)
        m = Mesh.box((n, n, 4 * n), (0, 0, 0), (1, 1, 4))
        m.set_fields(kappa=m.field("lognormal", [0.0, 1.0], seed=seed, stream=1000 + p),
                     Ip=m.field("const", [1.0]))
        meshes.append(m)
        albedo.append(0.95 * u[2 * p + 1])
        if progress:
            progress(p)
    return BatchProblem("batch%d" % count, meshes, Angular(maxk), albedo)
