"""CPU FP64 oracle of uranix/Transfer's hot path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
See oracle.c for the reference file:line each function restates.

Parity status of each oracle function (what pins it):
  or_angular        reference AngularData via oracle/_ref (1e-13), plus
                    orthonormality of the harmonics, sum_a Omega^aa = I,
                    <|Omega_n|> = 1/2, and completeness of the scattering form
  operator (lhs,    brute-force weak-form assembly on tiny meshes (tests),
   diag, rhs)       symmetry, constant-field and energy-balance identities
  PCG / invprec     residual of the returned solution, textbook CG identities
  E_n / step        scipy.special.expn
  moments / flux    closed forms in the infinite-medium limit
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle: %s missing; run `make -C oracle`" % path)
        _lib = C.CDLL(path)
        _lib.or_expn.restype = C.c_double
        _lib.or_expn.argtypes = [C.c_int, C.c_double]
        for n in ("or_step_U", "or_step_W", "or_step_T"):
            getattr(_lib, n).restype = C.c_double
            getattr(_lib, n).argtypes = [C.c_double]
        _lib.or_slab_U.restype = C.c_double
        _lib.or_slab_U.argtypes = [C.c_double, C.c_double]
        _lib.or_norm.restype = C.c_double
        _lib.or_dot.restype = C.c_double
        _lib.or_angular.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int] + [C.c_void_p] * 6
    return _lib


def ref():
    """The reference's own host code compiled from its sources (oracle/_ref),
    or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref_host.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        _ref.ref_xE1.restype = C.c_double
        _ref.ref_xE1.argtypes = [C.c_double]
    return _ref


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Problem(C.Structure):
    _fields_ = [("nP", C.c_uint32), ("nT", C.c_uint32), ("nF", C.c_uint32),
                ("tetstart", C.c_void_p), ("tetidx", C.c_void_p), ("tetpos", C.c_void_p),
                ("mesh", C.c_void_p), ("facestart", C.c_void_p), ("faceidx", C.c_void_p),
                ("facepos", C.c_void_p), ("bnd", C.c_void_p), ("slm", C.c_uint32),
                ("aslm", C.c_uint32), ("omega", C.c_void_p), ("omega_pos", C.c_void_p),
                ("Ox", C.c_void_p), ("Oy", C.c_void_p), ("Oz", C.c_void_p),
                ("massfac", C.c_void_p), ("nthreads", C.c_int), ("face_kind", C.c_void_p)]


class PcgParams(C.Structure):
    _fields_ = [("use_prec", C.c_int), ("maxit", C.c_int), ("tol_abs", C.c_double),
                ("tol_rel", C.c_double), ("inner_maxit", C.c_int), ("inner_tol", C.c_double),
                ("record", C.c_int)]


class PcgStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("preck", C.c_int), ("rnorm", C.c_double), ("bnorm", C.c_double)]


def slm_of(maxk):
    return (1 + maxk) * (1 + 2 * maxk)


def aslm_of(slm):
    return (slm + 7) & ~7


def angular(maxk, albedo=0.0, g=0.0, hg_lmax=0):
    """Host (unpadded) angular arrays by exact integration."""
    s = slm_of(maxk)
    om, pos = np.zeros(3 * s * s), np.zeros(3 * s * s, np.uint32)
    Ox, Oy, Oz, mf = np.zeros(s * s), np.zeros(s * s), np.zeros(s * s), np.zeros(s)
    lib().or_angular(maxk, albedo, g, hg_lmax, _p(om), _p(pos), _p(Ox), _p(Oy), _p(Oz), _p(mf))
    return dict(omega=om, omega_pos=pos, Ox=Ox, Oy=Oy, Oz=Oz, massfac=mf, slm=s)


def ref_angular(maxk):
    r = ref()
    if r is None:
        return None
    s = slm_of(maxk)
    om, pos = np.zeros(3 * s * s), np.zeros(3 * s * s, np.uint32)
    Ox, Oy, Oz = np.zeros(s * s), np.zeros(s * s), np.zeros(s * s)
    r.ref_angular(maxk, _p(om), _p(pos), _p(Ox), _p(Oy), _p(Oz))
    return dict(omega=om, omega_pos=pos, Ox=Ox, Oy=Oy, Oz=Oz, slm=s)


def sh_gram(lmax):
    n = (lmax + 1) ** 2
    g = np.zeros(n * n)
    lib().or_sh_gram(lmax, _p(g))
    return g.reshape(n, n)


def sh_eval(lmax, x, y, z):
    out = np.zeros((lmax + 1) ** 2)
    lib().or_sh_eval(lmax, C.c_double(x), C.c_double(y), C.c_double(z), _p(out))
    return out


class OracleProblem:
    """The reference's device data (DeviceMeshData/DeviceAngularData) on the
    host, from MeshData arrays and host angular arrays."""

    def __init__(self, arrays, ang, nthreads=0):
        s = int(ang["slm"])
        a = aslm_of(s)
        self.slm, self.aslm = s, a
        self.nP = len(arrays["tetstart"]) - 1
        self.nT, self.nF = len(arrays["mesh"]), len(arrays["bnd"])
        self._ang = ang
        self.omega_d = np.zeros(3 * a * a)
        self.pos_d = np.zeros(3 * a * a, np.uint32)
        self.Ox_d, self.Oy_d, self.Oz_d = np.zeros(a * a), np.zeros(a * a), np.zeros(a * a)
        lib().or_pad_angular(s, _p(ang["omega"]), _p(ang["omega_pos"]), _p(ang["Ox"]),
                             _p(ang["Oy"]), _p(ang["Oz"]), _p(self.omega_d), _p(self.pos_d),
                             _p(self.Ox_d), _p(self.Oy_d), _p(self.Oz_d))
        self.massfac = ang.get("massfac")
        self._arr = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        ar = self._arr
        self.pr = Problem(self.nP, self.nT, self.nF, _p(ar["tetstart"]), _p(ar["tetidx"]),
                          _p(ar["tetpos"]), _p(ar["mesh"]), _p(ar["facestart"]), _p(ar["faceidx"]),
                          _p(ar["facepos"]), _p(ar["bnd"]), s, a, _p(self.omega_d), _p(self.pos_d),
                          _p(self.Ox_d), _p(self.Oy_d), _p(self.Oz_d),
                          _p(self.massfac) if self.massfac is not None else None, nthreads,
                          _p(ar["face_kind"]) if "face_kind" in ar and len(ar["face_kind"]) else None)

    def N(self):
        return self.nP * self.aslm

    def zeros(self):
        return np.zeros(self.N())

    def _unary(self, name, f):
        r = self.zeros()
        getattr(lib(), name)(C.byref(self.pr), _p(np.ascontiguousarray(f, np.float64)), _p(r))
        return r

    def lhs(self, f):
        return self._unary("or_compute_lhs", f)

    def lhs_diag(self, f):
        return self._unary("or_compute_lhs_diag", f)

    def volume(self, f):
        return self._unary("or_volume_part", f)

    def rhs(self):
        b = self.zeros()
        lib().or_compute_rhs(C.byref(self.pr), _p(b))
        return b

    def norm(self, x):
        return lib().or_norm(C.byref(self.pr), _p(x))

    def dot(self, x, y):
        return lib().or_dot(C.byref(self.pr), _p(x), _p(y))

    def pcg(self, use_prec=False, maxit=1000, tol_abs=1e-4, tol_rel=0.0, inner_maxit=1000,
            inner_tol=1e-8, record=0):
        prm = PcgParams(int(use_prec), maxit, tol_abs, tol_rel, inner_maxit, inner_tol, record)
        st = PcgStats()
        f = self.zeros()
        hist = np.zeros(3 * max(record, 1))
        lib().or_pcg(C.byref(self.pr), C.byref(prm), _p(f), C.byref(st), _p(hist))
        return f, st, hist[:3 * record].reshape(-1, 3)

    def moments(self, f):
        U, T = np.zeros(self.nP), np.zeros(9 * self.nP)
        a = self._ang
        lib().or_compute_moments(C.byref(self.pr), _p(a["omega"]), _p(a["omega_pos"]),
                                 _p(np.ascontiguousarray(f)), _p(U), _p(T))
        return U, T.reshape(9, self.nP)

    def flux(self, T):
        W = np.zeros(3 * self.nT)
        lib().or_compute_flux(C.byref(self.pr), _p(np.ascontiguousarray(T)), _p(W))
        return W.reshape(3, self.nT)

    def boundary_flux(self, f):
        """Outward flux density per boundary face, face-averaged."""
        F = np.zeros(self.nF)
        lib().or_boundary_flux(C.byref(self.pr), _p(np.ascontiguousarray(f)), _p(F))
        return F


def expn(n, x):
    return lib().or_expn(int(n), float(x))


def step_U(x):
    return lib().or_step_U(float(x))


def step_W(x):
    return lib().or_step_W(float(x))


def step_T(x):
    return lib().or_step_T(float(x))


def slab_U(x, L):
    return lib().or_slab_U(float(x), float(L))
