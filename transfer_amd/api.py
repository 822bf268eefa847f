"""ctypes bindings of libtransfer_amd.so (include/transfer.h) and a Python
mirror of the reference's CudaContext interface (CudaContext.h:29-47).

The library is required: importing this module raises if it is missing, and
there is no Python or CPU fallback for anything that runs on the device.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TR_LIBRARY: another build of the library (same ABI), e.g. a variant under test
LIB_PATH = os.environ.get("TR_LIBRARY") or os.path.join(_HERE, "libtransfer_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("transfer_amd: %s is missing; run `python transfer_amd/build.py`" % LIB_PATH)
lib = C.CDLL(LIB_PATH)

u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p

TR_CTX_GENERAL_NORMALS = 1
TR_SOLVER_REFERENCE = 0
TR_SOLVER_PCG_FP64 = 1
TR_SOLVER_MIXED = 2

FIELD = dict(const=0, radial_exp=1, uniform=2, lognormal=3, blobs=4, planck=5, step=6)

TET_DTYPE = np.dtype([("p", "<u4", 4), ("kappa_volume", "<f8"), ("I_p", "<f8"), ("s", "<f8", (4, 3))])
FACE_DTYPE = np.dtype([("p", "<u4", 4), ("s", "<f8", 3), ("_padd", "<f8", 3)])
assert TET_DTYPE.itemsize == 128 and FACE_DTYPE.itemsize == 64


class MeshSpec(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("jitter", C.c_double), ("ball_radius", C.c_double),
                ("ball_center", C.c_double * 3), ("seed", C.c_uint64)]


class MeshView(C.Structure):
    _fields_ = [("nP", C.c_uint32), ("nF", C.c_uint32), ("nT", C.c_uint32),
                ("tetstart", u32p), ("tetidx", u32p), ("tetpos", u32p), ("mesh", vp),
                ("facestart", u32p), ("faceidx", u32p), ("facepos", u32p), ("bnd", vp),
                ("coords", f64p), ("face_kind", u8p)]


class AngularView(C.Structure):
    _fields_ = [("slm", C.c_uint32), ("omega", f64p), ("omega_pos", u32p),
                ("Ox", f64p), ("Oy", f64p), ("Oz", f64p), ("massfac", f64p)]


class SolveParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("use_prec", C.c_int), ("maxit", C.c_int),
                ("tol_abs", C.c_double), ("tol_rel", C.c_double),
                ("inner_maxit", C.c_int), ("inner_tol", C.c_double), ("record", C.c_int), ("x0", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("preck", C.c_int), ("rnorm", C.c_double),
                ("bnorm", C.c_double), ("true_rel", C.c_double), ("seconds", C.c_double),
                ("status", C.c_int), ("fallback", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("inner_iters", C.c_int), ("true_rel", C.c_double), ("status", C.c_int)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("tr_last_error", C.c_char_p)
_sig("tr_angular_create", C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(vp))
_sig("tr_angular_destroy", C.c_int, vp)
_sig("tr_angular_get_view", C.c_int, vp, C.POINTER(AngularView))
_sig("tr_mesh_create_box", C.c_int, C.POINTER(MeshSpec), C.POINTER(vp))
_sig("tr_mesh_create", C.c_int, C.c_uint32, f64p, C.c_uint32, u32p, f64p, f64p, C.POINTER(vp))
_sig("tr_mesh_destroy", C.c_int, vp)
_sig("tr_mesh_get_view", C.c_int, vp, C.POINTER(MeshView))
_sig("tr_mesh_get_geometry", C.c_int, vp, f64p, u32p, f64p, f64p)
_sig("tr_mesh_set_fields", C.c_int, vp, f64p, f64p)
_sig("tr_mesh_set_face_kind_plane", C.c_int, vp, C.c_int, C.c_double, C.c_double, C.c_int)
_sig("tr_field_eval", C.c_int, vp, C.c_int, f64p, C.c_uint64, C.c_uint64, f64p)
_sig("tr_rng_uniform", C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, f64p)
_sig("tr_config_load", C.c_int, C.c_char_p, C.POINTER(vp))
_sig("tr_config_destroy", C.c_int, vp)
_sig("tr_config_ints", C.c_int, vp, C.POINTER(C.c_int))
_sig("tr_config_region", C.c_int, vp, C.c_int, f64p, f64p)
_sig("tr_config_strings", C.c_int, vp, C.c_char_p, C.c_char_p, C.c_size_t)
_sig("tr_ctx_create", C.c_int, C.c_int, C.POINTER(MeshView), C.POINTER(AngularView), C.c_int, C.POINTER(vp))
_sig("tr_ctx_destroy", C.c_int, vp)
_sig("tr_ctx_N", C.c_uint32, vp)
_sig("tr_ctx_slm", C.c_uint32, vp)
_sig("tr_ctx_aslm", C.c_uint32, vp)
_sig("tr_ctx_nP", C.c_uint32, vp)
_sig("tr_ctx_permutation", C.c_int, vp, u32p)
_sig("tr_ctx_autotune", C.c_int, vp, C.c_int, C.c_char_p, C.c_size_t)
_sig("tr_ctx_schedule", C.c_int, vp, C.c_int, C.POINTER(C.c_int))
_sig("tr_ctx_set_schedule", C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_int)
_sig("tr_ctx_set_angular_precision", C.c_int, vp, C.c_int)
_sig("tr_ctx_set_refinement", C.c_int, vp, C.c_double, C.c_int)
_sig("tr_ctx_set_residual_replacement", C.c_int, vp, C.c_int)
_sig("tr_dot_noalloc", C.c_int, vp, vp, vp, vp)
_sig("tr_abs_normal_matrix", C.c_int, C.c_int, f64p, f64p)
_sig("tr_abs_normal_column", C.c_int, C.c_int, f64p, f64p)
_sig("tr_normal_rotation", C.c_int, C.c_int, f64p, f64p, u32p)
_sig("tr_device_alloc", vp, vp, C.c_size_t)
_sig("tr_alloc_vector", vp, vp)
_sig("tr_device_free", C.c_int, vp, vp)
_sig("tr_copy_to_dev", C.c_int, vp, vp, vp, C.c_size_t)
_sig("tr_copy_to_host", C.c_int, vp, vp, vp, C.c_size_t)
_sig("tr_get_host_small", C.c_int, vp, vp, f64p)
for _n in ("tr_compute_lhs", "tr_compute_lhs_diag"):
    _sig(_n, C.c_int, vp, vp, vp)
_sig("tr_compute_rhs", C.c_int, vp, vp)
_sig("tr_compute_lhs_prec", C.c_int, vp, C.c_int, vp, vp)
_sig("tr_compute_moments", C.c_int, vp, vp, vp, vp)
_sig("tr_ctx_set_multilevel", C.c_int, vp, C.c_uint32, C.c_uint32, C.c_int)
_sig("tr_mesh_load_vol", C.c_int, C.c_char_p, vp, C.POINTER(vp))
_sig("tr_mesh_get_regions", C.c_int, vp, u32p)
_sig("tr_write_vtk", C.c_int, vp, C.c_char_p, f64p, f64p, f64p, f64p, f64p, f64p)
_sig("tr_dump_slae", C.c_int, vp, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double))
_sig("tr_eig_estimate", C.c_int, vp, C.c_int, C.c_int, C.c_double, f64p)
_sig("tr_ctx_set_chebyshev", C.c_int, vp, C.c_int, C.c_double, C.c_int)
_sig("tr_compute_flux", C.c_int, vp, C.POINTER(MeshView), vp, vp)
_sig("tr_compute_boundary_flux", C.c_int, vp, C.POINTER(MeshView), vp, vp)
_sig("tr_expn", C.c_double, C.c_int, C.c_double)
_sig("tr_mlp_features", C.c_int, C.POINTER(MeshView), C.c_double, f32p)
_sig("tr_mlp_create", C.c_int, C.c_int, f32p, C.c_uint64, C.POINTER(vp))
_sig("tr_mlp_destroy", C.c_int, vp)
_sig("tr_mlp_set_norm", C.c_int, vp, f32p, f32p)
_sig("tr_mlp_alloc", C.c_int, vp, C.c_size_t, C.POINTER(vp))
_sig("tr_mlp_copy", C.c_int, vp, vp, vp, C.c_size_t, C.c_int)
_sig("tr_mlp_get_params", C.c_int, vp, f32p)
_sig("tr_mlp_predict", C.c_int, vp, vp, C.c_uint32, vp)
_sig("tr_mlp_grad", C.c_int, vp, vp, vp, C.c_uint32, f32p, f32p)
_sig("tr_mlp_train", C.c_int, vp, vp, vp, C.c_uint32, C.c_int, C.c_uint32, C.c_float, C.c_uint64, f32p)
_sig("tr_ctx_scalar_guess", C.c_int, vp, vp, C.c_double, vp)
_sig("tr_ctx_mlp_features", C.c_int, vp, C.c_double, vp, C.POINTER(C.c_int))
_sig("tr_ctx_learned_guess", C.c_int, vp, vp, C.c_double, C.c_double, vp)
_sig("tr_batch_create", C.c_int, C.c_int, C.c_int, C.POINTER(MeshView), C.POINTER(AngularView), f64p, C.c_int,
     C.POINTER(vp))
_sig("tr_batch_info", C.c_int, vp, u32p, u32p)
_sig("tr_batch_solve", C.c_int, vp, C.POINTER(SolveParams), vp, C.POINTER(SolveStats), C.POINTER(BatchStats))
_sig("tr_step_analytic", C.c_int, C.c_double, f64p)
_sig("tr_slab_analytic_U", C.c_double, C.c_double, C.c_double)
_sig("tr_scale", C.c_int, vp, vp, C.c_double)
_sig("tr_add_prod", C.c_int, vp, vp, vp, C.c_double)
_sig("tr_mul_add", C.c_int, vp, vp, C.c_double, vp)
_sig("tr_mul_add_prod", C.c_int, vp, vp, C.c_double, vp, C.c_double)
_sig("tr_norm", C.c_double, vp, vp)
_sig("tr_dot", C.c_double, vp, vp, vp)
_sig("tr_synchronize", C.c_int, vp)
_sig("tr_solve", C.c_int, vp, C.POINTER(SolveParams), vp, C.POINTER(SolveStats), f64p)
_sig("tr_bench_lhs", C.c_double, vp, C.c_int)
_sig("tr_bench_operator", C.c_int, vp, C.c_int, C.c_int, C.c_int, f64p)
_sig("tr_ctx_kernel_name", C.c_int, vp, C.c_int, C.c_char_p, C.c_size_t)
_sig("tr_bench_copy", C.c_int, vp, C.c_size_t, C.c_int, f64p)
_sig("tr_stencil_create", C.c_int, C.POINTER(MeshView), C.POINTER(AngularView), C.POINTER(vp))
_sig("tr_stencil_destroy", C.c_int, vp)
_sig("tr_stencil_info", C.c_int, vp, u32p)
_sig("tr_stencil_mask", C.c_int, vp, u32p, u32p)
_sig("tr_stencil_coarsen", C.c_int, vp, C.c_uint32, C.POINTER(vp), u32p)
_sig("tr_stencil_export", C.c_int, vp, u32p, u32p, f64p, u32p, u32p, f64p, f64p, f64p, f64p, f64p)


TR_OK, TR_ERR_ARG, TR_ERR_HIP, TR_ERR_STATE, TR_ERR_NOCONV = 0, -1, -2, -3, -4


class TransferError(RuntimeError):
    """A failed C-ABI call: `code` is its TR_ERR_* status."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


def check(rc):
    if rc != 0:
        raise TransferError("transfer_amd error %d: %s" % (rc, lib.tr_last_error().decode()), rc)
    return rc


def ptr(a, t=f64p):
    return None if a is None else a.ctypes.data_as(t)


def last_error():
    return lib.tr_last_error().decode()


# ---------------------------------------------------------------------------
class Angular:
    """AngularData(maxk) (AngularData.cpp:9-68) plus the scattering extension."""

    def __init__(self, maxk, albedo=0.0, g=0.0, hg_lmax=0):
        self._h = vp()
        check(lib.tr_angular_create(int(maxk), float(albedo), float(g), int(hg_lmax), C.byref(self._h)))
        self.maxk = maxk
        self.view = AngularView()
        check(lib.tr_angular_get_view(self._h, C.byref(self.view)))
        s = self.slm = int(self.view.slm)
        self.omega = np.ctypeslib.as_array(self.view.omega, (3 * s * s,)).copy()
        self.omega_pos = np.ctypeslib.as_array(self.view.omega_pos, (3 * s * s,)).copy()
        self.Ox = np.ctypeslib.as_array(self.view.Ox, (s * s,)).copy()
        self.Oy = np.ctypeslib.as_array(self.view.Oy, (s * s,)).copy()
        self.Oz = np.ctypeslib.as_array(self.view.Oz, (s * s,)).copy()
        self.massfac = np.ctypeslib.as_array(self.view.massfac, (s,)).copy()

    @staticmethod
    def view_of(omega, omega_pos, Ox, Oy, Oz, massfac=None):
        """An AngularView over caller-held arrays (e.g. the reference's own)."""
        v = AngularView()
        v.slm = int(round(np.sqrt(len(Ox))))
        v.omega, v.omega_pos = ptr(omega), ptr(omega_pos, u32p)
        v.Ox, v.Oy, v.Oz = ptr(Ox), ptr(Oy), ptr(Oz)
        v.massfac = ptr(massfac)
        v._keep = (omega, omega_pos, Ox, Oy, Oz, massfac)
        return v

    def __del__(self):
        if getattr(self, "_h", None):
            lib.tr_angular_destroy(self._h)
            self._h = None


class Mesh:
    """MeshData (MeshData.h:13-32) built from a generator or raw geometry."""

    def __init__(self, handle):
        self._h = handle
        self.view = MeshView()
        check(lib.tr_mesh_get_view(self._h, C.byref(self.view)))
        self.nP, self.nT, self.nF = int(self.view.nP), int(self.view.nT), int(self.view.nF)

    @classmethod
    def box(cls, n, lo, hi, jitter=0.0, ball_radius=0.0, ball_center=(0, 0, 0), seed=0):
        sp = MeshSpec()
        sp.nx, sp.ny, sp.nz = [int(x) for x in n]
        sp.lo[:] = list(map(float, lo))
        sp.hi[:] = list(map(float, hi))
        sp.jitter, sp.ball_radius = float(jitter), float(ball_radius)
        sp.ball_center[:] = list(map(float, ball_center))
        sp.seed = int(seed)
        h = vp()
        check(lib.tr_mesh_create_box(C.byref(sp), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, coords, tets, kappa=None, Ip=None):
        coords = np.ascontiguousarray(coords, np.float64)
        tets = np.ascontiguousarray(tets, np.uint32)
        kappa = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        Ip = None if Ip is None else np.ascontiguousarray(Ip, np.float64)
        h = vp()
        check(lib.tr_mesh_create(len(coords), ptr(coords), len(tets), ptr(tets, u32p),
                                 ptr(kappa), ptr(Ip), C.byref(h)))
        return cls(h)

    @classmethod
    def load_vol(cls, path, config=None):
        """A Netgen .vol mesh (tr_mesh_load_vol); kappa and I_p per tet from
        the config's [region N] sections when a Config is given."""
        h = vp()
        check(lib.tr_mesh_load_vol(str(path).encode(), None if config is None else config._h, C.byref(h)))
        return cls(h)

    def regions(self):
        out = np.empty(self.nT, np.uint32)
        check(lib.tr_mesh_get_regions(self._h, ptr(out, u32p)))
        return out

    def write_vtk(self, path, U=None, T=None, W=None, aU=None, aW=None, aT=None):
        """Legacy VTK of the solution fields (tr_write_vtk): W [3][nT], U [nP],
        T [9][nP], aU/aW/aT [nP]; host arrays, any may be None."""
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (U, T, W, aU, aW, aT)]
        check(lib.tr_write_vtk(self._h, str(path).encode(), *[ptr(a) for a in arrs]))

    def refresh_view(self):
        check(lib.tr_mesh_get_view(self._h, C.byref(self.view)))

    def set_fields(self, kappa=None, Ip=None):
        k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        i = None if Ip is None else np.ascontiguousarray(Ip, np.float64)
        check(lib.tr_mesh_set_fields(self._h, ptr(k), ptr(i)))
        self.refresh_view()

    def set_face_kind_plane(self, axis, value, kind, tol=1e-9):
        check(lib.tr_mesh_set_face_kind_plane(self._h, axis, value, tol, kind))

    def field(self, kind, params, seed=0, stream=0):
        p = np.zeros(8)
        p[:len(params)] = params
        out = np.empty(self.nT)
        check(lib.tr_field_eval(self._h, FIELD[kind], ptr(p), int(seed), int(stream), ptr(out)))
        return out

    def geometry(self):
        coords = np.empty((self.nP, 3))
        tets = np.empty((self.nT, 4), np.uint32)
        cen = np.empty((self.nT, 3))
        vol = np.empty(self.nT)
        check(lib.tr_mesh_get_geometry(self._h, ptr(coords), ptr(tets, u32p), ptr(cen), ptr(vol)))
        return coords, tets, cen, vol

    def arrays(self):
        """Copies of the MeshData arrays (the reference's layouts)."""
        v = self.view
        nP, nT, nF = self.nP, self.nT, self.nF
        tc = int(np.ctypeslib.as_array(v.tetstart, (nP + 1,))[-1])
        fc = int(np.ctypeslib.as_array(v.facestart, (nP + 1,))[-1])
        a = lambda p_, n: np.ctypeslib.as_array(p_, (n,)).copy()
        mesh = np.frombuffer((C.c_char * (128 * nT)).from_address(v.mesh), TET_DTYPE).copy()
        bnd = np.frombuffer((C.c_char * (64 * max(nF, 1))).from_address(v.bnd), FACE_DTYPE)[:nF].copy()
        out = dict(tetstart=a(v.tetstart, nP + 1), tetidx=a(v.tetidx, tc), tetpos=a(v.tetpos, tc),
                   mesh=mesh, facestart=a(v.facestart, nP + 1), faceidx=a(v.faceidx, fc),
                   facepos=a(v.facepos, fc), bnd=bnd, coords=a(v.coords, 3 * nP).reshape(nP, 3),
                   face_kind=a(v.face_kind, nF) if nF else np.zeros(0, np.uint8))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib.tr_mesh_destroy(self._h)
            self._h = None


def mesh_view_of(ar):
    """A MeshView over host arrays in the layout of Mesh.arrays() (the
    caller keeps `ar` alive while the view is used)."""
    v = MeshView()
    v.nP, v.nT, v.nF = len(ar["tetstart"]) - 1, len(ar["mesh"]), len(ar["bnd"])
    for k in ("tetstart", "tetidx", "tetpos", "facestart", "faceidx", "facepos"):
        setattr(v, k, ptr(ar[k], u32p))
    v.mesh = ar["mesh"].ctypes.data
    v.bnd = ar["bnd"].ctypes.data
    v.coords = ptr(ar["coords"])
    v.face_kind = ptr(ar["face_kind"], u8p) if len(ar["face_kind"]) else None
    v._keep = ar
    return v


def abs_normal(maxk, n):
    """(O(n) [slm][slm], its column 00, the packed rotation blocks R) of a
    face normal n (host only)."""
    s = (1 + maxk) * (1 + 2 * maxk)
    n = np.ascontiguousarray(n, np.float64)
    O, col = np.zeros(s * s), np.zeros(s)
    check(lib.tr_abs_normal_matrix(int(maxk), ptr(n), ptr(O)))
    check(lib.tr_abs_normal_column(int(maxk), ptr(n), ptr(col)))
    size = C.c_uint32()
    check(lib.tr_normal_rotation(int(maxk), ptr(n), None, C.byref(size)))
    R = np.zeros(size.value)
    check(lib.tr_normal_rotation(int(maxk), ptr(n), ptr(R), None))
    return O.reshape(s, s), col, R


def rng_uniform(seed, stream, n):
    out = np.empty(n)
    check(lib.tr_rng_uniform(int(seed), int(stream), int(n), ptr(out)))
    return out


class Config:
    """Config(cfgfile) (Config.cpp:8-131)."""

    def __init__(self, path):
        self._h = vp()
        check(lib.tr_config_load(path.encode(), C.byref(self._h)))
        v = (C.c_int * 6)()
        check(lib.tr_config_ints(self._h, v))
        self.maxk, self.device, self.dump, self.prec, self.soltype, self.nreg = list(v)
        self.kappa, self.Ip = {}, {}
        for r in range(0, self.nreg + 1):
            k, i = C.c_double(), C.c_double()
            check(lib.tr_config_region(self._h, r, C.byref(k), C.byref(i)))
            self.kappa[r], self.Ip[r] = k.value, i.value
        m, o = C.create_string_buffer(1024), C.create_string_buffer(1024)
        check(lib.tr_config_strings(self._h, m, o, 1024))
        self.meshfile, self.outfile = m.value.decode(), o.value.decode()

    def __del__(self):
        if getattr(self, "_h", None):
            lib.tr_config_destroy(self._h)
            self._h = None


class Stencil:
    """The assembled operator data (introspection for tests)."""

    def __init__(self, mesh_view, ang_view, _handle=None):
        self._h = vp()
        if _handle is not None:
            self._h = _handle
        else:
            check(lib.tr_stencil_create(C.byref(mesh_view), C.byref(ang_view), C.byref(self._h)))
        info = np.zeros(8, np.uint32)
        check(lib.tr_stencil_info(self._h, ptr(info, u32p)))
        nP, slm, NC, W, nB, WB, arow, brow = [int(x) for x in info]
        self.nP, self.slm, self.NC, self.W, self.nB, self.WB = nP, slm, NC, W, nB, WB
        self.perm = np.empty(nP, np.uint32)
        self.nbr = np.empty((W, nP), np.uint32)
        self.coef = np.empty((W, NC, nP))
        self.bvert = np.empty(nB, np.uint32)
        self.bnbr = np.empty((WB, nB), np.uint32)
        self.bcoef = np.empty((WB, 3, nB))
        self.omc = np.empty((NC - 1, slm, slm))
        self.onc = np.empty((3, slm, slm))
        self.massfac = np.empty(slm)
        self.rhs = np.empty(nP)
        cnt = np.zeros(1, np.uint32)
        check(lib.tr_stencil_mask(self._h, ptr(cnt, u32p), None))
        self.mask = np.empty(int(cnt[0]), np.uint32)
        if len(self.mask):
            check(lib.tr_stencil_mask(self._h, ptr(cnt, u32p), ptr(self.mask, u32p)))
        check(lib.tr_stencil_export(self._h, ptr(self.perm, u32p), ptr(self.nbr, u32p), ptr(self.coef),
                                    ptr(self.bvert, u32p), ptr(self.bnbr, u32p), ptr(self.bcoef),
                                    ptr(self.omc), ptr(self.onc), ptr(self.massfac), ptr(self.rhs)))

    def coarsen(self, size=8):
        """(coarse Stencil, aggregate starts) of the multilevel preconditioner."""
        h = vp()
        agg = np.zeros((self.nP + size - 1) // size + 1, np.uint32)
        check(lib.tr_stencil_coarsen(self._h, int(size), C.byref(h), ptr(agg, u32p)))
        return Stencil(None, None, _handle=h), agg

    def __del__(self):
        if getattr(self, "_h", None):
            lib.tr_stencil_destroy(self._h)
            self._h = None


# ---------------------------------------------------------------------------
class HipContext:
    """Mirror of CudaContext (CudaContext.h:29-47) over the C ABI.

    Device vectors are opaque handles (int addresses).  Host arrays use the
    reference layout [nP][aslm] in the caller's vertex order."""

    def __init__(self, mesh, ang, device=0, mesh_view=None, ang_view=None, general_normals=False):
        self._h = vp()
        mv = mesh_view if mesh_view is not None else mesh.view
        av = ang_view if ang_view is not None else ang.view
        flags = TR_CTX_GENERAL_NORMALS if general_normals else 0
        check(lib.tr_ctx_create(int(device), C.byref(mv), C.byref(av), flags, C.byref(self._h)))
        self._keep = (mesh, ang, mv, av)
        self._sizes()

    def _sizes(self):
        self.nP = int(lib.tr_ctx_nP(self._h))
        self.slm = int(lib.tr_ctx_slm(self._h))
        self.aslm = int(lib.tr_ctx_aslm(self._h))

    def N(self):
        return int(lib.tr_ctx_N(self._h))

    def autotune(self, precision=0):
        """Times the candidate operator schedules and keeps the fastest;
        returns the report text."""
        buf = C.create_string_buffer(16384)
        check(lib.tr_ctx_autotune(self._h, int(precision), buf, len(buf)))
        return buf.value.decode()

    def schedule(self, precision=0):
        out = (C.c_int * 5)()
        check(lib.tr_ctx_schedule(self._h, int(precision), out))
        return dict(zip(("J", "L", "VPW", "WPB", "variant"), list(out)))

    def set_schedule(self, precision=0, J=4, WPB=4, variant=2):
        check(lib.tr_ctx_set_schedule(self._h, int(precision), int(J), int(WPB), int(variant)))

    def set_angular_precision(self, prec):
        """0 FP32, 1 BF16 matrix-core projection, 3 BF16 hi/lo split, 4 FP32
        matrix-core products (33..80 harmonics under the class layout)"""
        check(lib.tr_ctx_set_angular_precision(self._h, int(prec)))

    def set_refinement(self, stagnation=0.5, inner_stall_window=100):
        """Mixed solver: FP64 fallback when a refinement step cuts the true
        residual by less than `stagnation`; inner CG stagnation window."""
        check(lib.tr_ctx_set_refinement(self._h, float(stagnation), int(inner_stall_window)))

    def set_residual_replacement(self, on=True):
        """Mixed solver: continue the inner CG across refinement steps
        (residual replacement, the default) or restart it."""
        check(lib.tr_ctx_set_residual_replacement(self._h, int(bool(on))))

    def permutation(self):
        p = np.empty(self.nP, np.uint32)
        check(lib.tr_ctx_permutation(self._h, ptr(p, u32p)))
        return p

    def deviceAlloc(self, nbytes):
        p = lib.tr_device_alloc(self._h, nbytes)
        if not p:
            raise TransferError(last_error())
        return p

    def allocVector(self):
        p = lib.tr_alloc_vector(self._h)
        if not p:
            raise TransferError(last_error())
        return p

    def deviceFree(self, p):
        check(lib.tr_device_free(self._h, p))

    def copyToDev(self, dst, host):
        host = np.ascontiguousarray(host)
        check(lib.tr_copy_to_dev(self._h, dst, host.ctypes.data, host.nbytes))

    def copyToHost(self, src, nbytes=None, dtype=np.float64):
        nbytes = nbytes if nbytes is not None else 8 * self.N()
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        check(lib.tr_copy_to_host(self._h, out.ctypes.data, src, nbytes))
        return out

    def vector_from(self, host):
        v = self.allocVector()
        self.copyToDev(v, np.ascontiguousarray(host, np.float64).reshape(-1))
        return v

    def set_chebyshev(self, degree=4, ratio=100.0, step_split=2):
        check(lib.tr_ctx_set_chebyshev(self._h, int(degree), float(ratio), int(step_split)))

    def set_multilevel(self, coarsest=2048, agg=8, nu=2):
        check(lib.tr_ctx_set_multilevel(self._h, int(coarsest), int(agg), int(nu)))

    def scalar_guess(self, y, scale, f):
        """f[v][00] = scale * y[v] (device y in caller order), the rest 0."""
        check(lib.tr_ctx_scalar_guess(self._h, y, float(scale), f))

    def mlp_features(self, albedo):
        """(features [nP][8] float32 in the caller's order, sweeps) computed on
        the device (tr_ctx_mlp_features)."""
        d = self.deviceAlloc(4 * 8 * self.nP)
        sw = C.c_int(0)
        check(lib.tr_ctx_mlp_features(self._h, float(albedo), d, C.byref(sw)))
        out = self.copyToHost(d, 4 * 8 * self.nP, np.float32).reshape(self.nP, 8)
        self.deviceFree(d)
        return out, sw.value

    def learned_guess(self, net, albedo, ip_scale, f):
        """f = the learned initial guess (tr_ctx_learned_guess); net is an MLP,
        ip_scale <= 0 takes each vertex's own source."""
        check(lib.tr_ctx_learned_guess(self._h, net._h, float(albedo), float(ip_scale), f))

    def moments(self, f):
        """ComputeMoments of the device vector f: host U[nP], T[9][nP] (T_ij at
        row 3i+j), caller's vertex order."""
        nP = self.nP
        dU, dT = self.deviceAlloc(8 * nP), self.deviceAlloc(8 * 9 * nP)
        try:
            check(lib.tr_compute_moments(self._h, f, dU, dT))
            return self.copyToHost(dU, 8 * nP), self.copyToHost(dT, 8 * 9 * nP).reshape(9, nP)
        finally:
            self.deviceFree(dU)
            self.deviceFree(dT)

    def flux(self, mesh, T):
        """ComputeFlux from host T[9][nP]: host W[3][nT]."""
        nP, nT = self.nP, mesh.nT
        dT, dW = self.deviceAlloc(8 * 9 * nP), self.deviceAlloc(8 * 3 * nT)
        try:
            self.copyToDev(dT, np.ascontiguousarray(T, np.float64).reshape(-1))
            check(lib.tr_compute_flux(self._h, C.byref(mesh.view), dT, dW))
            return self.copyToHost(dW, 8 * 3 * nT).reshape(3, nT)
        finally:
            self.deviceFree(dT)
            self.deviceFree(dW)

    def boundary_flux(self, mesh, f):
        """tr_compute_boundary_flux of the device vector f: host F[nF], the
        outward flux density through each boundary face, face-averaged."""
        nF = mesh.nF
        dF = self.deviceAlloc(8 * max(nF, 1))
        try:
            check(lib.tr_compute_boundary_flux(self._h, C.byref(mesh.view), f, dF))
            return self.copyToHost(dF, 8 * nF)
        finally:
            self.deviceFree(dF)

    def eig_estimate(self, steps=30, jacobi=False, tol=1e-6):
        """tr_eig_estimate: (smallest Ritz value, largest, extrapolated
        largest, steps) of A or, with jacobi, of D^-1 A."""
        out = np.zeros(4)
        check(lib.tr_eig_estimate(self._h, int(steps), int(bool(jacobi)), float(tol), ptr(out)))
        return out[0], out[1], out[2], int(out[3])

    def dump_slae(self, path):
        """main.cpp:73-114 (tr_dump_slae): the dense system to `path`; returns
        (pairs failing the symmetry check, largest asymmetry in units of
        eps (|Z_ij| + |Z_ji|))."""
        n, r = C.c_uint64(0), C.c_double(0)
        check(lib.tr_dump_slae(self._h, str(path).encode(), C.byref(n), C.byref(r)))
        return n.value, r.value

    def getHostSmall(self, src):
        out = np.empty(self.nP * self.slm)
        check(lib.tr_get_host_small(self._h, src, ptr(out)))
        return out

    def computeLhs(self, f, Af):
        check(lib.tr_compute_lhs(self._h, f, Af))

    def computeLhsPrec(self, precision, f, Af):
        check(lib.tr_compute_lhs_prec(self._h, int(precision), f, Af))

    def computeLhsDiag(self, f, Af):
        check(lib.tr_compute_lhs_diag(self._h, f, Af))

    def computeRhs(self, b):
        check(lib.tr_compute_rhs(self._h, b))

    def scale(self, x, wx):
        check(lib.tr_scale(self._h, x, wx))

    def addProd(self, x, y, wy):
        check(lib.tr_add_prod(self._h, x, y, wy))

    def mulAdd(self, x, wx, y):
        check(lib.tr_mul_add(self._h, x, wx, y))

    def mulAddProd(self, x, wx, y, wy):
        check(lib.tr_mul_add_prod(self._h, x, wx, y, wy))

    def norm(self, x):
        v = lib.tr_norm(self._h, x)
        if v != v:
            raise TransferError(last_error())
        return v

    def dot(self, x, y):
        v = lib.tr_dot(self._h, x, y)
        if v != v:
            raise TransferError(last_error())
        return v

    def dot_noalloc(self, x, y, result):
        """tr_dot_noalloc: x . y into the device double `result` (no
        allocation, no synchronisation)."""
        check(lib.tr_dot_noalloc(self._h, x, y, result))

    def synchronize(self):
        check(lib.tr_synchronize(self._h))

    def solve(self, f, solver=TR_SOLVER_REFERENCE, use_prec=False, maxit=1000, tol_abs=1e-4,
              tol_rel=0.0, inner_maxit=1000, inner_tol=1e-8, record=0, x0=False):
        prm = SolveParams(solver, int(use_prec), maxit, tol_abs, tol_rel, inner_maxit, inner_tol, record, int(x0))
        st = SolveStats()
        hist = np.zeros(3 * max(record, 1))
        check(lib.tr_solve(self._h, C.byref(prm), f, C.byref(st), ptr(hist)))
        return st, hist[:3 * record].reshape(-1, 3)

    def bench_lhs(self, reps=20):
        return lib.tr_bench_lhs(self._h, reps)

    def bench_copy(self, nbytes, reps=10):
        """ms per plain device copy of nbytes"""
        out = np.zeros(1)
        check(lib.tr_bench_copy(self._h, int(nbytes), int(reps), ptr(out)))
        return float(out[0])

    def bench_operator(self, warmup, steps, precision=0):
        """(wall s, mean, min, max ms per application) of the FP64 (0) or FP32
        (1) operator, or (2) the FP64 one on vectors in its own harmonic order
        (the K-blocked FP64 kernel alone); (3) the FP32 operator with its
        fused f.Af dot, (4) that after a 1 GiB copy before each application"""
        out = np.zeros(4)
        check(lib.tr_bench_operator(self._h, int(precision), warmup, steps, ptr(out)))
        return tuple(out)

    def kernel_name(self, precision=0):
        buf = C.create_string_buffer(256)
        check(lib.tr_ctx_kernel_name(self._h, int(precision), buf, len(buf)))
        return buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            lib.tr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def expn(n, x):
    """E_n(x), the library's exponential integral."""
    return lib.tr_expn(int(n), float(x))


def step_analytic(x):
    """(U, W, T_xx) of the step problem (main.cpp:498-515, correct E_n)."""
    out = np.zeros(3)
    check(lib.tr_step_analytic(float(x), ptr(out)))
    return tuple(out)


def slab_analytic_U(x, L):
    """Scalar flux of the purely absorbing slab [-L, L], kappa = 1, I_p = 1 on
    x < 0, vacuum at both ends."""
    return lib.tr_slab_analytic_U(float(x), float(L))


class Batch(HipContext):
    """Independent problems (meshes with the same angular data, per-problem
    isotropic albedo) on one context: the union operator plus per-problem CG
    (tr_batch_create / tr_batch_solve).  Problem p owns vertices
    voff[p] .. voff[p+1] of the union vectors."""

    def __init__(self, meshes, ang, albedo=None, device=0, general_normals=False):
        self._h = vp()
        views = (MeshView * len(meshes))(*[m.view for m in meshes])
        alb = None if albedo is None else np.ascontiguousarray(albedo, np.float64)
        flags = TR_CTX_GENERAL_NORMALS if general_normals else 0
        check(lib.tr_batch_create(int(device), len(meshes), views, C.byref(ang.view), ptr(alb), flags,
                                  C.byref(self._h)))
        self._keep = (meshes, ang, views, alb)
        self._sizes()
        n = np.zeros(1, np.uint32)
        check(lib.tr_batch_info(self._h, ptr(n, u32p), None))
        self.nprob = int(n[0])
        self.voff = np.zeros(self.nprob + 1, np.uint32)
        check(lib.tr_batch_info(self._h, ptr(n, u32p), ptr(self.voff, u32p)))

    def solve_batch(self, f, solver=TR_SOLVER_MIXED, use_prec=2, maxit=1000, tol_abs=0.0, tol_rel=1e-8,
                    inner_maxit=20000, inner_tol=1e-4):
        prm = SolveParams(solver, int(use_prec), maxit, tol_abs, tol_rel, inner_maxit, inner_tol, 0)
        st = SolveStats()
        per = (BatchStats * self.nprob)()
        check(lib.tr_batch_solve(self._h, C.byref(prm), f, C.byref(st), per))
        return st, [dict(iters=p.iters, inner_iters=p.inner_iters, true_rel=p.true_rel, status=p.status)
                    for p in per]


MLP_FEATURES, MLP_PARAMS = 8, 4801


def mlp_features(mesh, albedo):
    """Per-vertex features of the learned guess: float32 [nP][8]."""
    out = np.empty((mesh.nP, MLP_FEATURES), np.float32)
    check(lib.tr_mlp_features(C.byref(mesh.view), float(albedo), ptr(out, f32p)))
    return out


class MLP:
    """The learned initial guess: 8 -> 64 -> 64 -> 1 (SiLU, sigmoid), BF16
    weights with FP32 accumulation, trained with Adam by the library's HIP
    kernels.  Device arrays are passed as raw addresses (e.g. from
    HipContext.deviceAlloc)."""

    def __init__(self, params=None, seed=1, device=0):
        self._h = vp()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        check(lib.tr_mlp_create(int(device), ptr(p, f32p), int(seed), C.byref(self._h)))

    def upload(self, host):
        """A device copy (owned by the network) of a float32 host array."""
        host = np.ascontiguousarray(host, np.float32)
        p = vp()
        check(lib.tr_mlp_alloc(self._h, host.nbytes, C.byref(p)))
        check(lib.tr_mlp_copy(self._h, p, host.ctypes.data, host.nbytes, 0))
        return p.value

    def alloc(self, n):
        p = vp()
        check(lib.tr_mlp_alloc(self._h, 4 * int(n), C.byref(p)))
        return p.value

    def download(self, dev, n):
        out = np.empty(int(n), np.float32)
        check(lib.tr_mlp_copy(self._h, out.ctypes.data, dev, out.nbytes, 1))
        return out

    def set_norm(self, mu, sigma):
        mu = np.ascontiguousarray(mu, np.float32)
        sigma = np.ascontiguousarray(sigma, np.float32)
        check(lib.tr_mlp_set_norm(self._h, ptr(mu, f32p), ptr(sigma, f32p)))

    def params(self):
        out = np.empty(MLP_PARAMS, np.float32)
        check(lib.tr_mlp_get_params(self._h, ptr(out, f32p)))
        return out

    def predict(self, feat, n, y):
        check(lib.tr_mlp_predict(self._h, feat, int(n), y))

    def grad(self, feat, target, n):
        loss = np.zeros(1, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        check(lib.tr_mlp_grad(self._h, feat, target, int(n), ptr(loss, f32p), ptr(g, f32p)))
        return float(loss[0]), g

    def train(self, feat, target, n, steps, batch=65536, lr=1e-3, seed=1):
        hist = np.zeros(max(steps, 1), np.float32)
        check(lib.tr_mlp_train(self._h, feat, target, int(n), int(steps), int(batch), float(lr), int(seed),
                               ptr(hist, f32p)))
        return hist[:steps]

    def close(self):
        if getattr(self, "_h", None):
            lib.tr_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
