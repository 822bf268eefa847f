// c_api.cpp -- extern "C" entry points of the host front end: AngularData,
// MeshData (generators), Config, field generators and errors.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "internal.h"

namespace tr {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

}  // namespace tr

struct tr_angular {
    tr::Angular a;
};
struct tr_mesh {
    tr::Mesh m;
};
struct tr_config {
    tr::Config c;
};

using namespace tr;

extern "C" {

const char *tr_last_error(void) { return g_error.c_str(); }

int tr_angular_create(int maxk, double albedo, double g, int hg_lmax, tr_angular **out) {
    if (!out) return fail(TR_ERR_ARG, "tr_angular_create: null out");
    if (!(albedo >= 0 && albedo < 1) || !(g > -1 && g < 1) || hg_lmax < 0)
        return fail(TR_ERR_ARG, "tr_angular_create: need 0 <= albedo < 1, |g| < 1, hg_lmax >= 0");
    tr_angular *a = new tr_angular();
    try {
        build_angular(maxk, albedo, g, hg_lmax, a->a);
    } catch (const std::exception &e) {
        delete a;
        return fail(TR_ERR_ARG, e.what());
    }
    *out = a;
    return TR_OK;
}

int tr_angular_destroy(tr_angular *a) {
    delete a;
    return TR_OK;
}

int tr_angular_get_view(const tr_angular *a, tr_angular_view *v) {
    if (!a || !v) return fail(TR_ERR_ARG, "tr_angular_get_view: null argument");
    v->slm = a->a.slm;
    v->omega = a->a.omega.data();
    v->omega_pos = a->a.omega_pos.data();
    v->Ox = a->a.Ox.data();
    v->Oy = a->a.Oy.data();
    v->Oz = a->a.Oz.data();
    v->massfac = a->a.massfac.data();
    return TR_OK;
}

int tr_mesh_create_box(const tr_mesh_spec *spec, tr_mesh **out) {
    if (!spec || !out) return fail(TR_ERR_ARG, "tr_mesh_create_box: null argument");
    tr_mesh *m = new tr_mesh();
    int rc = generate_box(*spec, m->m);
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return TR_OK;
}

int tr_mesh_create(uint32_t nP, const double *coords, uint32_t nT, const uint32_t *tets,
                   const double *kappa, const double *Ip, tr_mesh **out) {
    if (!coords || !tets || !out || nP == 0 || nT == 0)
        return fail(TR_ERR_ARG, "tr_mesh_create: null or empty argument");
    tr_mesh *m = new tr_mesh();
    int rc = build_mesh(nP, coords, nT, tets, kappa, Ip, m->m);
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return TR_OK;
}

int tr_mesh_destroy(tr_mesh *m) {
    delete m;
    return TR_OK;
}

int tr_mesh_get_view(const tr_mesh *mm, tr_mesh_view *v) {
    if (!mm || !v) return fail(TR_ERR_ARG, "tr_mesh_get_view: null argument");
    const Mesh &m = mm->m;
    v->nP = m.nP;
    v->nF = m.nF;
    v->nT = m.nT;
    v->tetstart = m.tetstart.data();
    v->tetidx = m.tetidx.data();
    v->tetpos = m.tetpos.data();
    v->mesh = m.mesh.data();
    v->facestart = m.facestart.data();
    v->faceidx = m.faceidx.data();
    v->facepos = m.facepos.data();
    v->bnd = m.bnd.data();
    v->coords = m.coords.data();
    v->face_kind = m.face_kind.data();
    return TR_OK;
}

int tr_mesh_get_geometry(const tr_mesh *mm, double *coords, uint32_t *tets, double *centroids,
                         double *volumes) {
    if (!mm) return fail(TR_ERR_ARG, "tr_mesh_get_geometry: null mesh");
    const Mesh &m = mm->m;
    if (coords) std::memcpy(coords, m.coords.data(), sizeof(double) * 3 * m.nP);
    if (tets) std::memcpy(tets, m.tets.data(), sizeof(uint32_t) * 4 * m.nT);
    if (volumes) std::memcpy(volumes, m.volume.data(), sizeof(double) * m.nT);
    if (centroids)
        for (uint32_t t = 0; t < m.nT; t++)
            for (int a = 0; a < 3; a++) {
                double s = 0;
                for (int k = 0; k < 4; k++) s += m.coords[3 * (size_t)m.tets[4 * (size_t)t + k] + a];
                centroids[3 * (size_t)t + a] = 0.25 * s;
            }
    return TR_OK;
}

int tr_mesh_set_fields(tr_mesh *mm, const double *kappa, const double *Ip) {
    if (!mm) return fail(TR_ERR_ARG, "tr_mesh_set_fields: null mesh");
    Mesh &m = mm->m;
    if (kappa) {
        for (uint32_t t = 0; t < m.nT; t++)
            if (!(kappa[t] > 0)) return fail(TR_ERR_ARG, "tr_mesh_set_fields: kappa must be > 0");
        m.kappa.assign(kappa, kappa + m.nT);
    }
    if (Ip) m.Ip.assign(Ip, Ip + m.nT);
    refresh_fields(m);
    return TR_OK;
}

int tr_mesh_set_face_kind_plane(tr_mesh *mm, int axis, double value, double tol, int kind) {
    if (!mm || axis < 0 || axis > 2 || kind < 0 || kind > 1)
        return fail(TR_ERR_ARG, "tr_mesh_set_face_kind_plane: bad argument");
    Mesh &m = mm->m;
    for (uint32_t f = 0; f < m.nF; f++) {
        bool on = true;
        for (int k = 0; k < 3; k++)
            on = on && std::fabs(m.coords[3 * (size_t)m.bnd[f].p[k] + axis] - value) <= tol;
        if (on) m.face_kind[f] = (uint8_t)kind;
    }
    return TR_OK;
}

int tr_rng_uniform(uint64_t seed, uint64_t stream, uint64_t n, double *out) {
    if (!out) return fail(TR_ERR_ARG, "tr_rng_uniform: null out");
    for (uint64_t i = 0; i < n; i++) out[i] = rng_uniform(seed, stream, i);
    return TR_OK;
}

int tr_field_eval(const tr_mesh *mm, int kind, const double *p, uint64_t seed, uint64_t stream,
                  double *out) {
    if (!mm || !p || !out) return fail(TR_ERR_ARG, "tr_field_eval: null argument");
    const Mesh &m = mm->m;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t v = 0; v < m.nP; v++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], m.coords[3 * (size_t)v + a]);
            hi[a] = std::fmax(hi[a], m.coords[3 * (size_t)v + a]);
        }
    // blob parameters
    int nb = kind == TR_FIELD_BLOBS ? (int)p[1] : 0;
    std::vector<double> amp(nb), wid(nb), cen(3 * (size_t)nb);
    for (int b = 0; b < nb; b++) {
        double u0 = rng_uniform(seed, stream, 8 * (uint64_t)b + 0);
        double u1 = rng_uniform(seed, stream, 8 * (uint64_t)b + 1);
        amp[b] = std::exp(std::log(p[2]) + u0 * (std::log(p[3]) - std::log(p[2])));
        wid[b] = p[4] + u1 * (p[5] - p[4]);
        for (int a = 0; a < 3; a++)
            cen[3 * b + a] = lo[a] + rng_uniform(seed, stream, 8 * (uint64_t)b + 2 + a) * (hi[a] - lo[a]);
    }
    for (uint32_t t = 0; t < m.nT; t++) {
        double c[3];
        for (int a = 0; a < 3; a++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += m.coords[3 * (size_t)m.tets[4 * (size_t)t + k] + a];
            c[a] = 0.25 * s;
        }
        double r = std::sqrt((c[0] - p[2]) * (c[0] - p[2]) + (c[1] - p[3]) * (c[1] - p[3]) +
                             (c[2] - p[4]) * (c[2] - p[4]));
        double v;
        switch (kind) {
            case TR_FIELD_CONST: v = p[0]; break;
            case TR_FIELD_RADIAL_EXP: v = p[0] * std::exp(-r / p[1]); break;
            case TR_FIELD_UNIFORM: v = p[0] + p[1] * rng_uniform(seed, stream, t); break;
            case TR_FIELD_LOGNORMAL: v = std::exp(p[0] + p[1] * rng_normal(seed, stream, t)); break;
            case TR_FIELD_BLOBS: {
                v = p[0];
                for (int b = 0; b < nb; b++) {
                    double d2 = 0;
                    for (int a = 0; a < 3; a++) d2 += (c[a] - cen[3 * b + a]) * (c[a] - cen[3 * b + a]);
                    v += amp[b] * std::exp(-d2 / (2 * wid[b] * wid[b]));
                }
                break;
            }
            case TR_FIELD_PLANCK: {
                double q = std::fmax(0.0, 1.0 - p[0] * r);
                v = q * q * q * q;
                break;
            }
            case TR_FIELD_STEP: {
                int ax = (int)p[0];
                if (ax < 0 || ax > 2) return fail(TR_ERR_ARG, "tr_field_eval: bad axis");
                v = c[ax] < p[1] ? p[2] : p[3];
                break;
            }
            default: return fail(TR_ERR_ARG, "tr_field_eval: unknown kind");
        }
        out[t] = v;
    }
    return TR_OK;
}

int tr_config_load(const char *path, tr_config **out) {
    if (!path || !out) return fail(TR_ERR_ARG, "tr_config_load: null argument");
    tr_config *c = new tr_config();
    int rc = load_config(path, c->c);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return TR_OK;
}

int tr_config_destroy(tr_config *c) {
    delete c;
    return TR_OK;
}

int tr_config_ints(const tr_config *c, int *v) {
    if (!c || !v) return fail(TR_ERR_ARG, "tr_config_ints: null argument");
    v[0] = c->c.maxk;
    v[1] = c->c.dev;
    v[2] = c->c.dump;
    v[3] = c->c.prec;
    v[4] = c->c.soltype;
    v[5] = c->c.nreg;
    return TR_OK;
}

int tr_config_region(const tr_config *c, int region, double *kappa, double *Ip) {
    if (!c || region < 0 || region > c->c.nreg)
        return fail(TR_ERR_ARG, "tr_config_region: region out of range");
    if (kappa) *kappa = c->c.kappa[region];
    if (Ip) *Ip = c->c.Ip[region];
    return TR_OK;
}

int tr_config_strings(const tr_config *c, char *meshfile, char *outfile, size_t cap) {
    if (!c || cap == 0) return fail(TR_ERR_ARG, "tr_config_strings: bad argument");
    if (meshfile) {
        std::strncpy(meshfile, c->c.meshfn.c_str(), cap - 1);
        meshfile[cap - 1] = 0;
    }
    if (outfile) {
        std::strncpy(outfile, c->c.outfn.c_str(), cap - 1);
        outfile[cap - 1] = 0;
    }
    return TR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Introspection of the assembled operator (host data only).
#include "stencil.h"

struct tr_stencil {
    tr::HostOperator op;
};

extern "C" {

int tr_stencil_create(const tr_mesh_view *mesh, const tr_angular_view *ang, tr_stencil **out) {
    if (!mesh || !ang || !out) return fail(TR_ERR_ARG, "tr_stencil_create: null argument");
    tr_stencil *s = new tr_stencil();
    int rc = assemble_operator(*mesh, *ang, s->op);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return TR_OK;
}

int tr_stencil_coarsen(const tr_stencil *s, uint32_t size, tr_stencil **out, uint32_t *agg) {
    if (!s || !out) return fail(TR_ERR_ARG, "tr_stencil_coarsen: null argument");
    tr_stencil *c = new tr_stencil();
    std::vector<uint32_t> a, cv;
    int rc = coarsen_operator(s->op, {}, size, c->op, a, cv);
    if (rc) {
        delete c;
        return rc;
    }
    if (agg) std::memcpy(agg, a.data(), sizeof(uint32_t) * a.size());
    *out = c;
    return TR_OK;
}

int tr_stencil_destroy(tr_stencil *s) {
    delete s;
    return TR_OK;
}

int tr_stencil_mask(const tr_stencil *s, uint32_t *count, uint32_t *list) {
    if (!s || !count) return fail(TR_ERR_ARG, "tr_stencil_mask: null argument");
    *count = (uint32_t)s->op.mask.size();
    if (list && !s->op.mask.empty()) std::memcpy(list, s->op.mask.data(), sizeof(uint32_t) * s->op.mask.size());
    return TR_OK;
}

int tr_stencil_info(const tr_stencil *s, uint32_t *info) {
    if (!s || !info) return fail(TR_ERR_ARG, "tr_stencil_info: null argument");
    const HostOperator &o = s->op;
    uint32_t v[8] = {o.nP, o.slm, o.NC, o.W, o.nB, o.WB, o.arow, o.brow};
    std::memcpy(info, v, sizeof v);
    return TR_OK;
}

}  // extern "C"

template <typename T>
static void put(T *dst, const std::vector<T> &src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(T) * src.size());
}

extern "C" {

int tr_stencil_export(const tr_stencil *s, uint32_t *perm, uint32_t *nbr, double *coef,
                      uint32_t *bvert, uint32_t *bnbr, double *bcoef, double *omc, double *onc,
                      double *massfac, double *rhs) {
    if (!s) return fail(TR_ERR_ARG, "tr_stencil_export: null stencil");
    const HostOperator &o = s->op;
    put(perm, o.perm);
    put(nbr, o.nbr);
    put(coef, o.coef);
    put(bvert, o.bvert);
    if (o.nB) {
        put(bnbr, o.bnbr);
        put(bcoef, o.bcoef);
    }
    put(omc, o.omc);
    put(onc, o.onc);
    put(massfac, o.massfac);
    put(rhs, o.rhs);
    return TR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Angular matrices of a face normal off the axes (host only; test hooks).
extern "C" {

int tr_abs_normal_matrix(int maxk, const double *n, double *O) {
    if (maxk < 0 || maxk > 7 || !n || !O) return fail(TR_ERR_ARG, "tr_abs_normal_matrix: bad argument");
    std::vector<double> M;
    abs_normal_matrix(maxk, n, M);
    std::copy(M.begin(), M.end(), O);
    return TR_OK;
}

int tr_abs_normal_column(int maxk, const double *n, double *col) {
    if (maxk < 0 || maxk > 7 || !n || !col) return fail(TR_ERR_ARG, "tr_abs_normal_column: bad argument");
    abs_normal_column(maxk, n, col);
    return TR_OK;
}

int tr_normal_rotation(int maxk, const double *n, double *R, uint32_t *size) {
    if (maxk < 0 || maxk > 7 || !n) return fail(TR_ERR_ARG, "tr_normal_rotation: bad argument");
    if (size) *size = rotation_size(maxk);
    if (R) {
        std::vector<double> M;
        normal_rotation(maxk, n, M);
        std::copy(M.begin(), M.end(), R);
    }
    return TR_OK;
}

}  // extern "C"
