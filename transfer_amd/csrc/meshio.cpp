// meshio.cpp -- mesh input and solution output of the reference workflow,
// written here because the reference reads and writes through its
// un-vendored meshProcessor submodule (MeshData.cpp:10, main.cpp:209-216):
//
//   tr_mesh_load_vol   Netgen's neutral volume mesh (.vol): the `points` and
//                      `volumeelements` sections (1-based point numbers,
//                      material number = the cfg's [region N]); every other
//                      section is ignored.  kappa and I_p per tet come from
//                      the config's regions, as MeshData.cpp:62-63.
//   tr_write_vtk       legacy VTK unstructured grid with the fields
//                      main.cpp:209-216 saves: W per tet (vector), U, T
//                      (tensor), aU, aW, aT per vertex.
//   tr_dump_slae       the dense system dump and symmetry check of
//                      main.cpp:73-114, through the context's C ABI.
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"
#include "transfer.h"

struct tr_mesh {
    tr::Mesh m;
};
struct tr_config {
    tr::Config c;
};

using tr::fail;

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// next line that is neither empty nor a comment
bool data_line(std::istream &in, std::string &line) {
    while (std::getline(in, line)) {
        line = trim(line);
        if (!line.empty() && line[0] != '#') return true;
    }
    return false;
}

}  // namespace

extern "C" {

int tr_mesh_load_vol(const char *path, const tr_config *cfg, tr_mesh **out) {
    if (!path || !out) return fail(TR_ERR_ARG, "tr_mesh_load_vol: null argument");
    std::ifstream in(path);
    if (!in) return fail(TR_ERR_ARG, std::string("tr_mesh_load_vol: cannot open ") + path);
    std::vector<double> xyz;
    std::vector<uint32_t> tets, mat;
    bool have_points = false, have_elems = false;
    std::string line;
    while (std::getline(in, line)) {
        const std::string key = trim(line);
        if (key != "points" && key != "volumeelements") continue;
        std::string cnt;
        if (!data_line(in, cnt)) return fail(TR_ERR_ARG, "tr_mesh_load_vol: " + key + " without a count");
        char *end = nullptr;
        const unsigned long n = std::strtoul(cnt.c_str(), &end, 10);
        if (end == cnt.c_str()) return fail(TR_ERR_ARG, "tr_mesh_load_vol: bad count after " + key);
        for (unsigned long i = 0; i < n; i++) {
            std::string rec;
            if (!data_line(in, rec)) return fail(TR_ERR_ARG, "tr_mesh_load_vol: " + key + " section ends early");
            std::istringstream ss(rec);
            if (key == "points") {
                double x, y, z;
                if (!(ss >> x >> y >> z)) return fail(TR_ERR_ARG, "tr_mesh_load_vol: bad point record");
                xyz.insert(xyz.end(), {x, y, z});
            } else {
                long m, np;
                if (!(ss >> m >> np) || np < 4 || m < 0) return fail(TR_ERR_ARG, "tr_mesh_load_vol: bad element record");
                uint32_t p[4];
                for (int k = 0; k < 4; k++) {
                    long q;
                    if (!(ss >> q) || q < 1) return fail(TR_ERR_ARG, "tr_mesh_load_vol: bad element vertex");
                    p[k] = (uint32_t)(q - 1);
                }
                tets.insert(tets.end(), p, p + 4);
                mat.push_back((uint32_t)m);
            }
        }
        (key == "points" ? have_points : have_elems) = true;
    }
    if (!have_points || !have_elems || xyz.empty() || tets.empty())
        return fail(TR_ERR_ARG, "tr_mesh_load_vol: no points or no volume elements");
    const uint32_t nP = (uint32_t)(xyz.size() / 3), nT = (uint32_t)mat.size();
    for (uint32_t v : tets)
        if (v >= nP) return fail(TR_ERR_ARG, "tr_mesh_load_vol: element refers to a missing point");
    std::vector<double> kappa(nT, 1.0), Ip(nT, 0.0);
    if (cfg) {
        const tr::Config &c = cfg->c;
        for (uint32_t t = 0; t < nT; t++) {
            if ((int)mat[t] > c.nreg || mat[t] >= c.kappa.size())
                return fail(TR_ERR_ARG, "tr_mesh_load_vol: material " + std::to_string(mat[t]) +
                                            " has no [region] section in the config");
            kappa[t] = c.kappa[mat[t]];
            Ip[t] = c.Ip[mat[t]];
        }
    }
    tr_mesh *m = new tr_mesh();
    int rc = tr::build_mesh(nP, xyz.data(), nT, tets.data(), kappa.data(), Ip.data(), m->m);
    if (rc) {
        delete m;
        return rc;
    }
    m->m.region = mat;
    *out = m;
    return TR_OK;
}

int tr_mesh_get_regions(const tr_mesh *m, uint32_t *region) {
    if (!m || !region) return fail(TR_ERR_ARG, "tr_mesh_get_regions: null argument");
    for (uint32_t t = 0; t < m->m.nT; t++) region[t] = t < m->m.region.size() ? m->m.region[t] : 0u;
    return TR_OK;
}

int tr_write_vtk(const tr_mesh *mm, const char *path, const double *U, const double *T, const double *W,
                 const double *aU, const double *aW, const double *aT) {
    if (!mm || !path) return fail(TR_ERR_ARG, "tr_write_vtk: null argument");
    const tr::Mesh &m = mm->m;
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(TR_ERR_ARG, std::string("tr_write_vtk: cannot create ") + path);
    std::fprintf(f, "# vtk DataFile Version 3.0\ntransfer_amd solution\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(f, "POINTS %u double\n", m.nP);
    for (uint32_t v = 0; v < m.nP; v++)
        std::fprintf(f, "%.17g %.17g %.17g\n", m.coords[3 * (size_t)v], m.coords[3 * (size_t)v + 1],
                     m.coords[3 * (size_t)v + 2]);
    std::fprintf(f, "CELLS %u %u\n", m.nT, 5 * m.nT);
    for (uint32_t t = 0; t < m.nT; t++)
        std::fprintf(f, "4 %u %u %u %u\n", m.tets[4 * (size_t)t], m.tets[4 * (size_t)t + 1], m.tets[4 * (size_t)t + 2],
                     m.tets[4 * (size_t)t + 3]);
    std::fprintf(f, "CELL_TYPES %u\n", m.nT);
    for (uint32_t t = 0; t < m.nT; t++) std::fprintf(f, "10\n");
    if (W) {
        std::fprintf(f, "CELL_DATA %u\nVECTORS W double\n", m.nT);
        for (uint32_t t = 0; t < m.nT; t++)
            std::fprintf(f, "%.17g %.17g %.17g\n", W[t], W[(size_t)m.nT + t], W[2 * (size_t)m.nT + t]);
    }
    if (U || T || aU || aW || aT) std::fprintf(f, "POINT_DATA %u\n", m.nP);
    const struct {
        const char *name;
        const double *v;
    } sc[] = {{"U", U}, {"aU", aU}, {"aW", aW}, {"aT", aT}};
    for (const auto &s : sc) {
        if (!s.v) continue;
        std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", s.name);
        for (uint32_t v = 0; v < m.nP; v++) std::fprintf(f, "%.17g\n", s.v[v]);
    }
    if (T) {
        std::fprintf(f, "TENSORS T double\n");
        for (uint32_t v = 0; v < m.nP; v++)
            for (int a = 0; a < 3; a++)
                std::fprintf(f, "%.17g %.17g %.17g\n", T[(size_t)(3 * a) * m.nP + v], T[(size_t)(3 * a + 1) * m.nP + v],
                             T[(size_t)(3 * a + 2) * m.nP + v]);
    }
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok) return fail(TR_ERR_ARG, std::string("tr_write_vtk: write failed: ") + path);
    return TR_OK;
}

// main.cpp:73-114: the matrix column by column (A e_k for every unknown k
// outside the padding lanes), one text row per column (its nP*slm
// unpadded entries, then b[k]), and the symmetry check
// |Z_ij - Z_ji| > 100 eps (|Z_ij| + |Z_ji|) with eps the unit roundoff.
int tr_dump_slae(tr_ctx *ctx, const char *path, uint64_t *n_asym, double *max_ratio) {
    if (!ctx || !path) return fail(TR_ERR_ARG, "tr_dump_slae: null argument");
    const uint32_t N = tr_ctx_N(ctx), nP = tr_ctx_nP(ctx), slm = tr_ctx_slm(ctx), aslm = tr_ctx_aslm(ctx);
    if ((uint64_t)N * N > (uint64_t)1 << 28)
        return fail(TR_ERR_ARG, "tr_dump_slae: the dense matrix is for small meshes (N*N <= 2^28)");
    double *f = tr_alloc_vector(ctx), *Af = tr_alloc_vector(ctx), *b = tr_alloc_vector(ctx);
    if (!f || !Af || !b) return fail(TR_ERR_HIP, "tr_dump_slae: out of device memory");
    std::vector<double> Z((size_t)N * N, 0.0), e(N, 0.0), col(N), hb(N);
    FILE *out = std::fopen(path, "w");
    int rc = out ? TR_OK : fail(TR_ERR_ARG, std::string("tr_dump_slae: cannot create ") + path);
    if (!rc) rc = tr_compute_rhs(ctx, b);
    if (!rc) rc = tr_copy_to_host(ctx, hb.data(), b, sizeof(double) * N);
    for (uint32_t k = 0; k < N && !rc; k++) {
        if (k % aslm >= slm) continue;
        e[k] = 1.0;
        rc = tr_copy_to_dev(ctx, f, e.data(), sizeof(double) * N);
        e[k] = 0.0;
        if (!rc) rc = tr_compute_lhs(ctx, f, Af);
        if (!rc) rc = tr_copy_to_host(ctx, col.data(), Af, sizeof(double) * N);
        if (rc) break;
        std::memcpy(&Z[(size_t)k * N], col.data(), sizeof(double) * N);
        for (uint32_t i = 0; i < nP; i++)
            for (uint32_t j = 0; j < slm; j++) std::fprintf(out, "% 2.14g ", col[(size_t)i * aslm + j]);
        std::fprintf(out, "% 2.14g \n", hb[k]);
    }
    if (out && std::fclose(out) != 0 && !rc) rc = fail(TR_ERR_ARG, "tr_dump_slae: write failed");
    tr_device_free(ctx, f);
    tr_device_free(ctx, Af);
    tr_device_free(ctx, b);
    if (rc) return rc;
    const double eps = DBL_EPSILON / 2;  // Util.cpp:4-14: the largest x with 1 + x == 1 after halving
    uint64_t bad = 0;
    double worst = 0;
    for (uint32_t i = 0; i < N; i++)
        for (uint32_t j = i + 1; j < N; j++) {
            const double a = Z[(size_t)N * i + j], c = Z[(size_t)N * j + i];
            const double diff = std::fabs(a - c), allow = eps * (std::fabs(a) + std::fabs(c));
            if (allow > 0) worst = std::max(worst, diff / allow);
            if (diff > 100 * allow) bad++;
        }
    if (n_asym) *n_asym = bad;
    if (max_ratio) *max_ratio = worst;
    return TR_OK;
}

}  // extern "C"
