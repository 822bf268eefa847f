// mesh.cpp -- MeshData built from raw tetrahedral geometry, mesh generators
// and per-tet field generators.
//
// MeshData.cpp:9-94 takes its geometry from meshProcessor (a Netgen .vol
// reader that is not part of the reference tree); here the same arrays are
// derived from vertex coordinates by closed forms:
//   V = |det J| / 6, J = [p1-p0, p2-p0, p3-p0];
//   s_j = 3 V grad(e_j) = -S_j n_j (n_j the outward normal of the face
//   opposite vertex j), so sum_j s_j = 0 (the closure check of
//   MeshData.cpp:74-76);
//   boundary faces: s = S n with n outward (MeshData.cpp:79-91).
// The vertex->(tet, slot) and vertex->(face, slot) CSR arrays are built as in
// MeshData.cpp:17-54, with each vertex's incident tets in ascending order.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace tr {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

double rng_uniform(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t k = splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) + i);
    return (double)(k >> 11) * (1.0 / 9007199254740992.0);
}

// Box-Muller on two consecutive uniforms of a dedicated sub-stream.
double rng_normal(uint64_t seed, uint64_t stream, uint64_t i) {
    double u1 = rng_uniform(seed, stream ^ 0x5bd1e995ULL, 2 * i);
    double u2 = rng_uniform(seed, stream ^ 0x5bd1e995ULL, 2 * i + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * M_PI * u2);
}

static inline void cross(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

void refresh_fields(Mesh &m) {
    for (uint32_t t = 0; t < m.nT; t++) {
        m.mesh[t].kappa_volume = m.kappa[t] * m.volume[t];
        m.mesh[t].I_p = m.Ip[t];
    }
}

int build_mesh(uint32_t nP, const double *coords, uint32_t nT,
               const uint32_t *tets, const double *kappa, const double *Ip,
               Mesh &m) {
    m.nP = nP;
    m.nT = nT;
    m.coords.assign(coords, coords + (size_t)3 * nP);
    m.tets.assign(tets, tets + (size_t)4 * nT);
    for (size_t i = 0; i < (size_t)4 * nT; i++)
        if (tets[i] >= nP) return fail(TR_ERR_ARG, "tet vertex index out of range");
    m.kappa.assign(nT, 1.0);
    m.Ip.assign(nT, 0.0);
    if (kappa) m.kappa.assign(kappa, kappa + nT);
    if (Ip) m.Ip.assign(Ip, Ip + nT);
    m.volume.resize(nT);
    m.mesh.resize(nT);

    for (uint32_t t = 0; t < nT; t++) {
        const uint32_t *p = &tets[(size_t)4 * t];
        const double *x0 = &coords[3 * (size_t)p[0]];
        double c[3][3];
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) c[k][a] = coords[3 * (size_t)p[k + 1] + a] - x0[a];
        double r1[3], r2[3], r3[3];
        cross(c[1], c[2], r1);
        cross(c[2], c[0], r2);
        cross(c[0], c[1], r3);
        double det = c[0][0] * r1[0] + c[0][1] * r1[1] + c[0][2] * r1[2];
        if (!(std::fabs(det) > 0)) return fail(TR_ERR_ARG, "degenerate tetrahedron");
        double V = std::fabs(det) / 6.0;
        double h = (det > 0 ? 0.5 : -0.5);  // 3V/det
        tr_tetrahedron &T = m.mesh[t];
        std::memset(&T, 0, sizeof(T));
        for (int k = 0; k < 4; k++) T.p[k] = p[k];
        for (int a = 0; a < 3; a++) {
            T.s[1][a] = h * r1[a];
            T.s[2][a] = h * r2[a];
            T.s[3][a] = h * r3[a];
            T.s[0][a] = -(T.s[1][a] + T.s[2][a] + T.s[3][a]);
        }
        m.volume[t] = V;
    }
    refresh_fields(m);

    // vertex -> (tet, local slot)
    m.tetstart.assign(nP + 1, 0);
    for (size_t i = 0; i < (size_t)4 * nT; i++) m.tetstart[tets[i] + 1]++;
    for (uint32_t v = 0; v < nP; v++) m.tetstart[v + 1] += m.tetstart[v];
    m.tetidx.resize((size_t)4 * nT);
    m.tetpos.resize((size_t)4 * nT);
    {
        std::vector<uint32_t> fill(m.tetstart.begin(), m.tetstart.end() - 1);
        for (uint32_t t = 0; t < nT; t++)
            for (int k = 0; k < 4; k++) {
                uint32_t v = tets[(size_t)4 * t + k];
                m.tetidx[fill[v]] = t;
                m.tetpos[fill[v]] = k;
                fill[v]++;
            }
    }

    // boundary faces: faces owned by exactly one tet
    struct FaceRec {
        std::array<uint32_t, 3> key;
        uint32_t tet, loc;
    };
    auto face_key = [&](uint32_t t, int j) {
        std::array<uint32_t, 3> key;
        int q = 0;
        for (int k = 0; k < 4; k++)
            if (k != j) key[q++] = tets[(size_t)4 * t + k];
        std::sort(key.begin(), key.end());
        return key;
    };
    // faces in (key, tet) order: a counting sort by the smallest vertex,
    // then a sort of each (short) group by the other two vertices and tet
    std::vector<size_t> gstart((size_t)nP + 1, 0);
    for (uint32_t t = 0; t < nT; t++)
        for (int j = 0; j < 4; j++) gstart[face_key(t, j)[0] + 1]++;
    for (uint32_t v = 0; v < nP; v++) gstart[v + 1] += gstart[v];
    std::vector<FaceRec> fr((size_t)4 * nT);
    {
        std::vector<size_t> fill(gstart.begin(), gstart.end() - 1);
        for (uint32_t t = 0; t < nT; t++)
            for (int j = 0; j < 4; j++) {
                const std::array<uint32_t, 3> key = face_key(t, j);
                fr[fill[key[0]]++] = FaceRec{key, t, (uint32_t)j};
            }
    }
    for (uint32_t v = 0; v < nP; v++)
        std::sort(fr.begin() + gstart[v], fr.begin() + gstart[v + 1], [](const FaceRec &a, const FaceRec &b) {
            return a.key != b.key ? a.key < b.key : a.tet < b.tet;
        });
    m.bnd.clear();
    m.face_tet.clear();
    for (size_t i = 0; i < fr.size();) {
        size_t j = i + 1;
        while (j < fr.size() && fr[j].key == fr[i].key) j++;
        if (j - i > 2) return fail(TR_ERR_ARG, "non-manifold face shared by >2 tets");
        if (j - i == 1) {
            const FaceRec &f = fr[i];
            tr_face F;
            std::memset(&F, 0, sizeof(F));
            const uint32_t *p = &tets[(size_t)4 * f.tet];
            int q = 0;
            for (int k = 0; k < 4; k++)
                if (k != (int)f.loc) F.p[q++] = p[k];
            const double *a = &coords[3 * (size_t)F.p[0]];
            const double *b = &coords[3 * (size_t)F.p[1]];
            const double *c = &coords[3 * (size_t)F.p[2]];
            const double *o = &coords[3 * (size_t)p[f.loc]];
            double e1[3], e2[3], n[3], d[3];
            for (int k = 0; k < 3; k++) {
                e1[k] = b[k] - a[k];
                e2[k] = c[k] - a[k];
                d[k] = o[k] - a[k];
            }
            cross(e1, e2, n);
            double sg = (n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) > 0 ? -0.5 : 0.5;
            for (int k = 0; k < 3; k++) F.s[k] = sg * n[k];
            m.bnd.push_back(F);
            m.face_tet.push_back(f.tet);
        }
        i = j;
    }
    m.nF = (uint32_t)m.bnd.size();
    m.face_kind.assign(m.nF, 0);

    // vertex -> (boundary face, local slot)
    m.facestart.assign(nP + 1, 0);
    for (uint32_t f = 0; f < m.nF; f++)
        for (int k = 0; k < 3; k++) m.facestart[m.bnd[f].p[k] + 1]++;
    for (uint32_t v = 0; v < nP; v++) m.facestart[v + 1] += m.facestart[v];
    m.faceidx.resize((size_t)3 * m.nF);
    m.facepos.resize((size_t)3 * m.nF);
    {
        std::vector<uint32_t> fill(m.facestart.begin(), m.facestart.end() - 1);
        for (uint32_t f = 0; f < m.nF; f++)
            for (int k = 0; k < 3; k++) {
                uint32_t v = m.bnd[f].p[k];
                m.faceidx[fill[v]] = f;
                m.facepos[fill[v]] = k;
                fill[v]++;
            }
    }
    return TR_OK;
}

int generate_box(const tr_mesh_spec &sp, Mesh &m) {
    const int nx = sp.nx, ny = sp.ny, nz = sp.nz;
    if (nx < 1 || ny < 1 || nz < 1) return fail(TR_ERR_ARG, "box needs nx,ny,nz >= 1");
    const double h[3] = {(sp.hi[0] - sp.lo[0]) / nx, (sp.hi[1] - sp.lo[1]) / ny,
                         (sp.hi[2] - sp.lo[2]) / nz};
    const size_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    auto lat = [&](size_t i, size_t j, size_t k) { return i + NX * (j + NY * k); };
    // Kuhn split: v000 -> +e_a -> +e_a+e_b -> v111 over the 6 axis orders.
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                   {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<uint32_t> ltets;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (sp.ball_radius > 0) {
                    double c[3] = {sp.lo[0] + (i + 0.5) * h[0] - sp.ball_center[0],
                                   sp.lo[1] + (j + 0.5) * h[1] - sp.ball_center[1],
                                   sp.lo[2] + (k + 0.5) * h[2] - sp.ball_center[2]};
                    if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > sp.ball_radius * sp.ball_radius)
                        continue;
                }
                for (int q = 0; q < 6; q++) {
                    int ijk[3] = {i, j, k};
                    ltets.push_back((uint32_t)lat(ijk[0], ijk[1], ijk[2]));
                    for (int s = 0; s < 2; s++) {
                        ijk[perm[q][s]]++;
                        ltets.push_back((uint32_t)lat(ijk[0], ijk[1], ijk[2]));
                    }
                    ltets.push_back((uint32_t)lat(i + 1, j + 1, k + 1));
                }
            }
    if (ltets.empty()) return fail(TR_ERR_ARG, "mesh filter removed every cube");
    // compact the used lattice vertices, keeping lattice order
    std::vector<uint32_t> newid(NX * NY * NZ, UINT32_MAX);
    for (uint32_t v : ltets) newid[v] = 0;
    uint32_t nP = 0;
    std::vector<size_t> oldid;
    for (size_t v = 0; v < newid.size(); v++)
        if (newid[v] == 0) {
            newid[v] = nP++;
            oldid.push_back(v);
        }
    for (auto &v : ltets) v = newid[v];
    std::vector<double> xyz((size_t)3 * nP);
    for (uint32_t v = 0; v < nP; v++) {
        size_t L = oldid[v];
        size_t i = L % NX, j = (L / NX) % NY, k = L / (NX * NY);
        xyz[3 * v + 0] = (i == (size_t)nx) ? sp.hi[0] : sp.lo[0] + i * h[0];
        xyz[3 * v + 1] = (j == (size_t)ny) ? sp.hi[1] : sp.lo[1] + j * h[1];
        xyz[3 * v + 2] = (k == (size_t)nz) ? sp.hi[2] : sp.lo[2] + k * h[2];
    }
    uint32_t nT = (uint32_t)(ltets.size() / 4);
    int rc = build_mesh(nP, xyz.data(), nT, ltets.data(), nullptr, nullptr, m);
    if (rc) return rc;
    if (sp.jitter > 0) {
        std::vector<char> onb(nP, 0);
        for (const tr_face &f : m.bnd)
            for (int k = 0; k < 3; k++) onb[f.p[k]] = 1;
        for (uint32_t v = 0; v < nP; v++) {
            if (onb[v]) continue;
            for (int a = 0; a < 3; a++) {
                double u = rng_uniform(sp.seed, 1, 3 * (uint64_t)oldid[v] + a);
                xyz[3 * v + a] += (2 * u - 1) * sp.jitter * h[a];
            }
        }
        rc = build_mesh(nP, xyz.data(), nT, ltets.data(), nullptr, nullptr, m);
    }
    return rc;
}

}  // namespace tr
