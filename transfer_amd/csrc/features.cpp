// features.cpp -- per-vertex inputs of the learned initial guess (mlp.hip),
// from the element data and the element neighbourhood of each vertex:
//
//   0  log kappa_v            kappa averaged over the incident tets (by volume)
//   1  albedo
//   2  log(1 - albedo)
//   3  log(tau_b + 1e-3)      optical depth to the vacuum boundary: shortest
//                             path over the mesh edges with edge weight
//                             length * mean kappa of its ends (Dijkstra)
//   4  exp(-tau_b)
//   5  exp(-tau_b sqrt(3 (1 - albedo)))   (diffusion-length attenuation)
//   6  mean over edge neighbours of log kappa, minus log kappa_v
//   7  log I_p,v              source temperature term, averaged by kappa V
//
// The network's target is U / U_inf with U_inf = 4 pi I_p / (1 - albedo),
// the infinite-medium scalar radiation of the vertex's local source.
#include <algorithm>
#include <cmath>
#include <functional>
#include <queue>
#include <vector>

#include "internal.h"
#include "transfer.h"

extern "C" int tr_mlp_features(const tr_mesh_view *mv, double albedo, float *feat) {
    using tr::fail;
    if (!mv || !feat || !mv->coords || !mv->mesh) return fail(TR_ERR_ARG, "tr_mlp_features: mesh with coordinates required");
    if (!(albedo >= 0.0 && albedo < 1.0)) return fail(TR_ERR_ARG, "tr_mlp_features: albedo in [0, 1)");
    const uint32_t nP = mv->nP, nT = mv->nT;
    const double *X = mv->coords;
    std::vector<double> vol(nP, 0.0), kv(nP, 0.0), ipw(nP, 0.0), kw(nP, 0.0);
    std::vector<std::vector<uint32_t>> adj(nP);
    for (uint32_t t = 0; t < nT; t++) {
        const tr_tetrahedron &T = mv->mesh[t];
        const double *a = X + 3 * T.p[0], *b = X + 3 * T.p[1], *c = X + 3 * T.p[2], *d = X + 3 * T.p[3];
        double u[3], v[3], w[3];
        for (int k = 0; k < 3; k++) {
            u[k] = b[k] - a[k];
            v[k] = c[k] - a[k];
            w[k] = d[k] - a[k];
        }
        const double V = std::fabs(u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                                   u[2] * (v[0] * w[1] - v[1] * w[0])) / 6.0;
        if (!(V > 0)) return fail(TR_ERR_ARG, "tr_mlp_features: degenerate tetrahedron");
        const double kap = T.kappa_volume / V;
        for (int k = 0; k < 4; k++) {
            const uint32_t p = T.p[k];
            vol[p] += V;
            kv[p] += kap * V;
            ipw[p] += T.I_p * kap * V;
            kw[p] += kap * V;
            for (int l = 0; l < 4; l++)
                if (l != k) adj[p].push_back(T.p[l]);
        }
    }
    for (uint32_t p = 0; p < nP; p++) {
        std::sort(adj[p].begin(), adj[p].end());
        adj[p].erase(std::unique(adj[p].begin(), adj[p].end()), adj[p].end());
        kv[p] = vol[p] > 0 ? kv[p] / vol[p] : 1.0;
    }
    // optical depth to the vacuum boundary
    std::vector<double> tau(nP, 1e30);
    using Item = std::pair<double, uint32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    for (uint32_t f = 0; f < mv->nF; f++) {
        if (mv->face_kind && mv->face_kind[f] != 0) continue;
        for (int k = 0; k < 3; k++) {
            const uint32_t p = mv->bnd[f].p[k];
            if (tau[p] > 0) {
                tau[p] = 0;
                pq.push({0.0, p});
            }
        }
    }
    while (!pq.empty()) {
        const auto [d, p] = pq.top();
        pq.pop();
        if (d > tau[p]) continue;
        for (uint32_t q : adj[p]) {
            double len2 = 0;
            for (int k = 0; k < 3; k++) len2 += (X[3 * p + k] - X[3 * q + k]) * (X[3 * p + k] - X[3 * q + k]);
            const double nd = d + std::sqrt(len2) * 0.5 * (kv[p] + kv[q]);
            if (nd < tau[q]) {
                tau[q] = nd;
                pq.push({nd, q});
            }
        }
    }
    const double dl = std::sqrt(3.0 * (1.0 - albedo));
    for (uint32_t p = 0; p < nP; p++) {
        const double lk = std::log(kv[p]), tb = std::min(tau[p], 1e3);
        double nb = 0;
        for (uint32_t q : adj[p]) nb += std::log(kv[q]);
        nb = adj[p].empty() ? 0.0 : nb / adj[p].size() - lk;
        const double ip = kw[p] > 0 ? ipw[p] / kw[p] : 0.0;
        float *o = feat + 8 * (size_t)p;
        o[0] = (float)lk;
        o[1] = (float)albedo;
        o[2] = (float)std::log(1.0 - albedo);
        o[3] = (float)std::log(tb + 1e-3);
        o[4] = (float)std::exp(-tb);
        o[5] = (float)std::exp(-tb * dl);
        o[6] = (float)nb;
        o[7] = (float)std::log(std::max(ip, 1e-30));
    }
    return TR_OK;
}
