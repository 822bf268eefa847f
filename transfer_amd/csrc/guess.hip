// guess.hip -- the learned initial guess built on the device: the
// per-vertex features of the network (the same eight as tr_mlp_features,
// features.cpp) from the context's own data, the network's prediction, and
// the guess vector.
//
// The optical depth to the vacuum boundary, which the host version finds by
// Dijkstra, is the fixpoint of the Bellman equations
//   tau_v = min(tau_v, min_w tau_w + |x_v - x_w| (kappa_v + kappa_w) / 2)
// over the edge neighbours w (the operator's ELL slots), tau = 0 on vacuum
// faces.  The edge weights |x_v - x_w| (kappa_v + kappa_w) / 2 are formed
// once.  Sweeps update tau in place (every value read is the length of some
// path, so races only speed convergence); each launch relaxes its block's
// 256 consecutive vertices (a compact piece of the mesh in the operator's
// Morton order) several times, so distances cross a block in one launch.
// The iteration stops after a launch that changes nothing; the candidate
// values are the same path sums in the same left-to-right order as
// Dijkstra's, so the fixpoint is the same numbers.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx.h"
#include "internal.h"
#include "transfer.h"

namespace tr {

__global__ __launch_bounds__(256) void k_feat_init(uint32_t nP, const double *__restrict__ vsum,
                                                   const int32_t *__restrict__ bof, double *__restrict__ kap,
                                                   double *__restrict__ tau) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nP; v += gridDim.x * blockDim.x) {
        const double vol = vsum[v];
        kap[v] = vol > 0 ? vsum[(size_t)nP + v] / vol : 1.0;
        tau[v] = bof[v] >= 0 ? 0.0 : 1e30;
    }
}

// wgt[s-1][v] = |x_v - x_w| (kappa_w + kappa_v) / 2 for slot s >= 1 (the
// operations of the host's Dijkstra, unfused)
__global__ __launch_bounds__(256) void k_feat_wgt(uint32_t nP, uint32_t W, const uint32_t *__restrict__ nbr,
                                                  const double *__restrict__ xyz, const double *__restrict__ kap,
                                                  double *__restrict__ wgt) {
#pragma clang fp contract(off)
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nP; v += gridDim.x * blockDim.x) {
        const double xv = xyz[v], yv = xyz[(size_t)nP + v], zv = xyz[(size_t)2 * nP + v], kv = kap[v];
        for (uint32_t s = 1; s < W; s++) {
            const uint32_t w = nbr[(size_t)s * nP + v];
            const double dx = xyz[w] - xv, dy = xyz[(size_t)nP + w] - yv, dz = xyz[(size_t)2 * nP + w] - zv;
            double len2 = dx * dx;
            len2 = len2 + dy * dy;
            len2 = len2 + dz * dz;
            wgt[(size_t)(s - 1) * nP + v] = sqrt(len2) * 0.5 * (kap[w] + kv);
        }
    }
}

// PASSES relaxations of the block's 256 consecutive vertices (one thread
// each, one block per 256 vertices)
template <int PASSES>
__global__ __launch_bounds__(256) void k_feat_relax(uint32_t nP, uint32_t W, const uint32_t *__restrict__ nbr,
                                                    const double *__restrict__ wgt, double *tau,
                                                    int *__restrict__ changed) {
#pragma clang fp contract(off)
    const uint32_t v = blockIdx.x * 256u + threadIdx.x;
    int any = 0;
    for (int p = 0; p < PASSES; p++) {
        if (v < nP) {
            const double t = tau[v];
            double best = t;
            for (uint32_t s = 1; s < W; s++) {
                const uint32_t w = nbr[(size_t)s * nP + v];
                if (w == v) continue;
                const double nd = tau[w] + wgt[(size_t)(s - 1) * nP + v];
                best = nd < best ? nd : best;
            }
            if (best < t) {
                tau[v] = best;
                any = 1;
            }
        }
        __syncthreads();
    }
    if (any) *changed = 1;
}

__global__ __launch_bounds__(256) void k_feat_final(uint32_t nP, uint32_t W, const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ perm,
                                                    const double *__restrict__ vsum, const double *__restrict__ kap,
                                                    const double *__restrict__ tau, const double *__restrict__ mf0,
                                                    double albedo, float *__restrict__ feat) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nP; v += gridDim.x * blockDim.x) {
        const double lk = log(kap[v]), tb = fmin(tau[v], 1e3);
        double nb = 0;
        uint32_t cnt = 0;
        for (uint32_t s = 1; s < W; s++) {
            const uint32_t w = nbr[(size_t)s * nP + v];
            if (w == v) continue;
            nb += log(kap[w]);
            cnt++;
        }
        nb = cnt ? nb / cnt - lk : 0.0;
        const double a = mf0 ? 1.0 - mf0[v] : albedo;
        const double dl = sqrt(3.0 * (1.0 - a));
        const double kw = vsum[(size_t)nP + v], ip = kw > 0 ? vsum[(size_t)2 * nP + v] / kw : 0.0;
        float *o = feat + 8 * (size_t)perm[v];
        o[0] = (float)lk;
        o[1] = (float)a;
        o[2] = (float)log(1.0 - a);
        o[3] = (float)log(tb + 1e-3);
        o[4] = (float)exp(-tb);
        o[5] = (float)exp(-tb * dl);
        o[6] = (float)nb;
        o[7] = (float)log(fmax(ip, 1e-30));
    }
}

// tau = 0 on the vertices of vacuum faces off the axes (not in bof)
__global__ __launch_bounds__(256) void k_feat_gzero(uint32_t ngv, const uint32_t *__restrict__ gvert,
                                                    double *__restrict__ tau) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ngv; i += gridDim.x * blockDim.x) tau[gvert[i]] = 0.0;
}

// f[v][00] = y[v] * I_p / (1 - albedo), zero elsewhere; I_p = ip_scale, or
// exp(feature 7) (the vertex's own source) when ip_scale <= 0
__global__ __launch_bounds__(256) void k_learned_guess(const float *__restrict__ y, const float *__restrict__ feat,
                                                       const uint32_t *__restrict__ perm,
                                                       const double *__restrict__ mf0, uint32_t nP, uint32_t slm,
                                                       double albedo, double ip_scale, double *__restrict__ f) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nP * slm;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t v = i / slm;
        double val = 0.0;
        if (i - v * slm == 0) {
            const uint32_t p = perm[v];
            const double ip = ip_scale > 0 ? ip_scale : exp((double)feat[8 * (size_t)p + 7]);
            val = (double)y[p] * ip / (mf0 ? mf0[v] : 1.0 - albedo);
        }
        f[i] = val;
    }
}

static unsigned blocks_of(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

}  // namespace tr

using tr::fail;

// Per-vertex features of the learned guess (those of tr_mlp_features) into
// the device array feat[nP][8], in the caller's vertex order.  albedo is
// the problem's (batch contexts use each problem's own).  *sweeps (may be
// NULL) receives the number of optical-depth sweeps.
extern "C" int tr_ctx_mlp_features(tr_ctx *c, double albedo, float *feat, int *sweeps) {
    if (!c || !feat) return fail(TR_ERR_ARG, "tr_ctx_mlp_features: null argument");
    if (!c->d_vsum || !c->d_xyz)
        return fail(TR_ERR_ARG, "tr_ctx_mlp_features: the context's mesh has no coordinates");
    if (!c->d_mf0 && !(albedo >= 0.0 && albedo < 1.0))
        return fail(TR_ERR_ARG, "tr_ctx_mlp_features: albedo in [0, 1)");
    const uint32_t nP = c->op.nP, W = c->op.W;
    if (!c->d_feat_tmp) {
        HIPCHECK(hipMalloc((void **)&c->d_feat_tmp, sizeof(double) * ((size_t)W + 1) * nP + sizeof(int) * 8));
        c->owned.push_back(c->d_feat_tmp);
    }
    double *kap = c->d_feat_tmp, *tau = kap + nP, *wgt = tau + nP;
    int *flags = (int *)(wgt + (size_t)(W - 1) * nP);
    const unsigned g = tr::blocks_of(nP);
    hipLaunchKernelGGL(tr::k_feat_init, dim3(g), dim3(256), 0, c->stream, nP, c->d_vsum, c->d_bof, kap, tau);
    const uint32_t ngv = (uint32_t)c->op.gvert.size();
    if (ngv) hipLaunchKernelGGL(tr::k_feat_gzero, dim3(tr::blocks_of(ngv)), dim3(256), 0, c->stream, ngv, c->d_gvert, tau);
    hipLaunchKernelGGL(tr::k_feat_wgt, dim3(g), dim3(256), 0, c->stream, nP, W, c->d_nbr, c->d_xyz, kap, wgt);
    HIPCHECK(hipGetLastError());
    // chunks of 4 launches of 4 passes; converged when the chunk's last
    // launch changed nothing (at most nP launches: Bellman-Ford's bound)
    constexpr int PASSES = 4, CHUNK = 4;
    const unsigned gb = (nP + 255u) / 256u;
    int done = 0, total = 0;
    while (!done) {
        HIPCHECK(hipMemsetAsync(flags, 0, sizeof(int) * CHUNK, c->stream));
        for (int s = 0; s < CHUNK; s++)
            hipLaunchKernelGGL(tr::k_feat_relax<PASSES>, dim3(gb), dim3(256), 0, c->stream, nP, W, c->d_nbr, wgt,
                               tau, flags + s);
        HIPCHECK(hipGetLastError());
        int last = 0;
        HIPCHECK(hipMemcpyAsync(&last, flags + CHUNK - 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        total += CHUNK;
        done = last == 0 || (uint64_t)total > (uint64_t)nP + CHUNK;
    }
    total *= PASSES;
    hipLaunchKernelGGL(tr::k_feat_final, dim3(g), dim3(256), 0, c->stream, nP, W, c->d_nbr, c->d_perm, c->d_vsum, kap,
                       tau, c->d_mf0, albedo, feat);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (sweeps) *sweeps = total;
    return TR_OK;
}

// The learned initial guess in f (a device vector of the context): the
// features on the device, the network's prediction y = U / U_inf, and
// f[v][00] = y I_p / (1 - albedo) (I_p = ip_scale, or each vertex's own
// source when ip_scale <= 0).  Scratch space is kept by the context.
extern "C" int tr_ctx_learned_guess(tr_ctx *c, tr_mlp *m, double albedo, double ip_scale, double *f) {
    if (!c || !m || !f) return fail(TR_ERR_ARG, "tr_ctx_learned_guess: null argument");
    const uint32_t nP = c->op.nP;
    if (!c->d_guess_tmp) {
        HIPCHECK(hipMalloc((void **)&c->d_guess_tmp, sizeof(float) * 9 * (size_t)nP));
        c->owned.push_back(c->d_guess_tmp);
    }
    float *feat = c->d_guess_tmp, *y = feat + 8 * (size_t)nP;
    int rc = tr_ctx_mlp_features(c, albedo, feat, nullptr);
    if (rc) return rc;
    if ((rc = tr_mlp_predict(m, feat, nP, y))) return rc;
    hipLaunchKernelGGL(tr::k_learned_guess, dim3(tr::blocks_of((size_t)nP * c->op.slm)), dim3(256), 0, c->stream, y,
                       feat, c->d_perm, c->d_mf0, nP, c->op.slm, albedo, ip_scale, f);
    HIPCHECK(hipGetLastError());
    return TR_OK;
}
