// mg.hip -- the multilevel preconditioner (extension; use_prec = 3).
//
// Levels are contexts whose operators come from coarsen_operator (Galerkin
// products of the Kronecker-structured operator under aggregation of runs of
// consecutive Morton-ordered vertices), applied by the same kernels and
// sharing the finest context's stream.  One application is a symmetric
// V-cycle: on every level one damped-Jacobi step z = w D^-1 r, the residual
// restricted to the next level (a sum over each aggregate's rows), the
// coarse correction prolonged back (added to every row of the aggregate),
// and one damped-Jacobi step on the updated residual; the coarsest level
// takes `nu` damped-Jacobi steps.  D is diag(A) of the level and
// w = 1 / (1.1 lambda), lambda the largest eigenvalue of D^-1 A from a power
// iteration, so every smoothing step contracts in the A-norm and the
// V-cycle is symmetric positive definite: a valid CG preconditioner.  The
// finest level's output is projected on the reflection-symmetric subspace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.h"
#include "internal.h"
#include "kernels.h"
#include "stencil.h"

namespace tr {

struct Multilevel {
    std::vector<tr_ctx *> lev;          // lev[0] = the context itself (not owned here)
    std::vector<uint32_t *> agg;        // agg[k]: starts of level k+1's aggregates in level k's rows
    std::vector<double> omega;
    std::vector<double *> dinv64, t64, r64, z64;
    std::vector<float *> dinv32, t32, r32, z32;
    std::vector<void *> owned;
    int nu = 2;                          // damped-Jacobi steps on the coarsest level
    bool have32 = false;
};

template <typename T>
__global__ __launch_bounds__(256) void k_jac0(const T *__restrict__ r, const T *__restrict__ dinv, T w, T *__restrict__ z,
                                              size_t n, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = w * dinv[i] * r[i];
}

// z += w D^-1 (r - t)
template <typename T>
__global__ __launch_bounds__(256) void k_jacupd(const T *__restrict__ r, const T *__restrict__ t,
                                                const T *__restrict__ dinv, T w, T *__restrict__ z, size_t n,
                                                const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = fma(w * dinv[i], r[i] - t[i], z[i]);
}

// rc[I][lm] = sum over the aggregate's rows i of (r - t)[i][lm]
template <typename T>
__global__ __launch_bounds__(256) void k_restrict(const T *__restrict__ r, const T *__restrict__ t,
                                                  T *__restrict__ rc, const uint32_t *__restrict__ agg, uint32_t nC,
                                                  uint32_t slm, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nC * slm;
         q += (size_t)gridDim.x * blockDim.x) {
        const uint32_t I = (uint32_t)(q / slm), lm = (uint32_t)(q - (size_t)I * slm);
        T s = 0;
        for (uint32_t i = agg[I]; i < agg[I + 1]; i++) s += r[(size_t)i * slm + lm] - t[(size_t)i * slm + lm];
        rc[q] = s;
    }
}

// z[i][lm] += zc[I][lm] for the rows i of aggregate I
template <typename T>
__global__ __launch_bounds__(256) void k_prolong(T *__restrict__ z, const T *__restrict__ zc,
                                                 const uint32_t *__restrict__ agg, uint32_t nC, uint32_t slm,
                                                 const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nC * slm;
         q += (size_t)gridDim.x * blockDim.x) {
        const uint32_t I = (uint32_t)(q / slm), lm = (uint32_t)(q - (size_t)I * slm);
        const T c = zc[q];
        for (uint32_t i = agg[I]; i < agg[I + 1]; i++) z[(size_t)i * slm + lm] += c;
    }
}

static unsigned grid_of(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 4096); }

template <typename T>
static OpArgs<T> level_args(tr_ctx *c);
template <>
OpArgs<double> level_args<double>(tr_ctx *c) { return c->args(); }
template <>
OpArgs<float> level_args<float>(tr_ctx *c) { return c->args32(); }
template <typename T>
static const LaneMap &level_map(tr_ctx *c) { return sizeof(T) == 8 ? c->map : c->map32; }

template <typename T>
static hipError_t vcycle(Multilevel &mg, size_t k, const T *r, T *z, const double *skip, hipStream_t st) {
    tr_ctx *c = mg.lev[k];
    const OpArgs<T> a = level_args<T>(c);
    const LaneMap &m = level_map<T>(c);
    const int NC = (int)c->op.NC;
    const size_t n = c->nint();
    const T w = (T)mg.omega[k];
    const T *dinv = sizeof(T) == 8 ? (const T *)mg.dinv64[k] : (const T *)mg.dinv32[k];
    T *t = sizeof(T) == 8 ? (T *)mg.t64[k] : (T *)mg.t32[k];
    const unsigned g = grid_of(n);
    hipError_t e;
    hipLaunchKernelGGL(k_jac0<T>, dim3(g), dim3(256), 0, st, r, dinv, w, z, n, skip);
    const size_t L = mg.lev.size() - 1;
    if (k == L) {
        for (int it = 1; it < mg.nu; it++) {
            if ((e = launch_apply<T>(a, NC, m, z, t, nullptr, skip, st)) != hipSuccess) return e;
            hipLaunchKernelGGL(k_jacupd<T>, dim3(g), dim3(256), 0, st, r, t, dinv, w, z, n, skip);
        }
        return hipGetLastError();
    }
    tr_ctx *cc = mg.lev[k + 1];
    const uint32_t nC = cc->op.nP, slm = c->op.slm;
    T *rc = sizeof(T) == 8 ? (T *)mg.r64[k + 1] : (T *)mg.r32[k + 1];
    T *zc = sizeof(T) == 8 ? (T *)mg.z64[k + 1] : (T *)mg.z32[k + 1];
    if ((e = launch_apply<T>(a, NC, m, z, t, nullptr, skip, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_restrict<T>, dim3(grid_of((size_t)nC * slm)), dim3(256), 0, st, r, t, rc, mg.agg[k], nC, slm,
                       skip);
    if ((e = vcycle<T>(mg, k + 1, rc, zc, skip, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_prolong<T>, dim3(grid_of((size_t)nC * slm)), dim3(256), 0, st, z, zc, mg.agg[k], nC, slm, skip);
    if ((e = launch_apply<T>(a, NC, m, z, t, nullptr, skip, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_jacupd<T>, dim3(g), dim3(256), 0, st, r, t, dinv, w, z, n, skip);
    if (k == 0 && (e = launch_zero_mask<T>(a, z, skip, st)) != hipSuccess) return e;
    return hipGetLastError();
}

// Largest eigenvalue of D^-1 A by power iteration (FP64, 20 steps); the
// norm and the normalisation on the device (one scalar read per step)
int lambda_max(tr_ctx *c, const double *dinv, double *x, double *y, double &lam) {
    const size_t n = c->nint();
    {
        std::vector<double> h(n);
        for (size_t i = 0; i < n; i++) h[i] = 0.5 + rng_uniform(11, 7, i);
        // synchronous: h goes out of scope here
        HIPCHECK(hipMemcpy(x, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    lam = 0;
    for (int it = 0; it < 20; it++) {
        HIPCHECK(launch_apply<double>(c->args(), (int)c->op.NC, c->map, x, y, nullptr, nullptr, c->stream));
        hipLaunchKernelGGL(k_jac0<double>, dim3(grid_of(n)), dim3(256), 0, c->stream, y, dinv, 1.0, y, n, nullptr);
        HIPCHECK(launch_dot(y, y, n, c->d_part, c->d_red, c->stream));
        double s = 0;
        HIPCHECK(hipMemcpyAsync(&s, c->d_red, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        s = std::sqrt(s);
        if (!(s > 0) || !std::isfinite(s)) return fail(TR_ERR_HIP, "power iteration broke down");
        // x has unit norm after the first step, so |D^-1 A x| estimates lambda
        if (it > 0) lam = s;
        HIPCHECK(launch_axpby(x, 0.0, y, 1.0 / s, n, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return TR_OK;
}

template <typename T>
static T *dalloc(Multilevel &mg, size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
    mg.owned.push_back(p);
    return (T *)p;
}

void destroy_multilevel(Multilevel *mg) {
    if (!mg) return;
    for (size_t k = 1; k < mg->lev.size(); k++) tr_ctx_destroy(mg->lev[k]);
    for (void *p : mg->owned) hipFree(p);
    delete mg;
}

// Builds the hierarchy of c (down to at most `min_coarse` vertices on the
// coarsest level, aggregates of `size` vertices), FP32 data when `f32`.
static int build_levels(tr_ctx *c, Multilevel *mg, uint32_t min_coarse, uint32_t size);

int build_multilevel(tr_ctx *c, bool f32, uint32_t min_coarse, uint32_t size) {
    int rc;
    // faces off the axes (TR_CTX_GENERAL_NORMALS) stay on the fine level: the
    // coarse operators are the Galerkin products of the stencil without
    // them (still SPD), the smoother and residual use the full operator
    if (!c->mg) {
        // assembled aside and attached only when complete
        Multilevel *mg = new Multilevel();
        mg->nu = c->mg_nu;
        if ((rc = build_levels(c, mg, min_coarse, size))) {
            destroy_multilevel(mg);
            return rc;
        }
        c->mg = mg;
    }
    Multilevel *mg = c->mg;
    if (f32 && !mg->have32) {
        for (size_t k = 0; k < mg->lev.size(); k++) {
            tr_ctx *l = mg->lev[k];
            if ((rc = ensure_fp32(l))) return rc;
            const size_t n = l->nint();
            std::vector<double> d = operator_diagonal(l->op);
            std::vector<float> df(n);
            for (size_t i = 0; i < n; i++) df[i] = (float)(1.0 / d[i]);
            float *dv = dalloc<float>(*mg, n), *t = dalloc<float>(*mg, n), *r = dalloc<float>(*mg, n),
                  *z = dalloc<float>(*mg, n);
            if (!z) return fail(TR_ERR_HIP, "multilevel: out of device memory");
            HIPCHECK(hipMemcpyAsync(dv, df.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));  // df is a temporary
            mg->dinv32.push_back(dv);
            mg->t32.push_back(t);
            mg->r32.push_back(r);
            mg->z32.push_back(z);
        }
        mg->have32 = true;
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return TR_OK;
}

static int build_levels(tr_ctx *c, Multilevel *mg, uint32_t min_coarse, uint32_t size) {
    int rc;
    {
        mg->lev.push_back(c);
        while (mg->lev.size() < 10 && mg->lev.back()->op.nP > min_coarse) {
            tr_ctx *f = mg->lev.back();
            std::vector<uint32_t> agg, cvoff;
            tr_ctx *cc = new tr_ctx();
            cc->device = c->device;
            if ((rc = coarsen_operator(f->op, f->voff, size, cc->op, agg, cvoff))) {
                delete cc;
                return rc;
            }
            if (!f->voff.empty()) cc->voff = cvoff;
            cc->stream = c->stream;
            cc->own_stream = false;
            if ((rc = ctx_init(cc, c->device))) return rc;
            mg->lev.push_back(cc);
            uint32_t *d = dalloc<uint32_t>(*mg, agg.size());
            if (!d) return fail(TR_ERR_HIP, "multilevel: out of device memory");
            HIPCHECK(hipMemcpyAsync(d, agg.data(), sizeof(uint32_t) * agg.size(), hipMemcpyHostToDevice, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));  // agg is a temporary
            mg->agg.push_back(d);
        }
        for (size_t k = 0; k < mg->lev.size(); k++) {
            tr_ctx *l = mg->lev[k];
            const size_t n = l->nint();
            std::vector<double> d = operator_diagonal(l->op);
            for (double &v : d) v = 1.0 / v;
            double *dv = dalloc<double>(*mg, n), *t = dalloc<double>(*mg, n), *r = dalloc<double>(*mg, n),
                   *z = dalloc<double>(*mg, n);
            if (!z) return fail(TR_ERR_HIP, "multilevel: out of device memory");
            HIPCHECK(hipMemcpyAsync(dv, d.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
            mg->dinv64.push_back(dv);
            mg->t64.push_back(t);
            mg->r64.push_back(r);
            mg->z64.push_back(z);
            double lam = 0;
            if ((rc = lambda_max(l, dv, r, z, lam))) return rc;
            mg->omega.push_back(1.0 / (1.1 * lam));
        }
    }
    return TR_OK;
}

hipError_t multilevel_apply64(tr_ctx *c, const double *r, double *z, const double *skip) {
    return vcycle<double>(*c->mg, 0, r, z, skip, c->stream);
}
hipError_t multilevel_apply32(tr_ctx *c, const float *r, float *z, const double *skip) {
    return vcycle<float>(*c->mg, 0, r, z, skip, c->stream);
}

int multilevel_levels(const tr_ctx *c) { return c->mg ? (int)c->mg->lev.size() : 0; }

}  // namespace tr
