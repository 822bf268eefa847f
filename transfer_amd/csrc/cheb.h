// cheb.h -- the Chebyshev polynomial preconditioner (use_prec 4) shared by
// the single-problem solvers (context.hip) and the batched solver
// (batch.hip): z = the deg-th iterate of the Chebyshev iteration on
// D^-1 A z = D^-1 r from z = 0 over [lambda_max / ratio, lambda_max], in
// the three-term form of Saad, Iterative Methods for Sparse Linear Systems,
// Alg. 12.1.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "kernels.h"

namespace tr {

// z = s dinv r, the first Chebyshev iterate from z = 0
template <typename T>
__global__ __launch_bounds__(256) void k_cheb_first(const T *__restrict__ r, const T *__restrict__ dinv, T s,
                                                    T *__restrict__ z, size_t n, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = s * dinv[i] * r[i];
}

template <typename T>
struct ChebPrec {
    OpArgs<T> a;
    int NC = 7;
    LaneMap map;
    const T *dinv = nullptr;
    T *zb = nullptr, *tmp = nullptr;
    size_t n = 0;
    double inv_theta = 0;
    std::vector<double> om, ga;  // steps 1 .. deg-1
    void setup(int deg, double lmax, double ratio) {
        const double lmin = lmax / ratio, th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sig = th / de;
        double rho = 1.0 / sig;
        inv_theta = 1.0 / th;
        om.clear();
        ga.clear();
        for (int k = 1; k < deg; k++) {
            const double rn = 1.0 / (2.0 * sig - rho);
            om.push_back(rn * rho);
            ga.push_back(2.0 * rn / de);
            rho = rn;
        }
    }
    // the buffer that holds the first iterate when the result goes to z
    T *first_buffer(T *z) const { return (om.size() % 2 == 0) ? z : zb; }
    // r.z partials the last step writes (matrix-core schedules, at least
    // one step, at most 1024 workgroups), else 0
    int nparts() const {
        if (a.ngv) return 0;  // the fused r.z dot would miss the share of faces off the axes
        const uint32_t p = cheb_parts(a.nP, map);
        return (!om.empty() && p <= 1024) ? (int)p : 0;
    }
    hipError_t apply(const T *r, T *z, const double *skip, hipStream_t st) const {
        const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_cheb_first<T>, dim3(g), dim3(256), 0, st, r, dinv, (T)inv_theta, first_buffer(z), n,
                           skip);
        hipError_t e = hipGetLastError();
        return e != hipSuccess ? e : steps(r, z, skip, st, nullptr);
    }
    // the steps after the first iterate (already in first_buffer(z)); the
    // last one also writes nparts() r.z partials to rzpart when given
    hipError_t steps(const T *r, T *z, const double *skip, hipStream_t st, double *rzpart) const {
        const int ns = (int)om.size();
        T *cur = first_buffer(z), *oth = cur == z ? zb : z;
        hipError_t e = hipSuccess;
        for (int k = 0; k < ns && e == hipSuccess; k++) {
            ChebStep<T> cs;
            cs.rhs = r;
            cs.dinv = dinv;
            cs.zc = (T)(1.0 + om[k]);
            cs.om = k == 0 ? (T)0 : (T)om[k];  // z_0 = 0
            cs.ga = (T)ga[k];
            e = launch_cheb_step<T>(a, NC, map, cur, oth, cs, tmp, skip, st,
                                    (k == ns - 1 && rzpart && nparts()) ? rzpart : nullptr);
            std::swap(cur, oth);
        }
        return e;
    }
};


}  // namespace tr
