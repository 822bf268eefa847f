// coarse.hip -- the learned coarse correction (use_prec 5 of the device
// solvers): the preconditioner
//
//   z = D^-1 r + g (g . r) / (g . A g)
//
// adds to Jacobi the A-orthogonal projection onto the one-dimensional space
// spanned by g, the learned guess (the solver's scaled start).  D^-1 is SPD
// and g g^T / (g^T A g) is positive semidefinite, so M is a fixed SPD
// preconditioner for CG.  It sits in the slot main.cpp:223-248 gives
// invprec.  Two launches per application, captured into the CG iteration
// graph: the partials of g . r, then z with g . r summed per workgroup in
// a fixed order (every workgroup gets the same value).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "kernels.h"

namespace tr {

#define TR_CC_BLOCKS 1024

__device__ inline double cc_wave_sum(double s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    return s;
}

__device__ inline double cc_block_sum(double s) {
    __shared__ double w[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    s = cc_wave_sum(s);
    if (lane == 0) w[wave] = s;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / 64); i++) t += w[i];
    return t;  // valid in thread 0
}

template <typename T>
__global__ __launch_bounds__(256) void k_cc_dot(const T *__restrict__ g, const T *__restrict__ r, size_t n,
                                                double *__restrict__ part, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s = fma((double)g[i], (double)r[i], s);
    s = cc_block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_cc_apply(const T *__restrict__ r, T *__restrict__ z,
                                                  const T *__restrict__ dinv, const T *__restrict__ g, size_t n,
                                                  const double *__restrict__ part, int np,
                                                  const double *__restrict__ gAg, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    __shared__ double c_sh;
    {
        double s = 0;
        for (int i = threadIdx.x; i < np; i += blockDim.x) s += part[i];
        s = cc_block_sum(s);
        if (threadIdx.x == 0) c_sh = s / *gAg;
        __syncthreads();
    }
    const T c = (T)c_sh;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        z[i] = fma(c, g[i], dinv[i] * r[i]);
}

template <typename T>
hipError_t launch_learned_coarse(const T *r, T *z, const T *dinv, const T *g, size_t n, double *part,
                                 const double *gAg, const double *skip, hipStream_t s) {
    const unsigned G = (unsigned)std::min<size_t>(std::max<size_t>((n + 1023) / 1024, 1), TR_CC_BLOCKS);
    hipLaunchKernelGGL(k_cc_dot<T>, dim3(G), dim3(256), 0, s, g, r, n, part, skip);
    hipLaunchKernelGGL(k_cc_apply<T>, dim3(G), dim3(256), 0, s, r, z, dinv, g, n, part, (int)G, gAg, skip);
    return hipGetLastError();
}

int learned_coarse_partials() { return TR_CC_BLOCKS; }

template hipError_t launch_learned_coarse<double>(const double *, double *, const double *, const double *, size_t,
                                                  double *, const double *, const double *, hipStream_t);
template hipError_t launch_learned_coarse<float>(const float *, float *, const float *, const float *, size_t,
                                                 double *, const double *, const double *, hipStream_t);

}  // namespace tr
