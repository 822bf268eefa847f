// postproc.hip -- moments and flux of a solution on the device, replacing
// MeshData::ComputeMoments / ComputeFlux (MeshData.cpp:96-136), which the
// reference runs on the host after copying the solution back.
//
//   U_k     = 4 pi f_k[00]
//   T_ij,k  = 4 pi sum_lm omega[i][lm][00] f_k[lm],  j = omega_pos[i][lm][00]
//   W_j,t   = (1/3) sum_{k in t} sum_a s_k^a T_aj,p_k / (kappa V)_t
//
// U and T come out per vertex in the caller's numbering (T as [3i+j][nP]),
// W per tet as [j][nT], matching the reference's arrays.
#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.h"
#include "transfer.h"

namespace tr {

__global__ __launch_bounds__(256) void k_moments(const double *__restrict__ f, const uint32_t *__restrict__ perm,
                                                 const double *__restrict__ mval,
                                                 const uint32_t *__restrict__ mpos, uint32_t nP, uint32_t slm,
                                                 double *__restrict__ U, double *__restrict__ T) {
    const double fourpi = 12.5663706143591729538505735331;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nP; v += gridDim.x * blockDim.x) {
        const double *fv = f + (size_t)v * slm;
        const uint32_t k = perm[v];
        double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t lm = 0; lm < slm; lm++) {
            const double x = fv[lm];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const uint32_t j = mpos[i * slm + lm];
                const double c = fourpi * mval[i * slm + lm] * x;
                // j is 0..2 (checked at assembly); keep the sums in registers
                t[3 * i + 0] += j == 0 ? c : 0.0;
                t[3 * i + 1] += j == 1 ? c : 0.0;
                t[3 * i + 2] += j == 2 ? c : 0.0;
            }
        }
        U[k] = fourpi * fv[0];
#pragma unroll
        for (int q = 0; q < 9; q++) T[(size_t)q * nP + k] = t[q];
    }
}

__global__ __launch_bounds__(256) void k_flux(const tr_tetrahedron *__restrict__ mesh, uint32_t nT, uint32_t nP,
                                              const double *__restrict__ T, double *__restrict__ W) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nT; i += gridDim.x * blockDim.x) {
        const tr_tetrahedron &t = mesh[i];
        double w[3] = {0, 0, 0};
        for (int k = 0; k < 4; k++) {
            const uint32_t p = t.p[k];
            for (int a = 0; a < 3; a++) {
                const double s = t.s[k][a];
#pragma unroll
                for (int j = 0; j < 3; j++) w[j] += s * T[(size_t)(3 * a + j) * nP + p];
            }
        }
#pragma unroll
        for (int j = 0; j < 3; j++) W[(size_t)j * nT + i] = (1.0 / 3.0) * w[j] / t.kappa_volume;
    }
}

// Outward flux density of each boundary face, face-averaged:
//   F[j] = 4 pi sum_lm col[row_j][lm] (f_a + f_b + f_c)[lm] / 3
// rec[j] = {a, b, c (internal vertex ids), row}; row ~0u: no flux
// (reflecting face).  One lane per face; col rows are slm values.
__global__ __launch_bounds__(256) void k_boundary_flux(const double *__restrict__ f, const uint4 *__restrict__ rec,
                                                       const double *__restrict__ col, uint32_t nF, uint32_t slm,
                                                       double *__restrict__ F) {
    const double fourpi = 12.5663706143591729538505735331;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nF; j += gridDim.x * blockDim.x) {
        const uint4 r = rec[j];
        double acc = 0;
        if (r.w != ~0u) {
            const double *c = col + (size_t)r.w * slm;
            const double *fa = f + (size_t)r.x * slm, *fb = f + (size_t)r.y * slm, *fc = f + (size_t)r.z * slm;
            for (uint32_t lm = 0; lm < slm; lm++) acc += c[lm] * (fa[lm] + fb[lm] + fc[lm]);
        }
        F[j] = fourpi * acc * (1.0 / 3.0);
    }
}

hipError_t launch_boundary_flux(const double *f, const void *rec, const double *col, uint32_t nF, uint32_t slm,
                                double *F, hipStream_t s) {
    if (!nF) return hipSuccess;
    const uint32_t g = (nF + 255) / 256;
    hipLaunchKernelGGL(k_boundary_flux, dim3(g), dim3(256), 0, s, f, (const uint4 *)rec, col, nF, slm, F);
    return hipGetLastError();
}

hipError_t launch_moments(const double *f, const uint32_t *perm, const double *mval, const uint32_t *mpos,
                          uint32_t nP, uint32_t slm, double *U, double *T, hipStream_t s) {
    if (!nP) return hipSuccess;
    const uint32_t g = (nP + 255) / 256;
    hipLaunchKernelGGL(k_moments, dim3(g), dim3(256), 0, s, f, perm, mval, mpos, nP, slm, U, T);
    return hipGetLastError();
}

hipError_t launch_flux(const tr_tetrahedron *mesh, uint32_t nT, uint32_t nP, const double *T, double *W,
                       hipStream_t s) {
    if (!nT) return hipSuccess;
    const uint32_t g = (nT + 255) / 256;
    hipLaunchKernelGGL(k_flux, dim3(g), dim3(256), 0, s, mesh, nT, nP, T, W);
    return hipGetLastError();
}

}  // namespace tr
