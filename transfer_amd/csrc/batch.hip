// batch.hip -- many independent problems solved together (extension; the
// reference solves one mesh per process).
//
// The problems' operators are concatenated into one block-diagonal union
// (concat_operators), so one launch of the operator kernel covers every
// problem and the autotuned schedule applies to all of them.  CG runs per
// problem: the vector kernels walk blocks that never straddle two problems,
// every block writes its partial sums, and one thread per problem adds its
// blocks' partials in a fixed order (deterministic) into that problem's
// alpha, beta and convergence state.  A converged problem gets alpha = 0 and
// stops changing; the whole iteration is captured into one hipGraph of
// `chunk` iterations that the host relaunches until every problem has
// converged.  The mixed solver is FP64 iterative refinement per problem
// around the batched FP32 CG.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "cheb.h"
#include "ctx.h"
#include "internal.h"
#include "kernels.h"
#include "solver.h"
#include "stencil.h"

namespace tr {

#define TR_BLK_ENTRIES 4096  // vector entries per reduction block
#define TR_BLK_THREADS 256

// per-problem scalar slots
// BS_RZD: r.z of the direction the last iteration used; BS_B1, BS_C: the
// residual replacement's coefficients (k_brep_fin)
enum { BS_RZ, BS_RR, BS_BB, BS_ALPHA, BS_BETA, BS_THR, BS_CONV, BS_ITER, BS_RZD, BS_B1, BS_C, BS_N };

// Blocks of entries aligned to problems.
struct BlockTable {
    uint32_t nblk = 0, nprob = 0;
    uint32_t *lo = nullptr, *hi = nullptr, *prob = nullptr;  // [nblk]
    uint32_t *b0 = nullptr;                                   // [nprob+1] first block of each problem
};

__device__ inline double wsum(double s) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    return s;
}

// fixed-order workgroup sums of up to three values; valid in thread 0
__device__ inline void bsum3(double &a, double &b, double &c) {
    __shared__ double w[3][TR_BLK_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    a = wsum(a);
    b = wsum(b);
    c = wsum(c);
    if (lane == 0) {
        w[0][wave] = a;
        w[1][wave] = b;
        w[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = 0;
        for (int i = 0; i < TR_BLK_THREADS / 64; i++) {
            a += w[0][i];
            b += w[1][i];
            c += w[2][i];
        }
    }
}

// Blocks of problems that have stopped (converged, capped or broken down)
// skip the vector kernels: their alpha and beta are 0, their partials are
// no longer read, and their p no longer matters (alpha = 0 keeps x and r).
__device__ inline bool bstopped(const double *sc, const uint32_t *prob) {
    return sc && sc[(size_t)prob[blockIdx.x] * BS_N + BS_CONV] != 0.0;
}

// part[blk] = x . y over the block (sc, prob: the stopped problems' blocks
// skip, when given)
template <typename T>
__global__ __launch_bounds__(TR_BLK_THREADS) void k_bdot(const T *__restrict__ x, const T *__restrict__ y,
                                                         const uint32_t *__restrict__ lo,
                                                         const uint32_t *__restrict__ hi, double *__restrict__ part,
                                                         const double *__restrict__ skip,
                                                         const double *__restrict__ sc = nullptr,
                                                         const uint32_t *__restrict__ prob = nullptr) {
    if ((skip && *skip != 0.0) || bstopped(sc, prob)) return;
    double s = 0, u = 0, w = 0;
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS)
        s = fma((double)x[i], (double)y[i], s);
    bsum3(s, u, w);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha_p = rz_p / (p . Ap)_p for running problems, 0 otherwise
__global__ void k_bfin_alpha(const double *__restrict__ part, const uint32_t *__restrict__ b0, uint32_t nprob,
                             double *__restrict__ sc, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nprob; p += gridDim.x * blockDim.x) {
        double *s = sc + (size_t)p * BS_N;
        double pap = 0;
        for (uint32_t b = b0[p]; b < b0[p + 1]; b++) pap += part[b];
        if (s[BS_CONV] != 0.0) {
            s[BS_ALPHA] = 0;
        } else if (!(pap > 0)) {
            s[BS_ALPHA] = 0;
            s[BS_CONV] = 3;  // breakdown
        } else {
            s[BS_ALPHA] = s[BS_RZ] / pap;
        }
    }
}

// x += a p, r -= a Ap, z = D^-1 r; partials r.r and r.z; with z1 != null
// also z1 = cscale cdinv r (a Chebyshev preconditioner's first iterate)
template <typename T>
__global__ __launch_bounds__(TR_BLK_THREADS) void k_bupdate(T *__restrict__ x, const T *__restrict__ p,
                                                            T *__restrict__ r, const T *__restrict__ Ap,
                                                            const T *__restrict__ dinv,
                                                            const uint32_t *__restrict__ lo,
                                                            const uint32_t *__restrict__ hi,
                                                            const uint32_t *__restrict__ prob,
                                                            const double *__restrict__ sc, double *__restrict__ part,
                                                            uint32_t nblk, const double *__restrict__ skip,
                                                            const T *__restrict__ cdinv, T cscale, T *__restrict__ z1) {
    if ((skip && *skip != 0.0) || bstopped(sc, prob)) return;
    const double a = sc[(size_t)prob[blockIdx.x] * BS_N + BS_ALPHA];
    double rr = 0, rz = 0, w = 0;
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS) {
        x[i] = (T)((double)x[i] + a * (double)p[i]);
        const T ri = (T)((double)r[i] - a * (double)Ap[i]);
        r[i] = ri;
        if (z1) z1[i] = cscale * cdinv[i] * ri;
        const double zi = dinv ? (double)dinv[i] * (double)ri : (double)ri;
        rr = fma((double)ri, (double)ri, rr);
        rz = fma((double)ri, zi, rz);
    }
    bsum3(rr, rz, w);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[nblk + blockIdx.x] = rz;
    }
}

// beta, convergence test and iteration count per problem; all[0] = every
// problem stopped.  One workgroup.
__global__ __launch_bounds__(TR_BLK_THREADS) void k_bfin_beta(const double *__restrict__ part,
                                                              const uint32_t *__restrict__ b0, uint32_t nprob,
                                                              uint32_t nblk, double *__restrict__ sc, int maxit,
                                                              double *__restrict__ all) {
    if (all[0] != 0.0) return;
    __shared__ int running;
    if (threadIdx.x == 0) running = 0;
    __syncthreads();
    int mine = 0;
    for (uint32_t p = threadIdx.x; p < nprob; p += blockDim.x) {
        double *s = sc + (size_t)p * BS_N;
        if (s[BS_CONV] != 0.0) continue;
        double rr = 0, rz = 0;
        for (uint32_t b = b0[p]; b < b0[p + 1]; b++) {
            rr += part[b];
            rz += part[nblk + b];
        }
        s[BS_ITER] += 1;
        s[BS_RR] = rr;
        if (std::sqrt(rr) <= s[BS_THR])
            s[BS_CONV] = 1;
        else if (s[BS_ITER] >= maxit)
            s[BS_CONV] = 2;
        s[BS_BETA] = s[BS_CONV] != 0.0 ? 0.0 : rz / s[BS_RZ];
        s[BS_RZD] = s[BS_RZ];
        s[BS_RZ] = rz;
        mine |= s[BS_CONV] == 0.0;
    }
    if (mine) atomicOr(&running, 1);
    __syncthreads();
    if (threadIdx.x == 0 && !running) all[0] = 1;
}

// p = z + beta p, z = zb (a preconditioner's output) or D^-1 r
template <typename T>
__global__ __launch_bounds__(TR_BLK_THREADS) void k_bpupdate(T *__restrict__ p, const T *__restrict__ r,
                                                             const T *__restrict__ dinv,
                                                             const uint32_t *__restrict__ lo,
                                                             const uint32_t *__restrict__ hi,
                                                             const uint32_t *__restrict__ prob,
                                                             const double *__restrict__ sc,
                                                             const double *__restrict__ skip,
                                                             const T *__restrict__ zb) {
    if ((skip && *skip != 0.0) || bstopped(sc, prob)) return;
    const double be = sc[(size_t)prob[blockIdx.x] * BS_N + BS_BETA];
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS) {
        const double z = zb ? (double)zb[i] : (dinv ? (double)dinv[i] * (double)r[i] : (double)r[i]);
        p[i] = (T)(z + be * (double)p[i]);
    }
}

// r = b - r (r holds A x on entry), z = D^-1 r, p = z; partials r.r, r.z, b.b
template <typename T>
__global__ __launch_bounds__(TR_BLK_THREADS) void k_binit(const T *__restrict__ b, T *__restrict__ r,
                                                          T *__restrict__ p, const T *__restrict__ dinv,
                                                          const uint32_t *__restrict__ lo,
                                                          const uint32_t *__restrict__ hi, double *__restrict__ part,
                                                          uint32_t nblk) {
    double rr = 0, rz = 0, bb = 0;
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS) {
        const double ri = (double)b[i] - (double)r[i];
        r[i] = (T)ri;
        const double z = dinv ? (double)dinv[i] * ri : ri;
        p[i] = (T)z;
        rr = fma(ri, ri, rr);
        rz = fma(ri, z, rz);
        bb = fma((double)b[i], (double)b[i], bb);
    }
    bsum3(rr, rz, bb);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[nblk + blockIdx.x] = rz;
        part[2 * nblk + blockIdx.x] = bb;
    }
}

__global__ __launch_bounds__(TR_BLK_THREADS) void k_bfin_init(const double *__restrict__ part,
                                                              const uint32_t *__restrict__ b0, uint32_t nprob,
                                                              uint32_t nblk, double *__restrict__ sc, double tol_abs,
                                                              double tol_rel, double *__restrict__ all) {
    __shared__ int running;
    if (threadIdx.x == 0) running = 0;
    __syncthreads();
    int mine = 0;
    for (uint32_t p = threadIdx.x; p < nprob; p += blockDim.x) {
        double *s = sc + (size_t)p * BS_N;
        double rr = 0, rz = 0, bb = 0;
        for (uint32_t b = b0[p]; b < b0[p + 1]; b++) {
            rr += part[b];
            rz += part[nblk + b];
            bb += part[2 * nblk + b];
        }
        s[BS_RR] = rr;
        s[BS_RZ] = rz;
        s[BS_BB] = bb;
        s[BS_THR] = std::fmax(tol_abs, tol_rel * std::sqrt(bb));
        s[BS_ITER] = 0;
        s[BS_ALPHA] = s[BS_BETA] = 0;
        s[BS_CONV] = (std::sqrt(rr) <= s[BS_THR]) ? 1.0 : 0.0;
        mine |= s[BS_CONV] == 0.0;
    }
    if (mine) atomicOr(&running, 1);
    __syncthreads();
    if (threadIdx.x == 0) all[0] = running ? 0.0 : 1.0;
}

// Residual replacement per problem (cg_replace_residual of solver.hip, one
// problem at a time): rep[2p] = s, the stopped solve's scale over the new
// one (0: the problem restarts), rep[2p + 1] = the r.z of its last
// direction; part: the r.psave partials of each block.  Sets BS_B1 = r.z /
// (s rz_dir) and BS_C = r.psave / r.r for the problems that continue, 0
// for the others.
__global__ void k_brep_fin(const double *__restrict__ part, const uint32_t *__restrict__ b0, uint32_t nprob,
                           double *__restrict__ sc, const double *__restrict__ rep) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nprob; p += gridDim.x * blockDim.x) {
        double *s = sc + (size_t)p * BS_N;
        s[BS_B1] = 0;
        s[BS_C] = 0;
        const double sp = rep[2 * (size_t)p], rzd = rep[2 * (size_t)p + 1];
        if (!(sp > 0) || s[BS_CONV] != 0.0) continue;
        double rp = 0;
        for (uint32_t b = b0[p]; b < b0[p + 1]; b++) rp += part[b];
        const double rr = s[BS_RR], rz = s[BS_RZ];
        if (!(rr > 0) || !(rz > 0) || !(rzd > 0) || !isfinite(rp) || !isfinite(sp * rzd)) continue;
        s[BS_B1] = rz / (sp * rzd);
        s[BS_C] = rp / rr;
    }
}

// p += b1 (psave - c r) for the problems that continue (p holds z = M r
// after the start, psave the stopped solve's last direction)
template <typename T>
__global__ __launch_bounds__(TR_BLK_THREADS) void k_brep_p(T *__restrict__ p, const T *__restrict__ psave,
                                                           const T *__restrict__ r, const uint32_t *__restrict__ lo,
                                                           const uint32_t *__restrict__ hi,
                                                           const uint32_t *__restrict__ prob,
                                                           const double *__restrict__ sc) {
    const double *s = sc + (size_t)prob[blockIdx.x] * BS_N;
    const double b1 = s[BS_B1], c = s[BS_C];
    if (b1 == 0.0) return;
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS)
        p[i] = (T)((double)p[i] + b1 * ((double)psave[i] - c * (double)r[i]));
}

// dst = (float)(src * s_p);  x += s_p * (double)d   (per-problem factors)
__global__ __launch_bounds__(TR_BLK_THREADS) void k_bscale_f32(const double *__restrict__ src, float *__restrict__ dst,
                                                               const double *__restrict__ fac,
                                                               const uint32_t *__restrict__ lo,
                                                               const uint32_t *__restrict__ hi,
                                                               const uint32_t *__restrict__ prob) {
    const double f = fac[prob[blockIdx.x]];
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS)
        dst[i] = (float)(src[i] * f);
}
__global__ __launch_bounds__(TR_BLK_THREADS) void k_badd_f32(double *__restrict__ x, const float *__restrict__ d,
                                                             const double *__restrict__ fac,
                                                             const uint32_t *__restrict__ lo,
                                                             const uint32_t *__restrict__ hi,
                                                             const uint32_t *__restrict__ prob) {
    const double f = fac[prob[blockIdx.x]];
    if (f == 0.0) return;
    for (uint32_t i = lo[blockIdx.x] + threadIdx.x; i < hi[blockIdx.x]; i += TR_BLK_THREADS)
        x[i] += f * (double)d[i];
}
// r = b - r
__global__ __launch_bounds__(TR_BLK_THREADS) void k_bresid(const double *__restrict__ b, double *__restrict__ r,
                                                           size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        r[i] = b[i] - r[i];
}

// --------------------------------------------------------------------------
// host side

namespace {

struct DevMem {
    std::vector<void *> v;
    ~DevMem() {
        for (void *p : v) hipFree(p);
    }
    template <typename T>
    T *get(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
        v.push_back(p);
        return (T *)p;
    }
};

int build_blocks(const tr_ctx *c, DevMem &m, BlockTable &t, hipStream_t st) {
    const uint32_t P = (uint32_t)c->voff.size() - 1, slm = c->op.slm;
    std::vector<uint32_t> lo, hi, prob, b0(P + 1);
    for (uint32_t p = 0; p < P; p++) {
        b0[p] = (uint32_t)lo.size();
        const uint64_t e0 = (uint64_t)c->voff[p] * slm, e1 = (uint64_t)c->voff[p + 1] * slm;
        for (uint64_t e = e0; e < e1; e += TR_BLK_ENTRIES) {
            lo.push_back((uint32_t)e);
            hi.push_back((uint32_t)std::min<uint64_t>(e + TR_BLK_ENTRIES, e1));
            prob.push_back(p);
        }
    }
    b0[P] = (uint32_t)lo.size();
    t.nblk = (uint32_t)lo.size();
    t.nprob = P;
    if (!(t.lo = m.get<uint32_t>(t.nblk)) || !(t.hi = m.get<uint32_t>(t.nblk)) ||
        !(t.prob = m.get<uint32_t>(t.nblk)) || !(t.b0 = m.get<uint32_t>(P + 1)))
        return fail(TR_ERR_HIP, "batch: out of device memory");
    HIPCHECK(hipMemcpyAsync(t.lo, lo.data(), 4 * lo.size(), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(t.hi, hi.data(), 4 * hi.size(), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(t.prob, prob.data(), 4 * prob.size(), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(t.b0, b0.data(), 4 * b0.size(), hipMemcpyHostToDevice, st));
    // the host tables are local: the copies must have read them
    HIPCHECK(hipStreamSynchronize(st));
    return TR_OK;
}

struct BatchCg {
    int maxit = 1000, chunk = 16;
    double tol_abs = 0, tol_rel = 1e-8;
};

// Batched PCG from the x held on entry (T = double with the FP64 operator,
// float with the FP32 one).  sc: nprob*BS_N doubles; part >= 3*nblk.
// With cheb (z: its output buffer) the preconditioner is the Chebyshev
// polynomial over the union: one interval [lambda_max / ratio, lambda_max]
// with lambda_max of the whole union's D^-1 A bounds every problem's
// spectrum, so it is a valid SPD preconditioner for each of them.
// The operator launches of the iterations (and the Chebyshev steps, whose
// arguments the caller gives the same fields) skip the workgroups of
// stopped problems: a.vprob and a.pstop = sc + BS_CONV.  sc is cleared
// first, so the start's launches (before k_bfin_init) run everywhere.
// With psave and rep (residual replacement; b holds each problem's new
// residual and x is zero), the problems rep names continue the direction
// saved in psave (k_brep_fin, k_brep_p) instead of starting from p = z.
template <typename T>
int pcg_batch(tr_ctx *c, const BlockTable &t, const OpArgs<T> &a, const LaneMap &map, T *x, const T *b, T *r, T *p,
              T *Ap, const T *dinv, double *part, double *sc, double *all, const BatchCg &prm, hipStream_t st,
              int &launches, const ChebPrec<T> *cheb = nullptr, T *z = nullptr, const T *psave = nullptr,
              const double *rep = nullptr) {
    const int NC = (int)c->op.NC;
    const dim3 gb(t.nblk), tb(TR_BLK_THREADS);
    if (cheb && !z) return fail(TR_ERR_ARG, "batch: Chebyshev preconditioner without its buffer");
    HIPCHECK(hipMemsetAsync(sc, 0, sizeof(double) * (size_t)t.nprob * BS_N, st));
    HIPCHECK(launch_apply<T>(a, NC, map, x, r, nullptr, nullptr, st));
    hipLaunchKernelGGL(k_binit<T>, gb, tb, 0, st, b, r, p, dinv, t.lo, t.hi, part, t.nblk);
    if (cheb) {
        HIPCHECK(cheb->apply(r, z, nullptr, st));
        hipLaunchKernelGGL(k_bdot<T>, gb, tb, 0, st, r, z, t.lo, t.hi, part + t.nblk, nullptr);
    }
    hipLaunchKernelGGL(k_bfin_init, dim3(1), tb, 0, st, part, t.b0, t.nprob, t.nblk, sc, prm.tol_abs, prm.tol_rel,
                       all);
    if (cheb)  // p = z (beta = 0 after k_bfin_init)
        hipLaunchKernelGGL(k_bpupdate<T>, gb, tb, 0, st, p, r, dinv, t.lo, t.hi, t.prob, sc, all, (const T *)z);
    if (psave && rep) {
        hipLaunchKernelGGL(k_bdot<T>, gb, tb, 0, st, (const T *)r, psave, t.lo, t.hi, part, (const double *)nullptr,
                           (const double *)sc, (const uint32_t *)t.prob);
        hipLaunchKernelGGL(k_brep_fin, dim3((t.nprob + 255) / 256), dim3(256), 0, st, part, t.b0, t.nprob, sc, rep);
        hipLaunchKernelGGL(k_brep_p<T>, gb, tb, 0, st, p, psave, (const T *)r, t.lo, t.hi, t.prob,
                           (const double *)sc);
    }
    HIPCHECK(hipGetLastError());
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    OpArgs<T> ai = a;
    ai.vprob = c->d_vprob;
    ai.pstop = ai.vprob ? sc + BS_CONV : nullptr;
    ai.pstride = BS_N;
    for (int it = 0; it < prm.chunk; it++) {
        launch_apply<T>(ai, NC, map, p, Ap, nullptr, all, st);
        hipLaunchKernelGGL(k_bdot<T>, gb, tb, 0, st, p, Ap, t.lo, t.hi, part, all, (const double *)sc,
                           (const uint32_t *)t.prob);
        hipLaunchKernelGGL(k_bfin_alpha, dim3((t.nprob + 255) / 256), dim3(256), 0, st, part, t.b0, t.nprob, sc, all);
        hipLaunchKernelGGL(k_bupdate<T>, gb, tb, 0, st, x, p, r, Ap, cheb ? (const T *)nullptr : dinv, t.lo, t.hi,
                           t.prob, sc, part, t.nblk, all, cheb ? cheb->dinv : (const T *)nullptr,
                           cheb ? (T)cheb->inv_theta : (T)0, cheb ? cheb->first_buffer(z) : (T *)nullptr);
        if (cheb) {
            cheb->steps(r, z, all, st, nullptr);
            hipLaunchKernelGGL(k_bdot<T>, gb, tb, 0, st, r, z, t.lo, t.hi, part + t.nblk, all, (const double *)sc,
                               (const uint32_t *)t.prob);
        }
        hipLaunchKernelGGL(k_bfin_beta, dim3(1), tb, 0, st, part, t.b0, t.nprob, t.nblk, sc, prm.maxit, all);
        hipLaunchKernelGGL(k_bpupdate<T>, gb, tb, 0, st, p, r, dinv, t.lo, t.hi, t.prob, sc, all,
                           cheb ? (const T *)z : (const T *)nullptr);
    }
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(graph);
        return fail(TR_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(e));
    }
    double flag = 0;
    launches = 0;
    for (;;) {
        e = hipMemcpyAsync(&flag, all, sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess || flag != 0.0) break;
        if ((long)launches * prm.chunk > (long)prm.maxit + prm.chunk) break;  // the per-problem caps stop earlier
        if ((e = hipGraphLaunch(exec, st)) != hipSuccess) break;
        launches++;
    }
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("batch solve: ") + hipGetErrorString(e));
    return TR_OK;
}

}  // namespace

}  // namespace tr

using namespace tr;

extern "C" {

int tr_batch_create(int device, int nprob, const tr_mesh_view *meshes, const tr_angular_view *ang,
                    const double *albedo, int flags, tr_ctx **out) {
    if (nprob < 1 || !meshes || !ang || !out) return fail(TR_ERR_ARG, "tr_batch_create: bad argument");
    if (flags & ~TR_CTX_GENERAL_NORMALS) return fail(TR_ERR_ARG, "tr_batch_create: unknown flags");
    *out = nullptr;
    std::vector<HostOperator> ops((size_t)nprob);
    std::vector<int> rcs((size_t)nprob, 0);
    // the problems' stencils are independent: assemble them on the host cores
    const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; k++)
        th.emplace_back([&, k]() {
            for (int p = (int)k; p < nprob; p += (int)nt) {
                ops[p].general_normals = (flags & TR_CTX_GENERAL_NORMALS) != 0;
                rcs[p] = assemble_operator(meshes[p], *ang, ops[p]);
            }
        });
    for (auto &t : th) t.join();
    for (int p = 0; p < nprob; p++)
        if (rcs[p]) return fail(rcs[p], "tr_batch_create: problem " + std::to_string(p) + ": " + tr_last_error());
    std::vector<double> alb((size_t)nprob, 0.0);
    if (albedo)
        for (int p = 0; p < nprob; p++) {
            if (!(albedo[p] >= 0.0 && albedo[p] < 1.0)) return fail(TR_ERR_ARG, "tr_batch_create: albedo in [0, 1)");
            alb[p] = albedo[p];
        }
    tr_ctx *c = new tr_ctx();
    c->device = device;
    int rc = concat_operators(ops, alb, c->op, c->voff);
    ops.clear();
    if (rc) {
        delete c;
        return rc;
    }
    if ((rc = ctx_init(c, device))) return rc;
    *out = c;
    return TR_OK;
}

int tr_batch_info(const tr_ctx *c, uint32_t *nprob, uint32_t *voff) {
    if (!c || !nprob) return fail(TR_ERR_ARG, "tr_batch_info: null argument");
    const uint32_t P = c->voff.empty() ? 1 : (uint32_t)c->voff.size() - 1;
    *nprob = P;
    if (voff) {
        if (c->voff.empty()) {
            voff[0] = 0;
            voff[1] = c->op.nP;
        } else {
            std::memcpy(voff, c->voff.data(), sizeof(uint32_t) * (P + 1));
        }
    }
    return TR_OK;
}

int tr_batch_solve(tr_ctx *c, const tr_solve_params *prm, double *x, tr_solve_stats *total,
                   tr_batch_stats *per) {
    if (!c || !prm || !x) return fail(TR_ERR_ARG, "tr_batch_solve: null argument");
    if (c->voff.empty()) c->voff = {0, c->op.nP};
    if (prm->solver != TR_SOLVER_PCG_FP64 && prm->solver != TR_SOLVER_MIXED)
        return fail(TR_ERR_ARG, "tr_batch_solve: TR_SOLVER_PCG_FP64 or TR_SOLVER_MIXED");
    if (prm->use_prec != 0 && prm->use_prec != 2 && prm->use_prec != 4)
        return fail(TR_ERR_ARG, "tr_batch_solve: use_prec 0 (none), 2 (Jacobi) or 4 (Chebyshev)");
    const bool mixed = prm->solver == TR_SOLVER_MIXED;
    int rc;
    if (mixed && (rc = ensure_fp32(c))) return rc;
    hipStream_t st = c->stream;
    const size_t n = c->nint();
    const uint32_t P = (uint32_t)c->voff.size() - 1;
    DevMem m;
    BlockTable t;
    if ((rc = build_blocks(c, m, t, st))) return rc;
    double *b = m.get<double>(n), *r = m.get<double>(n), *p = m.get<double>(n), *Ap = m.get<double>(n);
    double *part = m.get<double>(3 * (size_t)t.nblk + 16), *sc = m.get<double>((size_t)P * BS_N + 16);
    double *all = m.get<double>(2), *fac = m.get<double>(P);
    if (!c->d_vprob) {
        // problem of every vertex, for the operator's skipping of stopped
        // problems (kept with the context)
        std::vector<uint32_t> vp(c->op.nP);
        for (uint32_t q = 0; q < P; q++)
            for (uint32_t v = c->voff[q]; v < c->voff[q + 1]; v++) vp[v] = q;
        if ((rc = upload(c, &c->d_vprob, vp))) return rc;
    }
    double *dinv = nullptr;
    float *dinv32 = nullptr;
    if (!b || !r || !p || !Ap || !part || !sc || !all || !fac) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
    auto t0 = std::chrono::steady_clock::now();
    ChebPrec<double> cheb;
    ChebPrec<float> cheb32;
    double *zc = nullptr;
    float *zc32 = nullptr;
    if (prm->use_prec == 2 || prm->use_prec == 4) {
        std::vector<double> d = operator_diagonal(c->op);
        for (double &v : d) v = 1.0 / v;
        if (!(dinv = m.get<double>(n))) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
        HIPCHECK(hipMemcpyAsync(dinv, d.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        if (mixed) {
            std::vector<float> d32(d.begin(), d.end());
            if (!(dinv32 = m.get<float>(n))) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
            HIPCHECK(hipMemcpyAsync(dinv32, d32.data(), sizeof(float) * n, hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));  // d32 is a temporary
        }
        HIPCHECK(hipStreamSynchronize(st));  // so is d
        if (prm->use_prec == 4) {
            // one lambda_max for the union bounds every problem's D^-1 A
            if ((rc = cheb_lambda(c, d))) return rc;
            if (!mixed) {
                cheb.a = c->args();
                cheb.NC = (int)c->op.NC;
                cheb.map = c->map;
                cheb.dinv = dinv;
                cheb.n = n;
                cheb.zb = m.get<double>(n);
                cheb.tmp = m.get<double>(n);
                zc = m.get<double>(n);
                if (!zc) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
                cheb.setup(c->cheb_deg, c->cheb_lmax, c->cheb_ratio);
                cheb.a.vprob = c->d_vprob;
                cheb.a.pstop = sc + BS_CONV;
                cheb.a.pstride = BS_N;
            } else {
                cheb32.a = c->args32();
                if (c->cheb_split > 0 && c->cheb_split < cheb32.a.nsplit &&
                    (c->map32.variant == 4 || c->map32.variant == 5) && c->map32.mfma_mode == 2)
                    cheb32.a.nsplit = c->cheb_split;
                cheb32.NC = (int)c->op.NC;
                cheb32.map = c->map32;
                cheb32.dinv = dinv32;
                cheb32.n = n;
                cheb32.zb = m.get<float>(n);
                cheb32.tmp = m.get<float>(n);
                zc32 = m.get<float>(n);
                if (!zc32) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
                cheb32.setup(c->cheb_deg, c->cheb_lmax, c->cheb_ratio);
                cheb32.a.vprob = c->d_vprob;
                cheb32.a.pstop = sc + BS_CONV;
                cheb32.a.pstride = BS_N;
            }
        }
    }
    HIPCHECK(launch_rhs(c->d_rhs, c->op.nP, c->op.slm, b, st));
    HIPCHECK(hipMemsetAsync(x, 0, sizeof(double) * n, st));
    std::vector<double> scv((size_t)P * BS_N);
    std::vector<int> outer(P, 0), inner(P, 0), done(P, 0);
    std::vector<double> bnorm(P, 0.0), rnorm(P, 0.0);
    int launches = 0;
    BatchCg cg;
    cg.maxit = prm->maxit;
    cg.tol_abs = prm->tol_abs;
    cg.tol_rel = prm->tol_rel;
    if (!mixed) {
        if ((rc = pcg_batch<double>(c, t, c->args(), c->map, x, b, r, p, Ap, dinv, part, sc, all, cg, st, launches,
                                    zc ? &cheb : nullptr, zc)))
            return rc;
        HIPCHECK(hipMemcpyAsync(scv.data(), sc, sizeof(double) * scv.size(), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        for (uint32_t q = 0; q < P; q++) {
            outer[q] = (int)scv[(size_t)q * BS_N + BS_ITER];
            done[q] = scv[(size_t)q * BS_N + BS_CONV] == 1.0;
        }
    } else {
        // FP64 iterative refinement per problem around the batched FP32 CG
        float *x32 = m.get<float>(n), *b32 = m.get<float>(n), *r32 = m.get<float>(n), *p32 = m.get<float>(n),
              *Ap32 = m.get<float>(n);
        if (!Ap32) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
        // residual replacement (as tr_solve's): each problem whose inner CG
        // stopped converged, and whose true residual has not drifted from
        // the recursive one, continues its last direction (saved in psave)
        float *psave = c->ir_reliable ? m.get<float>(n) : nullptr;
        double *rep = c->ir_reliable ? m.get<double>(2 * (size_t)P) : nullptr;
        if (c->ir_reliable && (!psave || !rep)) return fail(TR_ERR_HIP, "tr_batch_solve: out of device memory");
        std::vector<double> reph(2 * (size_t)P, 0.0), scale_prev(P, 0.0), rr_rec(P, 0.0), rz_dir(P, 0.0);
        std::vector<char> resumable(P, 0);
        BatchCg icg;
        icg.maxit = prm->inner_maxit;
        icg.tol_rel = prm->inner_tol;
        std::vector<double> f(P);
        const OpArgs<float> a32 = c->args32();
        for (int it = 0; it <= prm->maxit; it++) {
            // r = b - A64 x, per-problem norms
            HIPCHECK(launch_apply<double>(c->args(), (int)c->op.NC, c->map, x, r, nullptr, nullptr, st));
            hipLaunchKernelGGL(k_bresid, dim3(1024), dim3(256), 0, st, b, r, n);
            hipLaunchKernelGGL(k_bdot<double>, dim3(t.nblk), dim3(TR_BLK_THREADS), 0, st, r, r, t.lo, t.hi, part,
                               nullptr);
            if (it == 0)
                hipLaunchKernelGGL(k_bdot<double>, dim3(t.nblk), dim3(TR_BLK_THREADS), 0, st, b, b, t.lo, t.hi,
                                   part + t.nblk, nullptr);
            std::vector<double> hp(2 * (size_t)t.nblk);
            std::vector<uint32_t> b0(P + 1);
            HIPCHECK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipMemcpyAsync(b0.data(), t.b0, sizeof(uint32_t) * (P + 1), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            bool all_done = true;
            for (uint32_t q = 0; q < P; q++) {
                double rr = 0, bb = 0;
                for (uint32_t k = b0[q]; k < b0[q + 1]; k++) {
                    rr += hp[k];
                    bb += hp[t.nblk + k];
                }
                rnorm[q] = std::sqrt(rr);
                if (it == 0) bnorm[q] = std::sqrt(bb);
                done[q] = rnorm[q] <= std::fmax(prm->tol_abs, prm->tol_rel * bnorm[q]);
                all_done = all_done && done[q];
                f[q] = done[q] ? 0.0 : 1.0 / rnorm[q];
            }
            if (all_done || it == prm->maxit) break;
            for (uint32_t q = 0; q < P; q++) outer[q] += !done[q];
            // inner: A32 d = r / |r| per problem (zero for converged ones)
            // f is rewritten below: the copy must have read it (synchronous)
            HIPCHECK(hipMemcpyAsync(fac, f.data(), sizeof(double) * P, hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
            hipLaunchKernelGGL(k_bscale_f32, dim3(t.nblk), dim3(TR_BLK_THREADS), 0, st, r, b32, fac, t.lo, t.hi, t.prob);
            HIPCHECK(hipMemsetAsync(x32, 0, sizeof(float) * n, st));
            bool any_rep = false;
            if (psave) {
                for (uint32_t q = 0; q < P; q++) {
                    const double s_q = rnorm[q] > 0 ? scale_prev[q] / rnorm[q] : 0.0;
                    const bool go = !done[q] && resumable[q] && s_q > 0 &&
                                    1.0 / s_q <= kReplaceDrift * std::sqrt(rr_rec[q]);
                    reph[2 * (size_t)q] = go ? s_q : 0.0;
                    reph[2 * (size_t)q + 1] = go ? rz_dir[q] : 0.0;
                    any_rep = any_rep || go;
                }
                if (any_rep) {
                    // reph is rewritten next step: a synchronous copy
                    HIPCHECK(hipMemcpy(rep, reph.data(), sizeof(double) * reph.size(), hipMemcpyHostToDevice));
                    HIPCHECK(hipMemcpyAsync(psave, p32, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
                }
            }
            int l = 0;
            if ((rc = pcg_batch<float>(c, t, a32, c->map32, x32, b32, r32, p32, Ap32, dinv32, part, sc, all, icg, st,
                                       l, zc32 ? &cheb32 : nullptr, zc32, any_rep ? (const float *)psave : nullptr,
                                       any_rep ? (const double *)rep : nullptr)))
                return rc;
            launches += l;
            HIPCHECK(hipMemcpyAsync(scv.data(), sc, sizeof(double) * scv.size(), hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            for (uint32_t q = 0; q < P; q++) {
                const double *sq = scv.data() + (size_t)q * BS_N;
                inner[q] += (int)sq[BS_ITER];
                f[q] = done[q] ? 0.0 : rnorm[q];
                // converged with a direction: the next step may continue it
                resumable[q] = !done[q] && sq[BS_CONV] == 1.0 && sq[BS_ITER] > 0 && sq[BS_RZD] > 0;
                scale_prev[q] = rnorm[q];
                rr_rec[q] = sq[BS_RR];
                rz_dir[q] = sq[BS_RZD];
            }
            HIPCHECK(hipMemcpyAsync(fac, f.data(), sizeof(double) * P, hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
            hipLaunchKernelGGL(k_badd_f32, dim3(t.nblk), dim3(TR_BLK_THREADS), 0, st, x, x32, fac, t.lo, t.hi, t.prob);
        }
    }
    HIPCHECK(hipStreamSynchronize(st));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // true FP64 residuals per problem
    HIPCHECK(launch_apply<double>(c->args(), (int)c->op.NC, c->map, x, r, nullptr, nullptr, st));
    hipLaunchKernelGGL(k_bresid, dim3(1024), dim3(256), 0, st, b, r, n);
    hipLaunchKernelGGL(k_bdot<double>, dim3(t.nblk), dim3(TR_BLK_THREADS), 0, st, r, r, t.lo, t.hi, part, nullptr);
    hipLaunchKernelGGL(k_bdot<double>, dim3(t.nblk), dim3(TR_BLK_THREADS), 0, st, b, b, t.lo, t.hi, part + t.nblk,
                       nullptr);
    std::vector<double> hp(2 * (size_t)t.nblk);
    std::vector<uint32_t> b0(P + 1);
    HIPCHECK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(b0.data(), t.b0, sizeof(uint32_t) * (P + 1), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    int nconv = 0, maxo = 0, sumi = 0;
    double worst = 0, rtot = 0, btot = 0;
    for (uint32_t q = 0; q < P; q++) {
        double rr = 0, bb = 0;
        for (uint32_t k = b0[q]; k < b0[q + 1]; k++) {
            rr += hp[k];
            bb += hp[t.nblk + k];
        }
        const double rel = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        const bool ok = done[q] != 0;  // the problem's own stopping test (CG flag, or the FP64 refinement check)
        nconv += ok;
        worst = std::fmax(worst, rel);
        rtot += rr;
        btot += bb;
        maxo = std::max(maxo, outer[q]);
        sumi += inner[q];
        if (per) {
            per[q].iters = outer[q];
            per[q].inner_iters = inner[q];
            per[q].true_rel = rel;
            per[q].status = ok ? 0 : TR_ERR_NOCONV;
        }
    }
    if (total) {
        total->iters = maxo;
        total->preck = sumi;
        total->rnorm = std::sqrt(rtot);
        total->bnorm = std::sqrt(btot);
        total->true_rel = worst;
        total->seconds = secs;
        total->status = nconv == (int)P ? 0 : TR_ERR_NOCONV;
    }
    return TR_OK;
}

}  // extern "C"
