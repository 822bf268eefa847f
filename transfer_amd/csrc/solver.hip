// solver.hip -- device-resident preconditioned CG.
//
// The reference drives CG from the host (main.cpp:125-162): every dot
// product and norm ends in a blocking device-to-host copy of one REAL
// (CudaContext.cpp:175-198).  Here alpha, beta and the residual norm stay
// in a small device array; one iteration is three launches
//   apply A p (+ per-workgroup p.Ap)
//   alpha = rz / p.Ap; r -= alpha Ap, z = M r (+ r.r, r.z partials)
//   beta = rz_new / rz, convergence test; x += alpha p, p = z + beta p
// (with a general preconditioner z = M r runs between the last two, with
// its own r.z pass).  Every workgroup of the two vector launches sums the
// previous launch's partials itself, in one fixed order, so all agree on
// alpha and beta without a separate reduction launch; r.z alternates
// between two scalar slots by iteration parity, so workgroup 0 can record
// the new value while the others still read the old one.  The iterations
// are captured once into a HIP graph of `chunk` (even) iterations; the host
// reads the convergence flag once per graph launch.  After convergence
// every kernel of the remaining chunk returns at its first instruction.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "solver.h"

namespace tr {

#define TR_WAVE 64
#define TR_UPD_BLOCKS 1024

__device__ inline double wave_sum2(double s) {
#pragma unroll
    for (int off = TR_WAVE / 2; off > 0; off >>= 1) s += __shfl_down(s, off, TR_WAVE);
    return s;
}

// Two workgroup sums at once; results valid in thread 0.
__device__ inline void block_sum2(double &a, double &b) {
    __shared__ double w[2][16];
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    a = wave_sum2(a);
    b = wave_sum2(b);
    if (lane == 0) {
        w[0][wave] = a;
        w[1][wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0, t = 0;
        for (int i = 0; i < (int)(blockDim.x / TR_WAVE); i++) {
            s += w[0][i];
            t += w[1][i];
        }
        a = s;
        b = t;
    }
}

// sum of part[0..n) and part[n2..n2+n) in a fixed order (one workgroup)
__device__ inline void sum_parts(const double *part, int n, int n2, double &a, double &b) {
    a = 0;
    b = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        a += part[i];
        if (n2 >= 0) b += part[n2 + i];
    }
    block_sum2(a, b);
}

// r -= alpha Ap, z = M r and the r.r, r.z partials, alpha = rz / p.Ap from
// the nb p.Ap partials of the operator launch; workgroup 0 records p.Ap,
// alpha and whether the pass ran (SC_RUN).  With x != null also x +=
// alpha p (else the p update pass does it, reading p there anyway).
// With z1 != null also z1 = cscale cdinv r (the first Chebyshev iterate of
// a polynomial preconditioner) from the new r.  The old r and Ap are read
// non-temporally (neither is read again this iteration), so the caches
// keep the new r for the p update pass.
template <typename T>
__global__ __launch_bounds__(256) void k_update_fused(T *__restrict__ x, const T *__restrict__ p,
                                                      T *__restrict__ r, const T *__restrict__ Ap,
                                                      const T *__restrict__ dinv, size_t n,
                                                      double *__restrict__ sc, const double *__restrict__ pap_part,
                                                      int nb, int rzs, double *__restrict__ part,
                                                      const T *__restrict__ cdinv, T cscale, T *__restrict__ z1) {
    if (sc[SC_CONV] != 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) sc[SC_RUN] = 0.0;
        return;
    }
    __shared__ double al_sh;
    {
        double pap, dummy;
        sum_parts(pap_part, nb, -1, pap, dummy);
        if (threadIdx.x == 0) {
            al_sh = sc[rzs] / pap;
            if (blockIdx.x == 0) {
                sc[SC_PAP] = pap;
                sc[SC_ALPHA] = al_sh;
                sc[SC_RUN] = 1.0;
            }
        }
        __syncthreads();
    }
    const T alpha = (T)al_sh;
    double rr = 0, rz = 0;
    auto one = [&](size_t i) __attribute__((always_inline)) {
        if (x) x[i] = fma(alpha, p[i], x[i]);
        const T ri = fma(-alpha, Ap[i], r[i]);
        r[i] = ri;
        if (z1) z1[i] = cscale * cdinv[i] * ri;
        const T zi = dinv ? dinv[i] * ri : ri;
        rr = fma((double)ri, (double)ri, rr);
        rz = fma((double)ri, (double)zi, rz);
    };
    const size_t S = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "FP32 or FP64 vectors");
    if constexpr (sizeof(T) == 8) {
        // FP64: 16-byte accesses, two elements per thread and step
        typedef double d2 __attribute__((ext_vector_type(2)));
        const size_t n2 = n / 2;
        d2 *x2 = reinterpret_cast<d2 *>(x), *r2 = reinterpret_cast<d2 *>(r);
        const d2 *p2 = reinterpret_cast<const d2 *>(p), *a2 = reinterpret_cast<const d2 *>(Ap);
        const d2 *dd2 = reinterpret_cast<const d2 *>(dinv), *c2 = reinterpret_cast<const d2 *>(cdinv);
        d2 *z2 = reinterpret_cast<d2 *>(z1);
        for (size_t k = i; k < n2; k += S) {
            if (x) x2[k] = alpha * p2[k] + x2[k];
            const d2 rv = __builtin_nontemporal_load(&r2[k]) - alpha * __builtin_nontemporal_load(&a2[k]);
            r2[k] = rv;
            if (z1) z2[k] = cscale * c2[k] * rv;
            const d2 zv = dinv ? dd2[k] * rv : rv;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                rr = fma(rv[u], rv[u], rr);
                rz = fma(rv[u], zv[u], rz);
            }
        }
        for (size_t k = 2 * n2 + i; k < n; k += S) one(k);
    } else if constexpr (sizeof(T) == 4) {
        // FP32: 16-byte accesses, four elements per thread and step
        typedef float f4 __attribute__((ext_vector_type(4)));
        const size_t n4 = n / 4;
        f4 *x4 = reinterpret_cast<f4 *>(x), *r4 = reinterpret_cast<f4 *>(r);
        const f4 *p4 = reinterpret_cast<const f4 *>(p), *a4 = reinterpret_cast<const f4 *>(Ap);
        const f4 *d4 = reinterpret_cast<const f4 *>(dinv), *c4 = reinterpret_cast<const f4 *>(cdinv);
        f4 *z4 = reinterpret_cast<f4 *>(z1);
        for (size_t k = i; k < n4; k += S) {
            if (x) x4[k] = alpha * p4[k] + x4[k];
            const f4 rv = __builtin_nontemporal_load(&r4[k]) - alpha * __builtin_nontemporal_load(&a4[k]);
            r4[k] = rv;
            if (z1) z4[k] = cscale * c4[k] * rv;
            const f4 zv = dinv ? d4[k] * rv : rv;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                rr = fma((double)rv[u], (double)rv[u], rr);
                rz = fma((double)rv[u], (double)zv[u], rz);
            }
        }
        for (size_t k = 4 * n4 + i; k < n; k += S) one(k);
    }
    block_sum2(rr, rz);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[TR_UPD_BLOCKS + blockIdx.x] = rz;
    }
}

// beta = rz_new / rz from the G r.r / r.z partials (rz in slot rz_old, the
// new value into slot rz_new by workgroup 0, with the iteration count and
// the convergence flag); with x != null x += alpha p; p = z + beta p, z =
// M r (dinv) or z given (rr partials part[0 .. G), r.z partials
// part[TR_UPD_BLOCKS .. TR_UPD_BLOCKS + Grz)).  It runs when the
// iteration's update pass ran (SC_RUN, which no workgroup of this launch
// writes), so every workgroup updates x, also in the iteration that stops
// the solve.  r, the old p and x go through non-temporal accesses, so the
// caches keep the new p for the operator that reads it next.
template <typename T>
__global__ __launch_bounds__(256) void k_pupdate_fused(T *__restrict__ p, const T *__restrict__ r,
                                                       const T *__restrict__ dinv, const T *__restrict__ z,
                                                       T *__restrict__ x, size_t n, double *__restrict__ sc,
                                                       const double *__restrict__ part, int G, int Grz, int rz_old,
                                                       int rz_new, int maxit, int stall) {
    if (sc[SC_RUN] == 0.0) return;
    __shared__ double be_sh, al_sh;
    __shared__ int stop_sh;
    {
        double rr, rz, dummy;
        if (Grz == G) {
            sum_parts(part, G, TR_UPD_BLOCKS, rr, rz);
        } else {
            sum_parts(part, G, -1, rr, dummy);
            __syncthreads();
            sum_parts(part + TR_UPD_BLOCKS, Grz, -1, rz, dummy);
        }
        if (threadIdx.x == 0) {
            be_sh = rz / sc[rz_old];
            al_sh = sc[SC_ALPHA];
            // converged or broken down: every workgroup reaches this from
            // the same sums and from scalars no workgroup of this launch
            // writes, so none of them updates p, and the stopped solve
            // keeps its last direction whole (a resumed solve continues
            // from it)
            stop_sh = !(rr == rr) || !(rz == rz) || !(sc[SC_PAP] > 0.0) || rr <= sc[SC_THR];
            if (blockIdx.x == 0) {
                sc[SC_RR] = rr;
                sc[SC_BETA] = be_sh;
                sc[rz_new] = rz;
                sc[SC_ITER] += 1.0;
                const double it = sc[SC_ITER];
                if (rr < sc[SC_RRMIN]) sc[SC_RRMIN] = rr;
                if (!(rr == rr) || !(rz == rz) || !(sc[SC_PAP] > 0.0))
                    sc[SC_CONV] = 3.0;  // breakdown: NaN, or p.Ap <= 0
                else if (rr <= sc[SC_THR])
                    sc[SC_CONV] = 1.0;  // converged
                else if (it >= (double)maxit)
                    sc[SC_CONV] = 2.0;  // iteration cap
                else if (stall > 0 && fmod(it, (double)stall) == 0.0) {
                    if (sc[SC_RRMIN] > 0.81 * sc[SC_RRWIN]) sc[SC_CONV] = 4.0;  // stagnation
                    sc[SC_RRWIN] = sc[SC_RRMIN];
                }
            }
        }
        __syncthreads();
    }
    const T beta = (T)be_sh, alpha = (T)al_sh;
    const bool pupd = !stop_sh;
    if (!pupd && !x) return;
    auto one = [&](size_t i) __attribute__((always_inline)) {
        const T pi = p[i];
        if (x) x[i] = fma(alpha, pi, x[i]);
        if (pupd) {
            const T zi = z ? z[i] : (dinv ? dinv[i] * r[i] : r[i]);
            p[i] = fma(beta, pi, zi);
        }
    };
    const size_t S = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if constexpr (sizeof(T) == 8) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const size_t n2 = n / 2;
        d2 *p2 = reinterpret_cast<d2 *>(p), *x2 = reinterpret_cast<d2 *>(x);
        const d2 *r2 = reinterpret_cast<const d2 *>(r), *dd2 = reinterpret_cast<const d2 *>(dinv);
        const d2 *zz2 = reinterpret_cast<const d2 *>(z);
        for (size_t k = i; k < n2; k += S) {
            const d2 pv = __builtin_nontemporal_load(&p2[k]);
            if (x) __builtin_nontemporal_store(alpha * pv + __builtin_nontemporal_load(&x2[k]), &x2[k]);
            if (pupd) {
                const d2 rv = __builtin_nontemporal_load(&r2[k]);
                const d2 zv = z ? zz2[k] : (dinv ? dd2[k] * rv : rv);
                p2[k] = beta * pv + zv;
            }
        }
        for (size_t k = 2 * n2 + i; k < n; k += S) one(k);
    } else if constexpr (sizeof(T) == 4) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const size_t n4 = n / 4;
        f4 *p4 = reinterpret_cast<f4 *>(p), *x4 = reinterpret_cast<f4 *>(x);
        const f4 *r4 = reinterpret_cast<const f4 *>(r), *d4 = reinterpret_cast<const f4 *>(dinv);
        const f4 *zz4 = reinterpret_cast<const f4 *>(z);
        for (size_t k = i; k < n4; k += S) {
            const f4 pv = __builtin_nontemporal_load(&p4[k]);
            if (x) __builtin_nontemporal_store(alpha * pv + __builtin_nontemporal_load(&x4[k]), &x4[k]);
            if (pupd) {
                const f4 rv = __builtin_nontemporal_load(&r4[k]);
                const f4 zv = z ? zz4[k] : (dinv ? d4[k] * rv : rv);
                p4[k] = beta * pv + zv;
            }
        }
        for (size_t k = 4 * n4 + i; k < n; k += S) one(k);
    }
}

// r = b - r (r holds A x on entry); z = M r; p = z; partials of r.r, r.z, b.b
template <typename T>
__global__ __launch_bounds__(256) void k_init(const T *__restrict__ b, T *__restrict__ r,
                                              T *__restrict__ p, const T *__restrict__ dinv,
                                              size_t n, double *__restrict__ part) {
    double rr = 0, rz = 0, bb = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const T ri = b[i] - r[i];
        r[i] = ri;
        const T zi = dinv ? dinv[i] * ri : ri;
        p[i] = zi;
        rr = fma((double)ri, (double)ri, rr);
        rz = fma((double)ri, (double)zi, rz);
        bb = fma((double)b[i], (double)b[i], bb);
    }
    block_sum2(rr, rz);
    double bz = 0;
    __syncthreads();
    block_sum2(bb, bz);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[TR_UPD_BLOCKS + blockIdx.x] = rz;
        part[2 * TR_UPD_BLOCKS + blockIdx.x] = bb;
    }
}

// r.z partials (into part, the r.z slots of k_update / k_init), and p = z
// when p is given
template <typename T>
__global__ __launch_bounds__(256) void k_rz(const T *__restrict__ r, const T *__restrict__ z, T *__restrict__ p,
                                            size_t n, double *__restrict__ part, const double *__restrict__ sc) {
    if (sc && sc[SC_CONV] != 0.0) return;
    double rz = 0, dummy = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        rz = fma((double)r[i], (double)z[i], rz);
        if (p) p[i] = z[i];
    }
    block_sum2(rz, dummy);
    if (threadIdx.x == 0) part[blockIdx.x] = rz;
}

// Residual replacement: partials of r.r, r.z and r.p for a new residual r
// and the stopped solve's last direction p; z = M r given, or dinv r, or r
template <typename T>
__global__ __launch_bounds__(256) void k_replace_dots(const T *__restrict__ r, const T *__restrict__ z,
                                                      const T *__restrict__ dinv, const T *__restrict__ p,
                                                      size_t n, double *__restrict__ part) {
    double rr = 0, rz = 0, rp = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double ri = (double)r[i];
        const double zi = z ? (double)z[i] : (dinv ? (double)(dinv[i] * r[i]) : ri);
        rr = fma(ri, ri, rr);
        rz = fma(ri, zi, rz);
        rp = fma(ri, (double)p[i], rp);
    }
    block_sum2(rr, rz);
    double dummy = 0;
    __syncthreads();
    block_sum2(rp, dummy);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = rr;
        part[TR_UPD_BLOCKS + blockIdx.x] = rz;
        part[2 * TR_UPD_BLOCKS + blockIdx.x] = rp;
    }
}

// p = b1 p + z - cr r  (z = M r given, or dinv r, or r)
template <typename T>
__global__ __launch_bounds__(256) void k_replace_p(T *__restrict__ p, const T *__restrict__ r,
                                                   const T *__restrict__ z, const T *__restrict__ dinv, size_t n,
                                                   T b1, T cr) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T zi = z ? z[i] : (dinv ? dinv[i] * r[i] : r[i]);
        p[i] = fma(b1, p[i], fma(-cr, r[i], zi));
    }
}

// sc from the init partials: rr, rz, bb, and the stopping threshold
__global__ __launch_bounds__(256) void k_fin_init(const double *__restrict__ part, int n,
                                                  double *__restrict__ sc, double tol_abs,
                                                  double tol_rel) {
    double rr, rz, bb, dummy;
    sum_parts(part, n, TR_UPD_BLOCKS, rr, rz);
    __syncthreads();
    sum_parts(part + 2 * TR_UPD_BLOCKS, n, -1, bb, dummy);
    if (threadIdx.x == 0) {
        sc[SC_RR] = rr;
        sc[SC_RZ] = rz;
        sc[SC_BB] = bb;
        double thr = tol_abs * tol_abs;
        if (tol_rel > 0) thr = fmax(thr, tol_rel * tol_rel * bb);
        sc[SC_THR] = thr;
        sc[SC_ITER] = 0;
        sc[SC_RRMIN] = rr;
        sc[SC_RRWIN] = rr;
        sc[SC_CONV] = (rr <= thr) ? 1.0 : 0.0;
    }
}

static unsigned upd_grid(size_t n) {
    return (unsigned)std::min<size_t>(std::max<size_t>((n + 1023) / 1024, 1), TR_UPD_BLOCKS);
}

template <typename T>
int pcg_device(const CgSpace<T> &S, const CgParams &prm, hipStream_t st, CgResult &res) {
    const size_t n = S.n;
    const unsigned G = upd_grid(n);
    // dot partials written by the operator launch (the angular-diagonal
    // kernel is never persistent); with faces off the axes the boundary
    // term comes after the operator kernel, so p.Ap is a pass of its own
    const bool fdot = S.a.ngv == 0;
    const uint32_t nb = !fdot ? G : (prm.diag_op ? apply_tiles(S.a.nP, S.map) : apply_blocks(S.a.nP, S.map));
    auto t0 = std::chrono::steady_clock::now();
    hipError_t e;
#define HC(x)                                                                               \
    do {                                                                                    \
        e = (x);                                                                            \
        if (e != hipSuccess) return fail(TR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e)); \
    } while (0)
    const bool gen = (bool)S.prec;
    if (!prm.resume) {
        // r = b - A x ; z = M r ; p = z
        if (prm.diag_op)
            HC(launch_apply_diag<T>(S.a, S.NC, S.map, S.x, S.r, nullptr, nullptr, st));
        else
            HC(launch_apply<T>(S.a, S.NC, S.map, S.x, S.r, nullptr, nullptr, st));
        hipLaunchKernelGGL(k_init<T>, dim3(G), dim3(256), 0, st, S.b, S.r, S.p, gen ? nullptr : S.dinv, n, S.part);
        if (gen) {
            HC(S.prec(S.r, S.z, nullptr));
            hipLaunchKernelGGL(k_rz<T>, dim3(G), dim3(256), 0, st, S.r, S.z, S.p, n, S.part + TR_UPD_BLOCKS,
                               nullptr);
        }
        hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(256), 0, st, S.part, (int)G, S.sc, prm.tol_abs, prm.tol_rel);
        HC(hipGetLastError());
    }

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = prm.graph ? *prm.graph : nullptr;
    // an even number of iterations per graph: iteration k reads r.z from
    // slot k % 2 and writes slot (k + 1) % 2, and every launch starts at 0
    const int chunk = std::max(2, prm.chunk + (prm.chunk & 1));
    if (!exec) {  // capture the iteration graph (once, when the caller keeps it)
        HC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        // the first error of a launch made during the capture (the graph would
        // silently lack that node otherwise)
        hipError_t ce = hipSuccess;
        auto keep = [&ce](hipError_t x) {
            if (ce == hipSuccess && x != hipSuccess) ce = x;
        };
        for (int it = 0; it < chunk; it++) {
            const int rz_old = (it & 1) ? SC_RZ2 : SC_RZ, rz_new = (it & 1) ? SC_RZ : SC_RZ2;
            if (prm.diag_op)
                keep(launch_apply_diag<T>(S.a, S.NC, S.map, S.p, S.Ap, fdot ? S.part : nullptr, S.sc + SC_CONV, st));
            else
                keep(launch_apply<T>(S.a, S.NC, S.map, S.p, S.Ap, fdot ? S.part : nullptr, S.sc + SC_CONV, st));
            if (!fdot)
                hipLaunchKernelGGL(k_rz<T>, dim3(G), dim3(256), 0, st, (const T *)S.p, (const T *)S.Ap, (T *)nullptr, n,
                                   S.part, S.sc);
            // a polynomial preconditioner's first iterate comes out of the
            // update pass, and its last step may write the r.z partials itself
            const bool first = gen && S.prec_loop && S.pre_out;
            hipLaunchKernelGGL(k_update_fused<T>, dim3(G), dim3(256), 0, st, (T *)nullptr, S.p, S.r, S.Ap,
                               gen ? nullptr : S.dinv, n, S.sc, S.part, (int)nb, rz_old, S.part2,
                               first ? S.pre_dinv : (const T *)nullptr, S.pre_scale, first ? S.pre_out : (T *)nullptr);
            keep(hipGetLastError());
            int Grz = (int)G;
            if (gen) {
                if (first) {
                    keep(S.prec_loop(S.r, S.z, S.sc + SC_CONV, S.part2 + TR_UPD_BLOCKS));
                    if (S.prec_nparts) Grz = S.prec_nparts;
                } else {
                    keep(S.prec(S.r, S.z, S.sc + SC_CONV));
                }
                if (!(first && S.prec_nparts))
                    hipLaunchKernelGGL(k_rz<T>, dim3(G), dim3(256), 0, st, S.r, S.z, (T *)nullptr, n,
                                       S.part2 + TR_UPD_BLOCKS, S.sc);
            }
            hipLaunchKernelGGL(k_pupdate_fused<T>, dim3(G), dim3(256), 0, st, S.p, S.r, gen ? nullptr : S.dinv,
                               gen ? (const T *)S.z : nullptr, S.x, n, S.sc, S.part2, (int)G, Grz, rz_old, rz_new,
                               prm.maxit, prm.stall_window);
            keep(hipGetLastError());
        }
        e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        if (ce != hipSuccess) {
            hipGraphDestroy(graph);
            return fail(TR_ERR_HIP, std::string("graph capture: a captured launch failed: ") + hipGetErrorString(ce));
        }
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(TR_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        if (prm.graph) *prm.graph = exec;
    }
    double sc[SC_N];
    int launches = 0;
    for (;;) {
        HC(hipMemcpyAsync(sc, S.sc, sizeof(sc), hipMemcpyDeviceToHost, st));
        HC(hipStreamSynchronize(st));
        if (sc[SC_CONV] != 0.0) break;
        if (launches * chunk > prm.maxit + chunk) break;  // defensive: the flag caps earlier
        HC(hipGraphLaunch(exec, st));
        launches++;
    }
    if (!prm.graph) hipGraphExecDestroy(exec);
    auto t1 = std::chrono::steady_clock::now();
    res.iters = (int)sc[SC_ITER];
    res.rr = sc[SC_RR];
    res.bb = sc[SC_BB];
    res.flag = (int)sc[SC_CONV];
    // iteration k (from 0) read the r.z of its direction from slot k % 2
    res.rz_dir = res.iters > 0 ? sc[((res.iters - 1) & 1) ? SC_RZ2 : SC_RZ] : 0.0;
    res.seconds = std::chrono::duration<double>(t1 - t0).count();
    return TR_OK;
#undef HC
}

template <typename T>
int cg_replace_residual(const CgSpace<T> &S, const CgParams &prm, double s, double rz_dir, hipStream_t st,
                        bool &ok) {
    ok = false;
    const size_t n = S.n;
    const unsigned G = upd_grid(n);
    hipError_t e;
#define HC(x)                                                                               \
    do {                                                                                    \
        e = (x);                                                                            \
        if (e != hipSuccess) return fail(TR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e)); \
    } while (0)
    const bool gen = (bool)S.prec;
    if (gen) HC(S.prec(S.r, S.z, nullptr));
    hipLaunchKernelGGL(k_replace_dots<T>, dim3(G), dim3(256), 0, st, (const T *)S.r, gen ? (const T *)S.z : nullptr,
                       gen ? nullptr : S.dinv, (const T *)S.p, n, S.part);
    HC(hipGetLastError());
    std::vector<double> part(3 * TR_UPD_BLOCKS);
    HC(hipMemcpyAsync(part.data(), S.part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    HC(hipStreamSynchronize(st));
    double rr = 0, rz = 0, rp = 0;
    for (unsigned i = 0; i < G; i++) {
        rr += part[i];
        rz += part[TR_UPD_BLOCKS + i];
        rp += part[2 * TR_UPD_BLOCKS + i];
    }
    if (!(rr > 0) || !(rz > 0) || !(rz_dir > 0) || !(s > 0) || !std::isfinite(rr) || !std::isfinite(rz) ||
        !std::isfinite(rp) || !std::isfinite(s * rz_dir))
        return TR_OK;
    // the direction in the new scale is s p; take out its component along
    // the new r (zero for the recursive residual), then p = z + beta (s p)
    // with beta = r.z / (s^2 rz_dir)
    const double b1 = rz / (s * rz_dir), c = rp / rr;
    hipLaunchKernelGGL(k_replace_p<T>, dim3(G), dim3(256), 0, st, S.p, (const T *)S.r, gen ? (const T *)S.z : nullptr,
                       gen ? nullptr : S.dinv, n, (T)b1, (T)(b1 * c));
    HC(hipGetLastError());
    double sc[SC_N] = {};
    sc[SC_RZ] = rz;
    sc[SC_RR] = rr;
    sc[SC_BB] = rr;
    double thr = prm.tol_abs * prm.tol_abs;
    if (prm.tol_rel > 0) thr = std::max(thr, prm.tol_rel * prm.tol_rel * rr);
    sc[SC_THR] = thr;
    sc[SC_RRMIN] = rr;
    sc[SC_RRWIN] = rr;
    HC(hipMemcpyAsync(S.sc, sc, sizeof(sc), hipMemcpyHostToDevice, st));
    HC(hipStreamSynchronize(st));
    ok = true;
    return TR_OK;
#undef HC
}

template int pcg_device<double>(const CgSpace<double> &, const CgParams &, hipStream_t, CgResult &);
template int pcg_device<float>(const CgSpace<float> &, const CgParams &, hipStream_t, CgResult &);
template int cg_replace_residual<float>(const CgSpace<float> &, const CgParams &, double, double, hipStream_t, bool &);

}  // namespace tr

namespace tr {

// FP64 internal vectors <-> FP32 vectors whose harmonics at each vertex
// may be in another order: position k of a vertex's FP32 row holds
// harmonic hp[k] of its FP64 row (hp null: the same order).
__global__ void k_to_f32_scaled(const double *__restrict__ src, float *__restrict__ dst, double s, size_t n,
                                const uint16_t *__restrict__ hp, uint32_t slm) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t j = i;
        if (hp) {
            const size_t v = i / slm, k = i - v * slm;
            j = v * slm + hp[k];
        }
        dst[i] = (float)(src[j] * s);
    }
}

__global__ void k_add_f32_scaled(double *__restrict__ x, const float *__restrict__ d, double s, size_t n,
                                 const uint16_t *__restrict__ hp, uint32_t slm) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t j = i;
        if (hp) {
            const size_t v = i / slm, k = i - v * slm;
            j = v * slm + hp[k];
        }
        x[j] = fma(s, (double)d[i], x[j]);
    }
}

hipError_t launch_to_f32_scaled(const double *src, float *dst, double s, size_t n, hipStream_t st,
                                const uint16_t *hp, uint32_t slm) {
    if (hp && !slm) return hipErrorInvalidValue;
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_f32_scaled, dim3(g), dim3(256), 0, st, src, dst, s, n, hp, slm);
    return hipGetLastError();
}

hipError_t launch_add_f32_scaled(double *x, const float *d, double s, size_t n, hipStream_t st,
                                 const uint16_t *hp, uint32_t slm) {
    if (hp && !slm) return hipErrorInvalidValue;
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_add_f32_scaled, dim3(g), dim3(256), 0, st, x, d, s, n, hp, slm);
    return hipGetLastError();
}

}  // namespace tr
