// kernels.hip -- gfx950 kernels of the matrix-free operator, BLAS-1 and the
// FP64 reductions.
//
// Operator kernels (the roles of volumePart / surfacePart / *Diag in
// kernels.cu:200-493): a wavefront holds VPW vertices with L lanes each, a
// lane owning J consecutive harmonics.  Phase 1 gathers the ELL neighbours'
// harmonic slabs (contiguous, so the L lanes of a vertex read one coalesced
// run) and accumulates the spatial channels in registers.  Phase 2 stages the
// stiffness channels of the wave's vertices in LDS and applies the sparse
// angular couplings lane by lane from there.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.h"

namespace tr {

#define TR_WAVE 64

#define TR_RED_BLOCKS 1024

__device__ inline double wave_sum(double s) {
#pragma unroll
    for (int off = TR_WAVE / 2; off > 0; off >>= 1) s += __shfl_down(s, off, TR_WAVE);
    return s;
}

// Sum over the workgroup (wavefront shuffles, then the wavefronts' totals in
// a fixed order); the result is valid in thread 0.
__device__ inline double block_sum(double s) {
    __shared__ double w[16];
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    s = wave_sum(s);
    if (lane == 0) w[wave] = s;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x / TR_WAVE); i++) t += w[i];
    }
    return t;
}

// ---------------------------------------------------------------------------
// r = massfac * sum_k M_ik f_k + sum_c Omega_c (sum_k K^c_ik f_k)
//   + sum_n O_n (sum_k B^n_ik f_k)                       (boundary vertices)
template <typename T, int J, int NC, bool DOT>
__global__ __launch_bounds__(256) void k_apply(OpArgs<T> a, const T *__restrict__ f,
                                               T *__restrict__ r, double *__restrict__ part,
                                               const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const int vsub = lane / a.L, lv = lane - vsub * a.L;
    const uint32_t v = (blockIdx.x * (uint32_t)wpb + wave) * (uint32_t)a.VPW + vsub;
    const bool valid = vsub < a.VPW && v < a.nP;
    const int lm0 = lv * J;
    const size_t nP = a.nP, nB = a.nB;
    const int b = valid ? a.bof[v] : -1;
    T acc[NC][J], x0[J], bacc[3][J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        x0[j] = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c][j] = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) bacc[c][j] = 0;
    }
    if (valid) {
        for (uint32_t s = 0; s < a.W; s++) {
            const uint32_t k = a.nbr[s * nP + v];
            const T *fk = f + (size_t)k * a.slm + lm0;
            T x[J];
#pragma unroll
            for (int j = 0; j < J; j++) x[j] = fk[j];
            if (s == 0) {
#pragma unroll
                for (int j = 0; j < J; j++) x0[j] = x[j];
            }
            const T *cs = a.coef + (size_t)s * NC * nP + v;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const T cf = cs[(size_t)c * nP];
#pragma unroll
                for (int j = 0; j < J; j++) acc[c][j] = fma(cf, x[j], acc[c][j]);
            }
        }
        if (b >= 0) {
            for (uint32_t s = 0; s < a.WB; s++) {
                const uint32_t k = a.bnbr[s * nB + b];
                const T *fk = f + (size_t)k * a.slm + lm0;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T cf = a.bcoef[((size_t)s * 3 + c) * nB + b];
#pragma unroll
                    for (int j = 0; j < J; j++) bacc[c][j] = fma(cf, fk[j], bacc[c][j]);
                }
            }
        }
    }
    T *my = lds + (size_t)(wave * a.VPW + vsub) * (NC - 1 + 3) * a.slm;
    if (valid) {
#pragma unroll
        for (int c = 1; c < NC; c++)
#pragma unroll
            for (int j = 0; j < J; j++) my[(c - 1) * a.slm + lm0 + j] = acc[c][j];
        if (b >= 0) {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int j = 0; j < J; j++) my[(NC - 1 + c) * a.slm + lm0 + j] = bacc[c][j];
        }
    }
    __syncthreads();
    double d = 0;
    if (valid) {
        T R[J];
#pragma unroll
        for (int j = 0; j < J; j++)
            R[j] = ((a.mf0 && lm0 + j == 0) ? a.mf0[v] * a.massfac[0] : a.massfac[lm0 + j]) * acc[0][j];
        for (uint32_t e = 0; e < a.arow; e++) {
            const size_t q = (size_t)e * a.slm + lm0;
#pragma unroll
            for (int j = 0; j < J; j++) R[j] = fma(a.aval[q + j], my[a.aidx[q + j]], R[j]);
        }
        if (b >= 0) {
            const T *mb = my + (NC - 1) * a.slm;
            for (uint32_t e = 0; e < a.brow; e++) {
                const size_t q = (size_t)e * a.slm + lm0;
#pragma unroll
                for (int j = 0; j < J; j++) R[j] = fma(a.bval[q + j], mb[a.bidx[q + j]], R[j]);
            }
        }
        T *rv = r + (size_t)v * a.slm + lm0;
#pragma unroll
        for (int j = 0; j < J; j++) {
            rv[j] = R[j];
            if (DOT) d = fma((double)x0[j], (double)R[j], d);
        }
    }
    if (DOT) {
        d = block_sum(d);
        if (threadIdx.x == 0) part[blockIdx.x] = d;
    }
}

// Workgroup i is dispatched to XCD i % 8; give every XCD a contiguous range of
// logical blocks so that neighbouring vertices share that XCD's L2.
__device__ inline uint32_t xcd_block(uint32_t bid, uint32_t nblk) {
    const uint32_t X = 8, per = nblk / X, rem = nblk % X;
    const uint32_t xcd = bid % X, k = bid / X;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}


// Operator, second schedule.  Gathers: the ELL slots are unrolled to TR_WMAX
// so every neighbour index is loaded up front and the slab loads can be in
// flight together.  Angular contraction: instead of staging all of a
// wavefront's vertices at once, CB = 64/slm vertices at a time are staged in
// LDS ([vertex][channel][lm]) and contracted with lanes remapped to
// (vertex, output lm), so a wavefront needs CB*(NC+3)*slm words of LDS, not
// VPW*(NC+2)*slm; the coupling table sits in LDS, loaded once per workgroup.
template <typename T, int J, int NC, bool DOT>
__global__ __launch_bounds__(256) void k_apply2(OpArgs<T> a, const T *__restrict__ f,
                                                T *__restrict__ r, double *__restrict__ part,
                                                const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t slm = a.slm;
    const int CH = NC + 3;                       // mass, NC-1 stiffness, 3 boundary
    const int CB = slm >= TR_WAVE ? 1 : TR_WAVE / (int)slm;
    // LDS: table values, table indices, then per-wave staging
    T *tval = reinterpret_cast<T *>(smem_raw);
    const uint32_t tsz = (a.arow + a.brow) * slm;
    uint16_t *tidx = reinterpret_cast<uint16_t *>(tval + tsz);
    T *stage = reinterpret_cast<T *>(smem_raw + ((tsz * (sizeof(T) + 2) + 15) & ~(size_t)15)) +
               (size_t)wave * CB * CH * slm;
    for (uint32_t i = threadIdx.x; i < a.arow * slm; i += blockDim.x) {
        tval[i] = a.aval[i];
        tidx[i] = a.aidx[i];
    }
    for (uint32_t i = threadIdx.x; i < a.brow * slm; i += blockDim.x) {
        tval[a.arow * slm + i] = a.bval[i];
        tidx[a.arow * slm + i] = a.bidx[i];
    }
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int vsub = lane / a.L, lv = lane - vsub * a.L;
    const uint32_t v0 = (blk * (uint32_t)wpb + wave) * (uint32_t)a.VPW;
    const uint32_t v = v0 + vsub;
    const bool valid = vsub < a.VPW && v < a.nP;
    const int lm0 = lv * J;
    const size_t nP = a.nP, nB = a.nB;
    const int b = valid ? a.bof[v] : -1;
    T acc[NC][J], bacc[3][J];
#pragma unroll
    for (int j = 0; j < J; j++) {
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c][j] = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) bacc[c][j] = 0;
    }
    if (valid) {
        uint32_t nb[TR_WMAX];
#pragma unroll
        for (int s = 0; s < TR_WMAX; s++) nb[s] = s < (int)a.W ? a.nbr[s * nP + v] : 0;
#pragma unroll
        for (int s = 0; s < TR_WMAX; s++) {
            if (s < (int)a.W) {
                const T *fk = f + (size_t)nb[s] * slm + lm0;
                T x[J];
#pragma unroll
                for (int j = 0; j < J; j++) x[j] = fk[j];
                const T *cs = a.coef + (size_t)s * NC * nP + v;
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    const T cf = cs[(size_t)c * nP];
#pragma unroll
                    for (int j = 0; j < J; j++) acc[c][j] = fma(cf, x[j], acc[c][j]);
                }
            }
        }
        if (b >= 0) {
            for (uint32_t s = 0; s < a.WB; s++) {
                const uint32_t k = a.bnbr[s * nB + b];
                const T *fk = f + (size_t)k * slm + lm0;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T cf = a.bcoef[((size_t)s * 3 + c) * nB + b];
#pragma unroll
                    for (int j = 0; j < J; j++) bacc[c][j] = fma(cf, fk[j], bacc[c][j]);
                }
            }
        }
    }
    double d = 0;
    const int nbatch = (a.VPW + CB - 1) / CB;
    for (int q = 0; q < nbatch; q++) {
        __syncthreads();
        const int vb = vsub - q * CB;
        if (valid && vb >= 0 && vb < CB) {
            T *my = stage + (size_t)vb * CH * slm;
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < J; j++) my[c * slm + lm0 + j] = acc[c][j];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int j = 0; j < J; j++) my[(NC + c) * slm + lm0 + j] = bacc[c][j];
        }
        __syncthreads();
        // contraction lanes: (vertex cv of this batch, output lm)
        for (int o = lane; o < CB * (int)slm; o += TR_WAVE) {
            const int cv = o / (int)slm, lm = o - cv * (int)slm;
            const uint32_t vv = v0 + q * CB + cv;
            if (q * CB + cv >= a.VPW || vv >= a.nP) continue;
            const T *my = stage + (size_t)cv * CH * slm;
            T R = ((a.mf0 && lm == 0) ? a.mf0[vv] * a.massfac[0] : a.massfac[lm]) * my[lm];
            const T *st = my + slm;  // stiffness channels
            for (uint32_t e = 0; e < a.arow; e++)
                R = fma(tval[e * slm + lm], st[tidx[e * slm + lm]], R);
            if (a.bof[vv] >= 0) {
                const T *bt = my + NC * slm;
                for (uint32_t e = 0; e < a.brow; e++)
                    R = fma(tval[(a.arow + e) * slm + lm], bt[tidx[(a.arow + e) * slm + lm]], R);
            }
            r[(size_t)vv * slm + lm] = R;
            if (DOT) d = fma((double)f[(size_t)vv * slm + lm], (double)R, d);
        }
    }
    if (DOT) {
        d = block_sum(d);
        if (threadIdx.x == 0) part[blockIdx.x] = d;
    }
}

// ---------------------------------------------------------------------------
// Buffer-load schedule.  Every array is read through a buffer resource
// (SGPR descriptor): the per-slot offset is a scalar, the per-lane offset a
// loop-invariant 32-bit VGPR, and reads past the end of a buffer return
// zero, so the slot loop is straight-line code over WT slots.
#define TR_RSRC_FLAGS 0x00020000

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const T *p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), (short)0,
                                             (int)(bytes > 0xffffffffull ? 0xffffffffull : bytes),
                                             TR_RSRC_FLAGS);
}

typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// N consecutive elements of type T at byte offset voff + soff.  WIDE: the
// offset is a multiple of 16 bytes, so 16-byte loads may be used.  Loaded
// vectors are reinterpreted whole (bit-casting single vector elements
// miscompiles).
template <typename T, int N, bool WIDE>
__device__ inline void bload(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff, T *out) {
    if constexpr (sizeof(T) == 8) {
        int q = 0;
        if constexpr (WIDE) {
#pragma unroll
            for (; q + 1 < N; q += 2) {
                f64x2 v = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + q * 8, soff, 0));
                out[q] = v.x;
                out[q + 1] = v.y;
            }
        }
#pragma unroll
        for (; q < N; q++)
            out[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff + q * 8, soff, 0));
    } else {
        int q = 0;
        if constexpr (WIDE) {
#pragma unroll
            for (; q + 3 < N; q += 4) {
                f32x4 v = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + q * 4, soff, 0));
                out[q] = v.x;
                out[q + 1] = v.y;
                out[q + 2] = v.z;
                out[q + 3] = v.w;
            }
        }
#pragma unroll
        for (; q < N; q++)
            out[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff + q * 4, soff, 0));
    }
}

// EPI: 0 = r = A f; 1 = the same and the fused dot f . r; 2 = one Chebyshev
// step fused into the write, r = zc f - om r + ga dinv (rhs - A f) (r on
// entry holds the previous iterate, not read when om == 0)
template <typename T, int J, int NC, int WT, int EPI>
__global__ __launch_bounds__(256) void k_apply3(OpArgs<T> a, const T *__restrict__ f,
                                                T *__restrict__ r, double *__restrict__ part,
                                                const double *__restrict__ skip, ChebStep<T> cs) {
    constexpr bool DOT = EPI == 1;
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RS = rec_stride<T>(NC), TB = tab_rec_bytes<T>(), CH = NC + 3;
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t slm = a.slm, nP = a.nP;
    const int CB = slm >= TR_WAVE ? 1 : TR_WAVE / (int)slm;
    const uint32_t ntab = (a.arow + a.brow) * slm;
    // table records, then per-wavefront staging
    unsigned char *tab = smem_raw;
    T *stage = reinterpret_cast<T *>(smem_raw + (size_t)ntab * TB) + (size_t)wave * CB * CH * slm;
    {
        const i32x2 *src = reinterpret_cast<const i32x2 *>(a.tab);
        i32x2 *dst = reinterpret_cast<i32x2 *>(tab);
        for (uint32_t i = threadIdx.x; i < ntab * (TB / 8); i += blockDim.x) dst[i] = src[i];
    }
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    if (a.pstop) {
        // a workgroup whose vertices all belong to stopped problems has
        // nothing to do (the decision is the same for all its threads)
        const uint32_t vb0 = blk * (uint32_t)wpb * (uint32_t)a.VPW;
        const uint32_t vb1 = min(vb0 + (uint32_t)wpb * (uint32_t)a.VPW, nP);
        bool run = false;
        if (vb0 < nP)
            for (uint32_t p = a.vprob[vb0]; p <= a.vprob[vb1 - 1] && !run; p++) run = a.pstop[(size_t)p * a.pstride] == 0.0;
        if (!run) {
            if (DOT && threadIdx.x == 0) part[blockIdx.x] = 0;
            return;
        }
    }
    const int vsub = lane / a.L, lv = lane - vsub * a.L;
    const uint32_t v0 = (blk * (uint32_t)wpb + wave) * (uint32_t)a.VPW;
    const uint32_t v = v0 + vsub;
    const bool valid = vsub < a.VPW && v < nP;
    const int lm0 = lv * J;
    T acc[NC][J];
#pragma unroll
    for (int j = 0; j < J; j++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c][j] = 0;
    if (valid) {
        const auto frs = rsrc(f, (size_t)nP * slm * sizeof(T));
        const auto nrs = rsrc(a.nbr, (size_t)a.W * nP * 4);
        uint32_t nb[WT];
#pragma unroll
        for (int s = 0; s < WT; s++)
            nb[s] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(nrs, v * 4u, (uint32_t)s * nP * 4u, 0);
        const uint32_t cvoff = v * (uint32_t)(RS * sizeof(T));
#pragma unroll
        for (int s = 0; s < WT; s++) {
            T x[J], cf[RS];
            // J | slm, so a J whose slab is a multiple of 16 bytes keeps every
            // lane's offset 16-byte aligned; records are 16-byte multiples
            bload<T, J, (J * sizeof(T)) % 16 == 0>(frs, (nb[s] * slm + lm0) * (uint32_t)sizeof(T), 0u, x);
            // one descriptor per slot: a slot's records stay below 4 GiB
            // where all W of them need not (unions of many problems)
            const auto crs = rsrc(a.crec + (size_t)s * nP * RS, (size_t)nP * RS * sizeof(T));
            bload<T, RS, true>(crs, cvoff, 0u, cf);
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < J; j++) acc[c][j] = fma(cf[c], x[j], acc[c][j]);
        }
    }
    // vacuum boundary (few vertices): plain loads
    T bacc[3][J];
#pragma unroll
    for (int j = 0; j < J; j++) bacc[0][j] = bacc[1][j] = bacc[2][j] = 0;
    const int b = valid ? a.bof[v] : -1;
    if (b >= 0) {
        const size_t nB = a.nB;
        for (uint32_t s = 0; s < a.WB; s++) {
            const uint32_t k = a.bnbr[s * nB + b];
            const T *fk = f + (size_t)k * slm + lm0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T cf = a.bcoef[((size_t)s * 3 + c) * nB + b];
#pragma unroll
                for (int j = 0; j < J; j++) bacc[c][j] = fma(cf, fk[j], bacc[c][j]);
            }
        }
    }
    double d = 0;
    const int nbatch = (a.VPW + CB - 1) / CB;
    for (int q = 0; q < nbatch; q++) {
        __syncthreads();
        const int vb = vsub - q * CB;
        if (valid && vb >= 0 && vb < CB) {
            T *my = stage + (size_t)vb * CH * slm;
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < J; j++) my[c * slm + lm0 + j] = acc[c][j];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int j = 0; j < J; j++) my[(NC + c) * slm + lm0 + j] = bacc[c][j];
        }
        __syncthreads();
        for (int o = lane; o < CB * (int)slm; o += TR_WAVE) {
            const int cv = o / (int)slm, lm = o - cv * (int)slm;
            const uint32_t vv = v0 + q * CB + cv;
            if (q * CB + cv >= a.VPW || vv >= nP) continue;
            const unsigned char *my = reinterpret_cast<const unsigned char *>(stage + (size_t)cv * CH * slm);
            T R0 = ((a.mf0 && lm == 0) ? a.mf0[vv] * a.massfac[0] : a.massfac[lm]) * reinterpret_cast<const T *>(my)[lm],
              R1 = 0;
            const unsigned char *row = tab + (size_t)lm * TB;
            const uint32_t rstride = slm * TB;
            // rows are padded to multiples of 8 entries (zero couplings):
            // issue eight table reads, then eight staged reads, then the FMAs
            const uint32_t nent = a.arow + (a.bof[vv] >= 0 ? a.brow : 0);
            for (uint32_t e = 0; e < nent; e += 8) {
                T val[8];
                uint32_t off[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const unsigned char *rr = row + (e + u) * rstride;
                    val[u] = *reinterpret_cast<const T *>(rr);
                    off[u] = *reinterpret_cast<const uint32_t *>(rr + sizeof(T));
                }
                T t[8];
#pragma unroll
                for (int u = 0; u < 8; u++) t[u] = *reinterpret_cast<const T *>(my + off[u]);
#pragma unroll
                for (int u = 0; u < 8; u += 2) {
                    R0 = fma(val[u], t[u], R0);
                    R1 = fma(val[u + 1], t[u + 1], R1);
                }
            }
            const T R = R0 + R1;
            const size_t e = (size_t)vv * slm + lm;
            if constexpr (EPI == 2) {
                T nv = cs.zc * f[e] + cs.ga * cs.dinv[e] * (cs.rhs[e] - R);
                if (cs.om != (T)0) nv -= cs.om * r[e];
                r[e] = nv;
            } else {
                r[e] = R;
                if (DOT) d = fma((double)f[e], (double)R, d);
            }
        }
    }
    if (DOT) {
        d = block_sum(d);
        if (threadIdx.x == 0) part[blockIdx.x] = d;
    }
}

// angular-diagonal operator: per lane and slot the coupling
// massfac*M + sum_a Omega^aa[lm][lm] K^aa (+ sum_n O_n[lm][lm] B^n)
template <typename T, int J, bool DOT>
__global__ __launch_bounds__(256) void k_apply_diag(OpArgs<T> a, int NC, const T *__restrict__ f,
                                                    T *__restrict__ r, double *__restrict__ part,
                                                    const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const int vsub = lane / a.L, lv = lane - vsub * a.L;
    const uint32_t v = (blockIdx.x * (uint32_t)wpb + wave) * (uint32_t)a.VPW + vsub;
    const bool valid = vsub < a.VPW && v < a.nP;
    const int lm0 = lv * J;
    const size_t nP = a.nP, nB = a.nB;
    double d = 0;
    if (valid) {
        T mf[J], d0[J], d1[J], d2[J], acc[J], x0[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            mf[j] = a.massfac[lm0 + j];
            if (a.mf0 && lm0 + j == 0) mf[j] *= a.mf0[v];
            d0[j] = a.ddiag[lm0 + j];
            d1[j] = a.ddiag[a.slm + lm0 + j];
            d2[j] = a.ddiag[2 * a.slm + lm0 + j];
            acc[j] = 0;
            x0[j] = 0;
        }
        for (uint32_t s = 0; s < a.W; s++) {
            const uint32_t k = a.nbr[s * nP + v];
            const T *fk = f + (size_t)k * a.slm + lm0;
            const T *cs = a.coef + (size_t)s * NC * nP + v;
            const T c0 = cs[0], c1 = cs[nP], c2 = cs[2 * nP], c3 = cs[3 * nP];
#pragma unroll
            for (int j = 0; j < J; j++) {
                const T x = fk[j];
                if (s == 0) x0[j] = x;
                acc[j] = fma(mf[j] * c0 + d0[j] * c1 + d1[j] * c2 + d2[j] * c3, x, acc[j]);
            }
        }
        const int b = a.bof[v];
        if (b >= 0) {
#pragma unroll
            for (int j = 0; j < J; j++) {
                d0[j] = a.bdiag[lm0 + j];
                d1[j] = a.bdiag[a.slm + lm0 + j];
                d2[j] = a.bdiag[2 * a.slm + lm0 + j];
            }
            for (uint32_t s = 0; s < a.WB; s++) {
                const uint32_t k = a.bnbr[s * nB + b];
                const T *fk = f + (size_t)k * a.slm + lm0;
                const T c0 = a.bcoef[((size_t)s * 3 + 0) * nB + b];
                const T c1 = a.bcoef[((size_t)s * 3 + 1) * nB + b];
                const T c2 = a.bcoef[((size_t)s * 3 + 2) * nB + b];
#pragma unroll
                for (int j = 0; j < J; j++) acc[j] = fma(d0[j] * c0 + d1[j] * c1 + d2[j] * c2, fk[j], acc[j]);
            }
        }
        T *rv = r + (size_t)v * a.slm + lm0;
#pragma unroll
        for (int j = 0; j < J; j++) {
            rv[j] = acc[j];
            if (DOT) d = fma((double)x0[j], (double)acc[j], d);
        }
    }
    if (DOT) {
        d = block_sum(d);
        if (threadIdx.x == 0) part[blockIdx.x] = d;
    }
}

// ---------------------------------------------------------------------------
size_t apply_lds_bytes(const LaneMap &m, int NC, uint32_t slm, size_t elem) {
    return (size_t)m.WPB * m.VPW * (NC - 1 + 3) * slm * elem;
}

size_t apply2_lds_bytes(const LaneMap &m, int NC, uint32_t slm, uint32_t arow, uint32_t brow,
                        size_t elem) {
    const int CB = slm >= 64 ? 1 : 64 / (int)slm;
    size_t table = (((size_t)(arow + brow) * slm * (elem + 2)) + 15) & ~(size_t)15;
    return table + (size_t)m.WPB * CB * (NC + 3) * slm * elem;
}

size_t apply3_lds_bytes(const LaneMap &m, int NC, uint32_t slm, uint32_t arow, uint32_t brow,
                        size_t elem) {
    const int CB = slm >= 64 ? 1 : 64 / (int)slm;
    size_t table = (size_t)(arow + brow) * slm * (elem == 8 ? 16 : 8);
    return table + (size_t)m.WPB * CB * (NC + 3) * slm * elem;
}

uint32_t apply_tiles(uint32_t nP, const LaneMap &m) {
    uint32_t per = (uint32_t)(m.WPB * m.VPW);
    return (nP + per - 1) / per;
}

uint32_t apply_blocks(uint32_t nP, const LaneMap &m) {
    const uint32_t tiles = apply_tiles(nP, m);
    return ((m.variant == 4 || m.variant == 5) && m.grid_dot) ? std::min(m.grid_dot, tiles) : tiles;
}

// Specular reflection: the listed entries of r vanish (see stencil.h).
template <typename T>
__global__ __launch_bounds__(256) void k_zero_mask(T *__restrict__ r, const uint32_t *__restrict__ mask,
                                                   uint32_t n, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) r[mask[i]] = (T)0;
}

template <typename T>
hipError_t launch_zero_mask(const OpArgs<T> &a, T *r, const double *skip, hipStream_t s) {
    if (!a.nmask) return hipSuccess;
    const uint32_t g = std::min<uint32_t>((a.nmask + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero_mask<T>, dim3(g), dim3(256), 0, s, r, a.mask, a.nmask, skip);
    return hipGetLastError();
}

#define TR_J_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(11)

template <typename T>
hipError_t launch_apply(const OpArgs<T> &a, int NC, const LaneMap &m, const T *f, T *r,
                        double *part, const double *skip, hipStream_t s) {
    if (a.nP == 0) return hipSuccess;
    // a fused dot product would miss the faces off the axes
    if (part && a.ngv) return hipErrorInvalidValue;
    if (m.variant == 5) {
        if constexpr (sizeof(T) == 8) {
            hipError_t e = launch_apply5d(a, NC, m.WPB, part ? m.grid_dot : m.grid, f, r, part, skip, s);
            if (e == hipSuccess) e = launch_bgen(a, false, f, r, skip, s);
            return e != hipSuccess ? e : launch_zero_mask(a, r, skip, s);
        } else {
            hipError_t e = launch_apply5(a, NC, m.WPB, part ? m.grid_dot : m.grid, f, r, part, skip, s);
            if (e == hipSuccess) e = launch_bgen(a, false, f, r, skip, s);
            return e != hipSuccess ? e : launch_zero_mask(a, r, skip, s);
        }
    }
    if (m.variant == 4) {
        hipError_t e;
        if constexpr (sizeof(T) == 8)
            e = launch_apply4<T, 0>(a, NC, m.WPB, part ? m.grid_dot : m.grid, f, r, part, skip, s);
        else if (m.mfma_mode == 2)
            e = launch_apply4<T, 2>(a, NC, m.WPB, part ? m.grid_dot : m.grid, f, r, part, skip, s);
        else
            e = launch_apply4<T, 1>(a, NC, m.WPB, part ? m.grid_dot : m.grid, f, r, part, skip, s);
        if (e == hipSuccess) e = launch_bgen(a, false, f, r, skip, s);
        return e != hipSuccess ? e : launch_zero_mask(a, r, skip, s);
    }
    dim3 g(apply_blocks(a.nP, m)), b(TR_WAVE * m.WPB);
    // variant 3 needs W = 15 or 16 exactly (the host pads the stencil) so that
    // every unrolled slot reads inside the arrays
    const bool v3 = m.variant == 3 && (a.W == 15 || a.W == 16) && a.crec && a.tab;
    const bool v2 = !v3 && m.variant >= 2 && a.W <= TR_WMAX;
    size_t sh = v3 ? apply3_lds_bytes(m, NC, a.slm, a.arow, a.brow, sizeof(T))
                   : v2 ? apply2_lds_bytes(m, NC, a.slm, a.arow, a.brow, sizeof(T))
                        : apply_lds_bytes(m, NC, a.slm, sizeof(T));
#define TR_LAUNCH(JJ, NCC, D)                                                                   \
    do {                                                                                        \
        if (v3 && a.W == 15)                                                                    \
            hipLaunchKernelGGL((k_apply3<T, JJ, NCC, 15, D ? 1 : 0>), g, b, sh, s, a, f, r, part, skip, \
                               ChebStep<T>{});                                                  \
        else if (v3)                                                                            \
            hipLaunchKernelGGL((k_apply3<T, JJ, NCC, 16, D ? 1 : 0>), g, b, sh, s, a, f, r, part, skip, \
                               ChebStep<T>{});                                                  \
        else if (v2)                                                                            \
            hipLaunchKernelGGL((k_apply2<T, JJ, NCC, D>), g, b, sh, s, a, f, r, part, skip);      \
        else                                                                                    \
            hipLaunchKernelGGL((k_apply<T, JJ, NCC, D>), g, b, sh, s, a, f, r, part, skip);       \
    } while (0)
#define TR_CASE(JJ)                                                        \
    case JJ:                                                               \
        if (NC == 7) {                                                     \
            if (part) TR_LAUNCH(JJ, 7, true); else TR_LAUNCH(JJ, 7, false);   \
        } else {                                                           \
            if (part) TR_LAUNCH(JJ, 10, true); else TR_LAUNCH(JJ, 10, false); \
        }                                                                  \
        break;
    switch (m.J) {
        TR_J_CASES(TR_CASE)
        default: return hipErrorInvalidValue;
    }
#undef TR_CASE
#undef TR_LAUNCH
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_bgen(a, false, f, r, skip, s);
    return e != hipSuccess ? e : launch_zero_mask(a, r, skip, s);
}

// The vacuum faces off the axes (general normals): for the vertex v of
// block b, r_v += sum over its faces F of O(n_F) g_F, g_F = sum_{k in F}
// (S_F/12)(1 + delta_vk) f_k.  One workgroup per such vertex, a thread per
// harmonic, g_F staged in LDS.  O(n_F) g = R (O_z (R^T g)) with R the face's
// rotation (one block per even l, a block's harmonics contiguous) and O_z
// shared by all faces, so a face brings rotation_size values, not slm^2.
// DIAG: diag(O(n_F)) g, the Diag operator's boundary term.  cheb != 0: the
// operator kernel wrote a fused Chebyshev step into r, and this adds the
// faces' share of A f to it, r -= ga dinv (faces' A f), the step being
// linear in A f.
template <typename T, bool DIAG>
__global__ __launch_bounds__(128) void k_bgen(OpArgs<T> a, const T *__restrict__ f, T *__restrict__ r,
                                              const double *__restrict__ skip, ChebStep<T> cs, int cheb) {
    if (skip && *skip != 0.0) return;
    // dynamic LDS: O_z [slm][slm] when it takes at most 8 KB (else it is
    // read from global memory: staged per vertex it would cost occupancy),
    // then per face slot the face's rotation [nr], g [32|128] and y [32|128].
    // With slm <= 32 the 64 lanes take two faces at once (slot = lane / 32).
    extern __shared__ __align__(16) unsigned char bgen_smem[];
    const uint32_t slm = a.slm, b = blockIdx.x, nt = blockDim.x;
    const uint32_t ns = slm <= 32 ? 2u : 1u, hw = slm <= 32 ? 32u : 128u;  // slots, harmonics per slot
    const uint32_t slot = threadIdx.x / hw, t = threadIdx.x % hw, tps = nt / ns;  // tps: threads per slot
    const uint32_t v = a.gvert[b];
    // this thread's harmonic: block start s0 and size nb, block offset ro
    // in a face's packed rotation, rotation_size nr
    uint32_t s0 = 0, nb = 1, ro = 0, nr = 0;
    for (uint32_t l = 0, s = 0; s < slm; l += 2) {
        const uint32_t n = 2 * l + 1;
        if (t >= s && t < s + n) {
            s0 = s;
            nb = n;
            ro = nr;
        }
        s += n;
        nr += n * n;
    }
    const uint32_t p = t - s0;
    const bool ozl = !DIAG && (size_t)slm * slm * sizeof(T) <= 8192;
    T *Ozs = reinterpret_cast<T *>(bgen_smem);
    T *Rall = Ozs + (ozl ? (size_t)slm * slm : 0);
    T *gall = Rall + (DIAG ? 0 : (size_t)ns * nr);
    T *Rs = Rall + (DIAG ? 0 : (size_t)slot * nr), *g = gall + (size_t)slot * 2 * hw, *y = g + hw;
    if (ozl)
        for (uint32_t i = threadIdx.x; i < slm * slm; i += nt) Ozs[i] = a.gOz[i];
    const T *Ozp = ozl ? Ozs : a.gOz;
    const uint32_t j0 = a.gstart[b], j1 = a.gstart[b + 1];
    T acc = 0;
    for (uint32_t jb = j0; jb < j1; jb += ns) {
        const uint32_t j = jb + slot;
        const bool live = j < j1 && t < slm;
        uint32_t F = 0;
        if (j < j1) {
            F = a.gface[j];
            if (!DIAG) {
                const T *R = a.gR + (size_t)F * nr;
                for (uint32_t i = t; i < nr; i += tps) Rs[i] = R[i];
            }
        }
        if (live) {
            const uint32_t li = a.gpos[j];
            const T w = a.garea[F];
            T sum = 0;
            for (uint32_t k = 0; k < 3; k++) sum += (k == li ? 2 * w : w) * f[(size_t)a.gfv[3 * F + k] * slm + t];
            g[t] = sum;
        }
        __syncthreads();
        if (DIAG) {
            if (live) acc += a.gOd[(size_t)F * slm + t] * g[t];
        } else {
            const T *Rb = Rs + ro;  // this thread's block, [nb][nb]
            if (live) {  // y = R^T g
                T sum = 0;
                for (uint32_t c = 0; c < nb; c++) sum += Rb[c * nb + p] * g[s0 + c];
                y[t] = sum;
            }
            __syncthreads();
            if (live) {  // g = O_z y (g is free again)
                T sum = 0;
                const T *Oz = Ozp + (size_t)t * slm;
                for (uint32_t m = 0; m < slm; m++) sum += Oz[m] * y[m];
                g[t] = sum;
            }
            __syncthreads();
            if (live) {  // += R g
                T sum = 0;
                for (uint32_t c = 0; c < nb; c++) sum += Rb[p * nb + c] * g[s0 + c];
                acc += sum;
            }
        }
        __syncthreads();
    }
    // the slots' sums (slot 1 through the first g of slot 0's area)
    if (ns == 2) {
        if (slot == 1 && t < slm) gall[t] = acc;
        __syncthreads();
        if (slot == 0 && t < slm) acc += gall[t];
    }
    if (slot == 0 && t < slm) {
        const size_t o = (size_t)v * slm + t;
        if (cheb)
            r[o] -= cs.ga * cs.dinv[o] * acc;
        else
            r[o] += acc;
    }
}

template <typename T>
static hipError_t launch_bgen(const OpArgs<T> &a, bool diag, const T *f, T *r, const double *skip, hipStream_t s,
                              const ChebStep<T> *cs = nullptr) {
    if (!a.ngv) return hipSuccess;
    if (a.slm > 128 || (diag && cs)) return hipErrorInvalidValue;
    uint32_t nr = 0;
    for (uint32_t l = 0, q = 0; q < a.slm; l += 2) {
        q += 2 * l + 1;
        nr += (2 * l + 1) * (2 * l + 1);
    }
    // one wavefront per vertex up to 64 harmonics (two faces at once up to
    // 32); O_z, each slot's rotation and two vectors in LDS
    const unsigned nt = a.slm <= 64 ? 64 : 128;
    const size_t ns = a.slm <= 32 ? 2 : 1, hw = a.slm <= 32 ? 32 : 128;  // as the kernel
    const bool ozl = !diag && (size_t)a.slm * a.slm * sizeof(T) <= 8192;
    const size_t sh = sizeof(T) * ((ozl ? (size_t)a.slm * a.slm : 0) + (diag ? 0 : ns * nr) + ns * 2 * hw);
    if (sh > 65536) return hipErrorInvalidValue;
    const ChebStep<T> c = cs ? *cs : ChebStep<T>{};
    if (diag)
        hipLaunchKernelGGL((k_bgen<T, true>), dim3(a.ngv), dim3(nt), sh, s, a, f, r, skip, c, 0);
    else
        hipLaunchKernelGGL((k_bgen<T, false>), dim3(a.ngv), dim3(nt), sh, s, a, f, r, skip, c, cs ? 1 : 0);
    return hipGetLastError();
}

// out = zc z - om out + ga dinv (rhs - t), t = A z (the unfused Chebyshev
// step's elementwise pass)
template <typename T>
__global__ __launch_bounds__(256) void k_cheb_upd(const T *__restrict__ z, T *__restrict__ out,
                                                  const T *__restrict__ t, ChebStep<T> cs, size_t n,
                                                  const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        T nv = cs.zc * z[i] + cs.ga * cs.dinv[i] * (cs.rhs[i] - t[i]);
        if (cs.om != (T)0) nv -= cs.om * out[i];
        out[i] = nv;
    }
}

uint32_t cheb_parts(uint32_t nP, const LaneMap &m) {
    if (m.variant != 4 && m.variant != 5) return 0;
    const uint32_t per = 16u * (uint32_t)m.WPB, ntiles = (nP + per - 1) / per;
    return m.grid_cheb ? std::min(m.grid_cheb, ntiles) : ntiles;
}

template <typename T>
hipError_t launch_cheb_step(const OpArgs<T> &a, int NC, const LaneMap &m, const T *z, T *out,
                            const ChebStep<T> &cs, T *tmp, const double *skip, hipStream_t s, double *part) {
    if (a.nP == 0) return hipSuccess;
    hipError_t e;
    // (a fused dot product also covers masked entries: rhs, a CG residual,
    // is zero there).  Faces off the axes: k_bgen adds their share to the
    // fused step afterwards (the fused dot would miss it: not with part).
    if (part && a.ngv) return hipErrorInvalidValue;
    if (m.variant == 5) {
        if constexpr (sizeof(T) == 8) {
            e = launch_apply5d(a, NC, m.WPB, m.grid_cheb, z, out, part, skip, s, &cs);
            if (e == hipSuccess) e = launch_bgen(a, false, z, out, skip, s, &cs);
            return e != hipSuccess ? e : launch_zero_mask(a, out, skip, s);
        } else {
            e = launch_apply5(a, NC, m.WPB, m.grid_cheb, z, out, part, skip, s, &cs);
            if (e == hipSuccess) e = launch_bgen(a, false, z, out, skip, s, &cs);
            return e != hipSuccess ? e : launch_zero_mask(a, out, skip, s);
        }
    }
    if (m.variant == 4) {
        if constexpr (sizeof(T) == 8)
            e = launch_apply4<T, 0>(a, NC, m.WPB, m.grid_cheb, z, out, part, skip, s, &cs);
        else if (m.mfma_mode == 2)
            e = launch_apply4<T, 2>(a, NC, m.WPB, m.grid_cheb, z, out, part, skip, s, &cs);
        else
            e = launch_apply4<T, 1>(a, NC, m.WPB, m.grid_cheb, z, out, part, skip, s, &cs);
        if (e == hipSuccess) e = launch_bgen(a, false, z, out, skip, s, &cs);
        return e != hipSuccess ? e : launch_zero_mask(a, out, skip, s);
    }
    if (part) return hipErrorInvalidValue;
    if (m.variant == 3 && (a.W == 15 || a.W == 16) && a.crec && a.tab) {
        // the buffer-load VALU schedule with the step fused into its write
        dim3 g(apply_blocks(a.nP, m)), b(TR_WAVE * m.WPB);
        const size_t sh = apply3_lds_bytes(m, NC, a.slm, a.arow, a.brow, sizeof(T));
#define TR_CHEB3(JJ, NCC)                                                                                     \
    do {                                                                                                      \
        if (a.W == 15)                                                                                        \
            hipLaunchKernelGGL((k_apply3<T, JJ, NCC, 15, 2>), g, b, sh, s, a, z, out, nullptr, skip, cs);     \
        else                                                                                                  \
            hipLaunchKernelGGL((k_apply3<T, JJ, NCC, 16, 2>), g, b, sh, s, a, z, out, nullptr, skip, cs);     \
    } while (0)
#define TR_CASE(JJ)              \
    case JJ:                     \
        if (NC == 7)             \
            TR_CHEB3(JJ, 7);     \
        else                     \
            TR_CHEB3(JJ, 10);    \
        break;
        switch (m.J) {
            TR_J_CASES(TR_CASE)
            default:
                return hipErrorInvalidValue;
        }
#undef TR_CASE
#undef TR_CHEB3
        e = hipGetLastError();
        if (e == hipSuccess) e = launch_bgen(a, false, z, out, skip, s, &cs);
        return e != hipSuccess ? e : launch_zero_mask(a, out, skip, s);
    }
    if (!tmp) return hipErrorInvalidValue;
    if ((e = launch_apply<T>(a, NC, m, z, tmp, nullptr, skip, s)) != hipSuccess) return e;
    const size_t n = (size_t)a.nP * a.slm;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_cheb_upd<T>, dim3(g), dim3(256), 0, s, z, out, tmp, cs, n, skip);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_apply_diag(const OpArgs<T> &a, int NC, const LaneMap &m, const T *f, T *r,
                             double *part, const double *skip, hipStream_t s) {
    if (a.nP == 0) return hipSuccess;
    if (part && a.ngv) return hipErrorInvalidValue;
    dim3 g(apply_tiles(a.nP, m)), b(TR_WAVE * m.WPB);
#define TR_CASE(JJ)                                                                          \
    case JJ:                                                                                 \
        if (part)                                                                            \
            hipLaunchKernelGGL((k_apply_diag<T, JJ, true>), g, b, 0, s, a, NC, f, r, part, skip);  \
        else                                                                                 \
            hipLaunchKernelGGL((k_apply_diag<T, JJ, false>), g, b, 0, s, a, NC, f, r, part, skip); \
        break;
    switch (m.J) {
        TR_J_CASES(TR_CASE)
        default: return hipErrorInvalidValue;
    }
#undef TR_CASE
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_bgen(a, true, f, r, skip, s);
    return e != hipSuccess ? e : launch_zero_mask(a, r, skip, s);
}

template hipError_t launch_cheb_step<double>(const OpArgs<double> &, int, const LaneMap &, const double *, double *,
                                             const ChebStep<double> &, double *, const double *, hipStream_t,
                                             double *);
template hipError_t launch_cheb_step<float>(const OpArgs<float> &, int, const LaneMap &, const float *, float *,
                                            const ChebStep<float> &, float *, const double *, hipStream_t, double *);
template hipError_t launch_zero_mask<double>(const OpArgs<double> &, double *, const double *, hipStream_t);
template hipError_t launch_zero_mask<float>(const OpArgs<float> &, float *, const double *, hipStream_t);
template hipError_t launch_apply<double>(const OpArgs<double> &, int, const LaneMap &, const double *,
                                         double *, double *, const double *, hipStream_t);
template hipError_t launch_apply<float>(const OpArgs<float> &, int, const LaneMap &, const float *,
                                        float *, double *, const double *, hipStream_t);
template hipError_t launch_apply_diag<double>(const OpArgs<double> &, int, const LaneMap &,
                                              const double *, double *, double *, const double *,
                                              hipStream_t);
template hipError_t launch_apply_diag<float>(const OpArgs<float> &, int, const LaneMap &,
                                             const float *, float *, double *, const double *,
                                             hipStream_t);

// ---------------------------------------------------------------------------
__global__ void k_rhs(const double *__restrict__ rhs, uint32_t nP, uint32_t slm,
                      double *__restrict__ b) {
    size_t n = (size_t)nP * slm;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t v = i / slm, lm = i - v * slm;
        b[i] = lm == 0 ? rhs[v] : 0.0;
    }
}

hipError_t launch_rhs(const double *rhs, uint32_t nP, uint32_t slm, double *b, hipStream_t s) {
    size_t n = (size_t)nP * slm;
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_rhs, dim3(g), dim3(256), 0, s, rhs, nP, slm, b);
    return hipGetLastError();
}

__global__ void k_axpby(double *__restrict__ x, double wx, const double *__restrict__ y,
                        double wy, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        x[i] = wx * x[i] + (y ? wy * y[i] : 0.0);
}

hipError_t launch_axpby(double *x, double wx, const double *y, double wy, size_t n,
                        hipStream_t s) {
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    if (g == 0) return hipSuccess;
    hipLaunchKernelGGL(k_axpby, dim3(g), dim3(256), 0, s, x, wx, y, wy, n);
    return hipGetLastError();
}

// Deterministic FP64 reductions: per-workgroup partials, summed in a fixed
// order by one workgroup.
__global__ __launch_bounds__(256) void k_dot_part(const double *__restrict__ x,
                                                  const double *__restrict__ y, size_t n,
                                                  double *__restrict__ part) {
    double s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        s = fma(x[i], y[i], s);
    s = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_sum_final(const double *__restrict__ part, int n,
                                                   double *__restrict__ out) {
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
    s = block_sum(s);
    if (threadIdx.x == 0) out[0] = s;
}

int red_partials() { return TR_RED_BLOCKS; }

hipError_t launch_dot(const double *x, const double *y, size_t n, double *part, double *out,
                      hipStream_t s) {
    unsigned g = (unsigned)std::min<size_t>(std::max<size_t>((n + 1023) / 1024, 1), TR_RED_BLOCKS);
    hipLaunchKernelGGL(k_dot_part, dim3(g), dim3(256), 0, s, x, y, n, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, part, (int)g, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ void k_to_internal(const double *__restrict__ src, double *__restrict__ dst,
                              const uint32_t *__restrict__ perm, uint32_t nP, uint32_t slm,
                              uint32_t aslm) {
    size_t n = (size_t)nP * slm;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t v = i / slm, lm = i - v * slm;
        dst[i] = src[(size_t)perm[v] * aslm + lm];
    }
}

__global__ void k_to_external(const double *__restrict__ src, double *__restrict__ dst,
                              const uint32_t *__restrict__ iperm, uint32_t nP, uint32_t slm,
                              uint32_t aslm) {
    size_t n = (size_t)nP * aslm;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        size_t v = i / aslm, lm = i - v * aslm;
        dst[i] = lm < slm ? src[(size_t)iperm[v] * slm + lm] : 0.0;
    }
}

hipError_t launch_to_internal(const double *src, double *dst, const uint32_t *perm, uint32_t nP,
                              uint32_t slm, uint32_t aslm, hipStream_t s) {
    size_t n = (size_t)nP * slm;
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_internal, dim3(g), dim3(256), 0, s, src, dst, perm, nP, slm, aslm);
    return hipGetLastError();
}

hipError_t launch_to_external(const double *src, double *dst, const uint32_t *iperm, uint32_t nP,
                              uint32_t slm, uint32_t aslm, hipStream_t s) {
    size_t n = (size_t)nP * aslm;
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_external, dim3(g), dim3(256), 0, s, src, dst, iperm, nP, slm, aslm);
    return hipGetLastError();
}

}  // namespace tr
