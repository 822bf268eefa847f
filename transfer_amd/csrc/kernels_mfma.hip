// kernels_mfma.hip -- the operator with its angular contraction on the matrix
// cores (schedule variant 4).
//
// Lane layout: a wavefront holds 16 vertices; lane l owns vertex l % 16 and
// the harmonic chunk q = l / 16, i.e. harmonics q*J .. q*J+J-1 with
// J = ceil(slm / 4).  Phase 1 (as in the other schedules) gathers the ELL
// neighbours and accumulates the spatial channels T_c[v][lm] in registers.
// Phase 2 is
//   R[v][n] = sum_c sum_m T_c[v][m] Mat_c[n][m]
// (c = mass (Mat = diag(massfac)), the NC-1 stiffness channels and, for
// wavefronts holding boundary vertices, the 3 boundary channels), a
// [16 vertices x K] x [K x 16 outputs] product per output tile.  With the
// K index ordered as (channel, j, chunk q), the accumulator acc[c][j] of
// lane l is exactly the A operand of one MFMA step (A[v][q]), so the
// channel sums never leave the registers.  B slices (4 x 16 couplings per
// step and tile) are staged in LDS once per workgroup.
//   FP64 operator: v_mfma_f64_16x16x4f64, K = 4 per step (one j).
//   FP32 operator: v_mfma_f32_16x16x4f32 (exact FP32), or BF16 MFMA
//   v_mfma_f32_16x16x16bf16_1k with FP32 accumulation, K = 16 per step
//   (4 consecutive j of every chunk), optionally with the hi/lo split
//   A_hi B_hi + A_lo B_hi + A_hi B_lo for ~16 significant bits.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"

namespace tr {

#define TR_WAVE 64

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));
typedef int i32x2m __attribute__((ext_vector_type(2)));
typedef int i32x4m __attribute__((ext_vector_type(4)));
typedef double f64x2m __attribute__((ext_vector_type(2)));
typedef float f32x4m __attribute__((ext_vector_type(4)));

__device__ inline uint32_t xcd_block4(uint32_t bid, uint32_t nblk) {
    const uint32_t X = 8, per = nblk / X, rem = nblk % X;
    const uint32_t xcd = bid % X, k = bid / X;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t rsrc4(const T *p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(p), (short)0,
                                             (int)(bytes > 0xffffffffull ? 0xffffffffull : bytes),
                                             0x00020000);
}

template <typename T>
__device__ inline T bload1(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}

// coefficient record of one slot (RS elements, a 16-byte multiple)
template <typename T, int RS>
__device__ inline void brec(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff, T *out) {
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int q = 0; q < RS; q += 2) {
            f64x2m v = __builtin_bit_cast(f64x2m, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + q * 8, soff, 0));
            out[q] = v.x;
            out[q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < RS; q += 4) {
            f32x4m v = __builtin_bit_cast(f32x4m, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + q * 4, soff, 0));
            out[q] = v.x;
            out[q + 1] = v.y;
            out[q + 2] = v.z;
            out[q + 3] = v.w;
        }
    }
}

__device__ inline short bf16_rne(float x) {
    unsigned u = __builtin_bit_cast(unsigned, x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline float bf16_to_f(short h) {
    return __builtin_bit_cast(float, ((unsigned)(unsigned short)h) << 16);
}

// MODE: 0 = FP64 (T = double), 1 = FP32 MFMA, 2 = BF16 MFMA (T = float)
template <typename T, int MODE, int J, int NC, int WT, bool DOT>
__global__ __launch_bounds__(512) void k_apply4(OpArgs<T> a, const T *__restrict__ f,
                                                T *__restrict__ r, double *__restrict__ part,
                                                const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RS = rec_stride<T>(NC);
    constexpr int S = (MODE == 2) ? (J + 3) / 4 : J;         // K-steps per channel
    constexpr int BW = (MODE == 2) ? 8 : (int)sizeof(T);      // B bytes per lane and step
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t slm = a.slm, nP = a.nP;
    const int NT = (int)((slm + 15) / 16);
    const int nsplit = a.nsplit;
    // stage the volume B slices: NC channels x S steps x NT tiles x 64 lanes
    const uint32_t nvol = (uint32_t)NC * S * NT * TR_WAVE * BW * (MODE == 2 && nsplit > 1 ? 2 : 1);
    {
        const i32x4m *src = reinterpret_cast<const i32x4m *>(a.tab);
        i32x4m *dst = reinterpret_cast<i32x4m *>(smem_raw);
        for (uint32_t i = threadIdx.x; i < nvol / 16; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const uint32_t blk = xcd_block4(blockIdx.x, gridDim.x);
    const int vloc = lane & 15, q = lane >> 4;
    const uint32_t v0 = (blk * (uint32_t)wpb + wave) * 16u;
    const uint32_t v = v0 + vloc;
    const bool valid = v < nP;
    const int lm0 = q * J;
    T acc[NC][J];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int j = 0; j < J; j++) acc[c][j] = 0;
    if (valid) {
        const auto frs = rsrc4(f, (size_t)nP * slm * sizeof(T));
        const auto nrs = rsrc4(a.nbr, (size_t)a.W * nP * 4);
        const auto crs = rsrc4(a.crec, (size_t)a.W * nP * RS * sizeof(T));
        uint32_t nb[WT];
#pragma unroll
        for (int s = 0; s < WT; s++)
            nb[s] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(nrs, v * 4u, (uint32_t)s * nP * 4u, 0);
        const uint32_t cvoff = v * (uint32_t)(RS * sizeof(T));
        const uint32_t cstride = nP * (uint32_t)(RS * sizeof(T));
        // two slots in flight: slot s+1's loads are issued before slot s's
        // FMAs.  Left alone, the optimiser issues every slot's loads first and
        // sinks all FMAs below them, which holds W slots of data live and
        // spills; empty volatile asm statements (which keep their relative
        // order) pin each slot's FMAs before the next slot's load offsets.
        T xa[J], ca[RS], xb[J], cb[RS];
        auto load = [&](int s, T *x, T *cf) {
            uint32_t base = (nb[s] * slm + lm0) * (uint32_t)sizeof(T);
            uint32_t cv = cvoff;
            asm volatile("" : "+v"(base), "+v"(cv));
#pragma unroll
            for (int j = 0; j < J; j++) x[j] = bload1<T>(frs, base + j * (uint32_t)sizeof(T), 0u);
            brec<T, RS>(crs, cv, (uint32_t)s * cstride, cf);
        };
        auto accum = [&](const T *x, const T *cf) {
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < J; j++) acc[c][j] = fma(cf[c], x[j], acc[c][j]);
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < J; j++) asm volatile("" : "+v"(acc[c][j]));
        };
        load(0, xa, ca);
#pragma unroll
        for (int s = 0; s < WT; s += 2) {
            if (s + 1 < WT) load(s + 1, xb, cb);
            __builtin_amdgcn_sched_barrier(0);
            accum(xa, ca);
            __builtin_amdgcn_sched_barrier(0);
            if (s + 2 < WT) load(s + 2, xa, ca);
            __builtin_amdgcn_sched_barrier(0);
            if (s + 1 < WT) accum(xb, cb);
            __builtin_amdgcn_sched_barrier(0);
        }
        // harmonics past slm (the last chunk) carry neighbouring data: zero them
#pragma unroll
        for (int j = 0; j < J; j++)
            if (lm0 + j >= (int)slm)
#pragma unroll
                for (int c = 0; c < NC; c++) acc[c][j] = 0;
    }
    // ---- angular contraction on the matrix cores
    constexpr int NTMAX = (J * 4 + 15) / 16;
    typedef typename std::conditional<MODE == 0, d4v, f4v>::type acc4;
    acc4 D[NTMAX];
#pragma unroll
    for (int t = 0; t < NTMAX; t++) D[t] = (acc4){0, 0, 0, 0};
    const unsigned char *Bv = smem_raw;
#pragma unroll
    for (int c = 0; c < NC; c++) {
#pragma unroll
        for (int st = 0; st < S; st++) {
#pragma unroll
            for (int t = 0; t < NTMAX; t++) {
                if (t >= NT) break;
                const size_t bo = ((((size_t)c * S + st) * NT + t) * TR_WAVE + lane) * BW;
                if constexpr (MODE == 0) {
                    const double bv = *reinterpret_cast<const double *>(Bv + bo);
                    D[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[c][st], bv, D[t], 0, 0, 0);
                } else if constexpr (MODE == 1) {
                    const float bv = *reinterpret_cast<const float *>(Bv + bo);
                    D[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(acc[c][st], bv, D[t], 0, 0, 0);
                } else {
                    s4v ah, al, bh = *reinterpret_cast<const s4v *>(Bv + bo);
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const float x = (4 * st + u < J) ? (float)acc[c][4 * st + u] : 0.0f;
                        ah[u] = bf16_rne(x);
                        al[u] = bf16_rne(x - bf16_to_f(ah[u]));
                    }
                    D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, D[t], 0, 0, 0);
                    if (nsplit > 1) {
                        const s4v bl = *reinterpret_cast<const s4v *>(Bv + (size_t)NC * S * NT * TR_WAVE * BW + bo);
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, D[t], 0, 0, 0);
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, D[t], 0, 0, 0);
                    }
                }
            }
        }
        // one channel's B slices at a time: hoisting all of them above the
        // MFMAs would hold every slice in registers at once
        __builtin_amdgcn_sched_barrier(0);
    }
    // ---- vacuum boundary: only wavefronts that hold boundary vertices
    const int b = valid ? a.bof[v] : -1;
    if (__any(b >= 0)) {
        T bacc[3][J];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < J; j++) bacc[c][j] = 0;
        if (b >= 0) {
            const size_t nB = a.nB;
            for (uint32_t s = 0; s < a.WB; s++) {
                const uint32_t k = a.bnbr[s * nB + b];
                const T *fk = f + (size_t)k * slm + lm0;
                T x[J];
#pragma unroll
                for (int j = 0; j < J; j++) x[j] = (lm0 + j < (int)slm) ? fk[j] : (T)0;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const T cf = a.bcoef[((size_t)s * 3 + c) * nB + b];
#pragma unroll
                    for (int j = 0; j < J; j++) bacc[c][j] = fma(cf, x[j], bacc[c][j]);
                }
            }
        }
        // boundary B slices come from global memory (few wavefronts read them)
        const unsigned char *Bb = reinterpret_cast<const unsigned char *>(a.tab) + nvol;
#pragma unroll
        for (int c = 0; c < 3; c++) {
#pragma unroll
            for (int st = 0; st < S; st++) {
#pragma unroll
                for (int t = 0; t < NTMAX; t++) {
                    if (t >= NT) break;
                    const size_t bo = ((((size_t)c * S + st) * NT + t) * TR_WAVE + lane) * BW;
                    if constexpr (MODE == 0) {
                        D[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(bacc[c][st], *reinterpret_cast<const double *>(Bb + bo), D[t], 0, 0, 0);
                    } else if constexpr (MODE == 1) {
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bacc[c][st], *reinterpret_cast<const float *>(Bb + bo), D[t], 0, 0, 0);
                    } else {
                        s4v ah, al, bh = *reinterpret_cast<const s4v *>(Bb + bo);
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const float x = (4 * st + u < J) ? (float)bacc[c][4 * st + u] : 0.0f;
                            ah[u] = bf16_rne(x);
                            al[u] = bf16_rne(x - bf16_to_f(ah[u]));
                        }
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, D[t], 0, 0, 0);
                        if (nsplit > 1) {
                            const s4v bl = *reinterpret_cast<const s4v *>(Bb + (size_t)3 * S * NT * TR_WAVE * BW + bo);
                            D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, D[t], 0, 0, 0);
                            D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, D[t], 0, 0, 0);
                        }
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // ---- write: FP64 D holds rows (l/16) + 4i, FP32 rows 4(l/16) + i; column l % 16
    double d = 0;
#pragma unroll
    for (int t = 0; t < NTMAX; t++) {
        if (t >= NT) break;
        const uint32_t n = 16u * t + (uint32_t)vloc;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t vr = v0 + (MODE == 0 ? (uint32_t)(q + 4 * i) : (uint32_t)(4 * q + i));
            if (n < slm && vr < nP) {
                const T val = (T)D[t][i];
                r[(size_t)vr * slm + n] = val;
                if (DOT) d = fma((double)f[(size_t)vr * slm + n], (double)val, d);
            }
        }
    }
    if (DOT) {
        __shared__ double w[8];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, TR_WAVE);
        if (lane == 0) w[wave] = d;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int i = 0; i < wpb; i++) t += w[i];
            part[blockIdx.x] = t;
        }
    }
}

int mfma_J(uint32_t slm) { return (int)((slm + 3) / 4); }

bool mfma_supported(uint32_t slm, int precision_mode) {
    const int J = mfma_J(slm);
    (void)precision_mode;
    return J == 2 || J == 4 || J == 7;
}

// bytes of the volume B table staged in LDS
size_t apply4_lds_bytes(int mode, int NC, uint32_t slm, int nsplit) {
    const int J = mfma_J(slm), S = mode == 2 ? (J + 3) / 4 : J, BW = mode == 2 ? 8 : (mode == 0 ? 8 : 4);
    const int NT = (int)((slm + 15) / 16);
    return (size_t)NC * S * NT * 64 * BW * (mode == 2 && nsplit > 1 ? 2 : 1);
}

template <typename T, int MODE>
hipError_t launch_apply4(const OpArgs<T> &a, int NC, int wpb, const T *f, T *r, double *part,
                         const double *skip, hipStream_t s) {
    if (a.nP == 0) return hipSuccess;
    const uint32_t per = 16u * (uint32_t)wpb;
    dim3 g((a.nP + per - 1) / per), b(TR_WAVE * wpb);
    const size_t sh = apply4_lds_bytes(MODE, NC, a.slm, a.nsplit);
    const int J = mfma_J(a.slm);
    // the unrolled slot loop reads exactly W = 15 or 16 slots; coefficient
    // records and B slices must be present
    if ((a.W != 15 && a.W != 16) || !a.crec || !a.tab || wpb < 1 || wpb > 8) return hipErrorInvalidValue;
#define TR_L4(JJ, NCC, WW, D) \
    hipLaunchKernelGGL((k_apply4<T, MODE, JJ, NCC, WW, D>), g, b, sh, s, a, f, r, part, skip)
#define TR_L4W(JJ, NCC, D)                        \
    do {                                          \
        if (a.W == 15) TR_L4(JJ, NCC, 15, D);     \
        else TR_L4(JJ, NCC, 16, D);               \
    } while (0)
#define TR_L4N(JJ, D)                             \
    do {                                          \
        if (NC == 7) TR_L4W(JJ, 7, D);            \
        else TR_L4W(JJ, 10, D);                   \
    } while (0)
#define TR_L4J(D)                                 \
    do {                                          \
        if (J == 2) TR_L4N(2, D);                 \
        else if (J == 4) TR_L4N(4, D);            \
        else if (J == 7) TR_L4N(7, D);            \
        else return hipErrorInvalidValue;         \
    } while (0)
    if (part)
        TR_L4J(true);
    else
        TR_L4J(false);
#undef TR_L4J
#undef TR_L4N
#undef TR_L4W
#undef TR_L4
    return hipGetLastError();
}

template hipError_t launch_apply4<double, 0>(const OpArgs<double> &, int, int, const double *, double *,
                                             double *, const double *, hipStream_t);
template hipError_t launch_apply4<float, 1>(const OpArgs<float> &, int, int, const float *, float *,
                                            double *, const double *, hipStream_t);
template hipError_t launch_apply4<float, 2>(const OpArgs<float> &, int, int, const float *, float *,
                                            double *, const double *, hipStream_t);

}  // namespace tr
