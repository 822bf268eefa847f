// mfma_common.h -- device helpers shared by the matrix-core operator
// kernels (kernels_mfma.hip, kernels_mfma5s.hip): buffer descriptors and
// loads, the pipelined ELL gathers, DPP quad broadcasts and the move from
// the gather layout to the MFMA operand layout (see kernels_mfma.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <utility>

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

// AUX: the cache policy bits; kNonTemporal for streams that are read once
// per operator application (the coefficient records), so they do not
// displace the vectors that the next kernel reads again
constexpr int kNonTemporal = 2;
template <typename T, int AUX = 0>
__device__ inline T bload1(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, AUX));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, AUX));
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

typedef __bf16 bf4v __attribute__((ext_vector_type(4)));

// One channel group of the angular contraction: D[t] += A_c B_c over the
// NCH channels of acc (lane layout as above), B slices at B (hi) and B + lo
// (lo part of the BF16 split).  BF16 nsplit: 1 = A_hi B_hi, 2 = + A_lo B_hi
// (the A operand split only), 3 = + A_hi B_lo.
template <typename T, int MODE, int J, int NCH, typename ACC4, int NT>
__device__ __forceinline__ void contract(const T (&acc)[NCH][J], const unsigned char *B, size_t lo, int lane,
                                         int nsplit, ACC4 (&D)[NT]) {
    constexpr int S = mfma_steps(MODE, J);
    constexpr int BW = (MODE == 2) ? 8 : (int)sizeof(T);
#pragma unroll
    for (int c = 0; c < NCH; c++) {
#pragma unroll
        for (int st = 0; st < S; st++) {
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const size_t bo = ((((size_t)c * S + st) * NT + t) * TR_WAVE + lane) * BW;
                if constexpr (MODE == 0) {
                    const double bv = *reinterpret_cast<const double *>(B + bo);
                    D[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[c][st], bv, D[t], 0, 0, 0);
                } else if constexpr (MODE == 1) {
                    const float bv = *reinterpret_cast<const float *>(B + bo);
                    D[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(acc[c][st], bv, D[t], 0, 0, 0);
                } else {
                    // A operand: registers 4st .. 4st+3, rounded to BF16 (the
                    // hardware conversion rounds to nearest even) and, for the
                    // split, the BF16 residual x - hi
                    f4v x;
#pragma unroll
                    for (int u = 0; u < 4; u++) x[u] = (4 * st + u < J) ? (float)acc[c][4 * st + u] : 0.0f;
                    const bf4v hb = __builtin_convertvector(x, bf4v);
                    const s4v ah = __builtin_bit_cast(s4v, hb), bh = *reinterpret_cast<const s4v *>(B + bo);
                    D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, D[t], 0, 0, 0);
                    if (nsplit > 1) {
                        const s4v al = __builtin_bit_cast(s4v, __builtin_convertvector(x - __builtin_convertvector(hb, f4v), bf4v));
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, D[t], 0, 0, 0);
                        if (nsplit > 2) {
                            const s4v bl = *reinterpret_cast<const s4v *>(B + lo + bo);
                            D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, D[t], 0, 0, 0);
                        }
                    }
                }
            }
        }
        // one channel's B slices at a time: hoisting all of them above the
        // MFMAs would hold every slice in registers at once
        __builtin_amdgcn_sched_barrier(0);
    }
}

// A wavefront-uniform value the optimiser cannot see through (so it is not
// hoisted out of loops and kept live): laundered through a VGPR, back in an
// SGPR with readfirstlane.
__device__ inline uint32_t opaque_uniform(uint32_t x) {
    asm volatile("" : "+v"(x));
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// Neighbour indices of `nslot` slots (nslot <= NS, wavefront-uniform): slot
// s at nrs (voff, s*nstride).
template <int NS>
__device__ inline void load_idx(uint32_t (&nb)[NS], __amdgpu_buffer_rsrc_t nrs, uint32_t voff, uint32_t nstride,
                                int nslot) {
    nstride = opaque_uniform(nstride);
#pragma unroll
    for (int s = 0; s < NS; s++)
        nb[s] = s < nslot ? (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(nrs, voff, (uint32_t)s * nstride, 0) : 0u;
}

// Broadcast of lane `src` of each quad (4 adjacent lanes) to the quad.
__device__ inline int quad_bcast32(int x, int src) {
    switch (src) {
        case 0: return __builtin_amdgcn_mov_dpp(x, 0x00, 0xf, 0xf, false);
        case 1: return __builtin_amdgcn_mov_dpp(x, 0x55, 0xf, 0xf, false);
        case 2: return __builtin_amdgcn_mov_dpp(x, 0xaa, 0xf, 0xf, false);
        default: return __builtin_amdgcn_mov_dpp(x, 0xff, 0xf, 0xf, false);
    }
}
__device__ inline float quad_bcast(float x, int src) {
    return __builtin_bit_cast(float, quad_bcast32(__builtin_bit_cast(int, x), src));
}
__device__ inline double quad_bcast(double x, int src) {
    i32x2m v = __builtin_bit_cast(i32x2m, x);
    v.x = quad_bcast32(v.x, src);
    v.y = quad_bcast32(v.y, src);
    return __builtin_bit_cast(double, v);
}

// Pipelined ELL gather of one lane (chunk q of its vertex, the vertex's four
// lanes adjacent): acc[c][j] += sum_s cf_s[c] f[nb_s][mfma_harm(E, q, j)]
// over `nslot` slots (nslot <= NS, a wavefront-uniform count).  The
// coefficient record (RS elements) of slot s sits at byte coff + s*cstride
// of the record array (one buffer descriptor: per-slot descriptors, 4 SGPRs
// each, get hoisted out of the tile loop and spilled); lane q loads elements
// q*PE .. q*PE+PE-1 of it and channel c comes from lane c/PE of the quad.
// Reads past the last record (PE*4 > RS) fall outside the descriptor's
// range, which returns zeros.  DEPTH slots are in flight: slot s+DEPTH's loads are issued
// right after slot s's FMAs, so a wavefront waits for one memory round trip
// per DEPTH slots.  Left alone, the optimiser issues every slot's loads
// first and sinks all FMAs below them, which holds NS slots of data live and
// spills; empty volatile asm statements (which keep their relative order)
// pin each slot's FMAs before the next load offsets are formed.
// hoff / h0: byte and harmonic offset of the harmonic block gathered (the
// K-blocked schedule of kernels with more than 32 harmonics).
// SLOTRS: the records of slot s are read through a descriptor of their own
// (base cbase + s * cstride bytes, cstride bytes long) instead of crs at
// offset s * cstride, so that one slot's records, not all of them, must fit
// the 32-bit buffer offsets (unions of many problems).  The descriptors are
// formed per call from the opaque stride, one slot at a time.
template <typename T, int J, int NCH, int RS, int NS, int DEPTH, bool SLOTRS = false>
__device__ __forceinline__ void gather(T (&acc)[NCH][J], const uint32_t (&nb)[NS], __amdgpu_buffer_rsrc_t frs,
                                       __amdgpu_buffer_rsrc_t crs, uint32_t cstride, uint32_t coff, int nslot,
                                       uint32_t slm, int q, uint32_t hoff = 0, uint32_t h0 = 0,
                                       uint32_t hw = 0xffffffffu, const T *cbase = nullptr) {
    constexpr int E = 16 / (int)sizeof(T), PE = (RS + 3) / 4;
    // opaque here, so that the per-slot products and tests are formed where
    // used instead of hoisted out of the tile loop (kept live, they spill)
    cstride = opaque_uniform(cstride);
    slm = opaque_uniform(slm);
    T xs[DEPTH][J], cs[DEPTH][PE];
    auto load = [&](int s, T *x, T *cf) __attribute__((always_inline)) {
        uint32_t base = nb[s] * slm * (uint32_t)sizeof(T) + 16u * (uint32_t)q + hoff;
        uint32_t cv = coff + (uint32_t)(q * PE * (int)sizeof(T)) + (SLOTRS ? 0u : (uint32_t)s * cstride);
        asm volatile("" : "+v"(base), "+v"(cv));
#pragma unroll
        for (int j = 0; j < J; j++)
            x[j] = bload1<T>(frs, base + (uint32_t)((64 * (j / E) + (j % E) * (int)sizeof(T))), 0u);
        if constexpr (SLOTRS) {
            const auto cs1 = rsrc4(reinterpret_cast<const unsigned char *>(cbase) + (size_t)s * cstride, cstride);
#pragma unroll
            for (int e = 0; e < PE; e++) cf[e] = bload1<T>(cs1, cv + (uint32_t)(e * (int)sizeof(T)), 0u);
        } else {
#pragma unroll
            for (int e = 0; e < PE; e++) cf[e] = bload1<T>(crs, cv + (uint32_t)(e * (int)sizeof(T)), 0u);
        }
    };
    auto accum = [&](const T *x, const T *cf) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const T w = quad_bcast(cf[c % PE], c / PE);
#pragma unroll
            for (int j = 0; j < J; j++) acc[c][j] = fma(w, x[j], acc[c][j]);
        }
#pragma unroll
        for (int c = 0; c < NCH; c++)
#pragma unroll
            for (int j = 0; j < J; j++) asm volatile("" : "+v"(acc[c][j]));
    };
#pragma unroll
    for (int d = 0; d < DEPTH; d++)
        if (d < NS && d < nslot) load(d, xs[d], cs[d]);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        if (s < nslot) {
            accum(xs[s % DEPTH], cs[s % DEPTH]);
            if (s + DEPTH < NS && s + DEPTH < nslot) load(s + DEPTH, xs[s % DEPTH], cs[s % DEPTH]);
        }
    }
    // harmonics past slm (padding of the last pieces) or past the block's
    // hw read neighbouring data: zero them
#pragma unroll
    for (int j = 0; j < J; j++)
        if (mfma_harm(E, q, j) + h0 >= slm || (uint32_t)mfma_harm(E, q, j) >= hw)
#pragma unroll
            for (int c = 0; c < NCH; c++) acc[c][j] = 0;
}

// acc[c][j] += (w of lane S of the quad) * x[j], j = 0..3, for two channels:
// the quad broadcast rides on the FMAs as their DPP source operand
// (v_fmac_f32_dpp) instead of a separate v_mov_b32_dpp per channel.  The
// leading s_nop 1 gives the two wait states a DPP read needs after a VALU
// write of its source register, whatever the compiler placed before.
#define TR_DPP_FMAC2(S0, S1)                                                                            \
    asm volatile("s_nop 1\n"                                                                          \
                 "v_fmac_f32_dpp %0, %8, %10 quad_perm:[" #S0 "," #S0 "," #S0 "," #S0 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %1, %8, %11 quad_perm:[" #S0 "," #S0 "," #S0 "," #S0 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %2, %8, %12 quad_perm:[" #S0 "," #S0 "," #S0 "," #S0 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %3, %8, %13 quad_perm:[" #S0 "," #S0 "," #S0 "," #S0 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %4, %9, %10 quad_perm:[" #S1 "," #S1 "," #S1 "," #S1 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %5, %9, %11 quad_perm:[" #S1 "," #S1 "," #S1 "," #S1 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %6, %9, %12 quad_perm:[" #S1 "," #S1 "," #S1 "," #S1 "] row_mask:0xf bank_mask:0xf\n" \
                 "v_fmac_f32_dpp %7, %9, %13 quad_perm:[" #S1 "," #S1 "," #S1 "," #S1 "] row_mask:0xf bank_mask:0xf" \
                 : "+v"(a0[0]), "+v"(a0[1]), "+v"(a0[2]), "+v"(a0[3]), "+v"(a1[0]), "+v"(a1[1]), "+v"(a1[2]), \
                   "+v"(a1[3])                                                                          \
                 : "v"(w0), "v"(w1), "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]))

template <int S0, int S1>
__device__ __forceinline__ void dpp_fmac2(float (&a0)[4], float (&a1)[4], float w0, float w1, const float *x) {
    static_assert(S0 >= 0 && S0 < 4 && S1 >= 0 && S1 < 4, "quad lane");
    if constexpr (S0 == 0 && S1 == 0) TR_DPP_FMAC2(0, 0);
    else if constexpr (S0 == 0 && S1 == 1) TR_DPP_FMAC2(0, 1);
    else if constexpr (S0 == 1 && S1 == 1) TR_DPP_FMAC2(1, 1);
    else if constexpr (S0 == 1 && S1 == 2) TR_DPP_FMAC2(1, 2);
    else if constexpr (S0 == 2 && S1 == 2) TR_DPP_FMAC2(2, 2);
    else if constexpr (S0 == 2 && S1 == 3) TR_DPP_FMAC2(2, 3);
    else if constexpr (S0 == 3 && S1 == 3) TR_DPP_FMAC2(3, 3);
    else if constexpr (S0 == 0 && S1 == 2) TR_DPP_FMAC2(0, 2);
    else if constexpr (S0 == 1 && S1 == 3) TR_DPP_FMAC2(1, 3);
    else if constexpr (S0 == 0 && S1 == 3) TR_DPP_FMAC2(0, 3);
    else static_assert(S0 <= S1, "channel pairs in order");
}
#undef TR_DPP_FMAC2

// all NCH channels of one slot, two at a time (channel c's weight is
// element c % PE of quad lane c / PE)
template <int PE, int NCH, int... P>
__device__ __forceinline__ void accum_dpp(float (&acc)[NCH][4], const float *c0, const float *x,
                                          std::integer_sequence<int, P...>) {
    (dpp_fmac2<(2 * P) / PE, (2 * P + 1) / PE>(acc[2 * P], acc[2 * P + 1], c0[(2 * P) % PE], c0[(2 * P + 1) % PE], x),
     ...);
}

// The gather with the slots' coefficient records already in registers
// (cf[s][e]: element q*PE + e of slot s's record, loaded once per tile by
// the K-blocked schedule); only the neighbour rows stream, DEPTH slots in
// flight.
template <typename T, int J, int NCH, int PE, int NS, int DEPTH>
__device__ __forceinline__ void gather_pre(T (&acc)[NCH][J], const uint32_t (&nb)[NS], __amdgpu_buffer_rsrc_t frs,
                                           const T (&cf)[NS][PE], uint32_t slm, int q, uint32_t hoff, uint32_t hw) {
    constexpr int E = 16 / (int)sizeof(T);
    slm = opaque_uniform(slm);
    T xs[DEPTH][J];
    auto load = [&](int s, T *x) __attribute__((always_inline)) {
        uint32_t base = nb[s] * slm * (uint32_t)sizeof(T) + 16u * (uint32_t)q + hoff;
        asm volatile("" : "+v"(base));
#pragma unroll
        for (int j = 0; j < J; j++)
            x[j] = bload1<T>(frs, base + (uint32_t)((64 * (j / E) + (j % E) * (int)sizeof(T))), 0u);
    };
    auto accum = [&](const T *x, const T *c0) __attribute__((always_inline)) {
        if constexpr (std::is_same<T, float>::value && J == 4 && NCH % 2 == 0) {
            accum_dpp<PE, NCH>(acc, c0, x, std::make_integer_sequence<int, NCH / 2>{});
        } else {
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const T w = quad_bcast(c0[c % PE], c / PE);
#pragma unroll
                for (int j = 0; j < J; j++) acc[c][j] = fma(w, x[j], acc[c][j]);
            }
#pragma unroll
            for (int c = 0; c < NCH; c++)
#pragma unroll
                for (int j = 0; j < J; j++) asm volatile("" : "+v"(acc[c][j]));
        }
    };
#pragma unroll
    for (int d = 0; d < DEPTH && d < NS; d++) load(d, xs[d]);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        accum(xs[s % DEPTH], cf[s]);
        if (s + DEPTH < NS) load(s + DEPTH, xs[s % DEPTH]);
    }
    // harmonics past the block's hw read the next block's (or the next
    // row's) values: zero them
#pragma unroll
    for (int j = 0; j < J; j++)
        if ((uint32_t)mfma_harm(E, q, j) >= hw)
#pragma unroll
            for (int c = 0; c < NCH; c++) acc[c][j] = 0;
}

__device__ inline double bperm(int addr, double x) {
    i32x2m v = __builtin_bit_cast(i32x2m, x);
    v.x = __builtin_amdgcn_ds_bpermute(addr, v.x);
    v.y = __builtin_amdgcn_ds_bpermute(addr, v.y);
    return __builtin_bit_cast(double, v);
}
__device__ inline float bperm(int addr, float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, x)));
}

// Moves the accumulators from the gather layout (lane 4v + q) to the MFMA
// operand layout (lane v + 16q).  MODE 0 at J = 8: register 6 of chunk q
// becomes harmonic 24 + q (register 6 + q%2 of chunk q/2), see mfma_steps.
template <typename T, int MODE, int J, int NCH>
__device__ __forceinline__ void to_mfma_layout(T (&acc)[NCH][J], int lane) {
    constexpr bool MERGE = MODE == 0 && J == 8;
    const int src = 4 * (4 * (lane & 15) + (lane >> 4));
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int j = 0; j < (MERGE ? 6 : J); j++) acc[c][j] = bperm(src, acc[c][j]);
    if constexpr (MERGE) {
        const int src2 = 4 * (4 * (lane & 15) + (lane >> 5));
        const bool odd = (lane >> 4) & 1;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const T a6 = bperm(src2, acc[c][6]), a7 = bperm(src2, acc[c][7]);
            acc[c][6] = odd ? a7 : a6;
        }
    }
}

}  // namespace tr
