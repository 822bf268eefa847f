// kernels_mfma5s.hip -- variant 5 of the FP32 operator (33..80 harmonics,
// BF16 matrix-core projection) under a static parity-class layout.
#include "mfma_common.h"

namespace tr {

// ---- Variant 5 with a static parity-class schedule (k_apply5s) ----
//
// For the two class layouts of kernels.h (K = 5: blocks 16, 5 | 15 | 15 |
// 15; K = 4: 15 | 10 | 10 | 10) every (input block, channel, output block)
// triple that can couple is known at compile time from the classes'
// mirror parities, so the contraction is straight-line code: no mask
// tests, the B reads at constant LDS offsets (issued ahead of their
// MFMAs), and the harmonic blocks unrolled, so one pipelined stream of
// neighbour-row loads runs through all of them (the next block's first
// rows load while the current block contracts).  The first slot of each
// block writes the accumulators with a multiply instead of zeroing them.

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// One slot of the static schedule's gather: acc[c][j] (+)= w_c x[j] for all
// NC channels, w_c (element c % PE of quad lane c / PE) broadcast to the
// quad by a DPP move; FIRST writes instead of accumulating (no zeroing).
// A VALU instruction with a DPP operand issues at half the rate of a plain
// one on gfx950 (4 against 2 cycles per wave64 FP32 FMA, measured), so the
// moves stay separate from the four FMAs each feeds: the empty asm keeps
// the compiler from folding them back.
template <int PE, int NC, bool FIRST>
__device__ __forceinline__ void slot_accum(float (&acc)[NC][4], const float *cw, const float *x) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
        float w = quad_bcast(cw[c % PE], c / PE);
        asm volatile("" : "+v"(w));
#pragma unroll
        for (int j = 0; j < 4; j++) acc[c][j] = FIRST ? w * x[j] : fmaf(w, x[j], acc[c][j]);
    }
    // the slot's products stay ahead of the next slot's (the optimiser
    // would otherwise batch every slot's moves and loads and spill)
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int j = 0; j < 4; j++) asm volatile("" : "+v"(acc[c][j]));
}

// The gather of a block five harmonics wide (harmonics h0 .. h0 + 4, the
// second block of class 0 at K = 5): lane gq of a vertex's quad takes
// harmonic h0 + gq for every channel (acc1, the coefficients broadcast as
// in slot_accum) and harmonic h0 + 4 for its own record's channels
// gq * PE + e (acc4, no broadcast); x[0], x[1] hold those two values of the
// neighbour row.  NC + PE FMAs and NC moves per slot instead of 4 NC FMAs
// and NC moves.
template <int PE, int NC, bool FIRST>
__device__ __forceinline__ void slot_accum5(float (&acc1)[NC], float (&acc4)[PE], const float *cw, const float *x) {
#pragma unroll
    for (int c = 0; c < NC; c++) {
        float w = quad_bcast(cw[c % PE], c / PE);
        asm volatile("" : "+v"(w));
        acc1[c] = FIRST ? w * x[0] : fmaf(w, x[0], acc1[c]);
    }
#pragma unroll
    for (int e = 0; e < PE; e++) acc4[e] = FIRST ? cw[e] * x[1] : fmaf(cw[e], x[1], acc4[e]);
#pragma unroll
    for (int c = 0; c < NC; c++) asm volatile("" : "+v"(acc1[c]));
#pragma unroll
    for (int e = 0; e < PE; e++) asm volatile("" : "+v"(acc4[e]));
}

// The five-wide block's sums in the MFMA operand layout (lane v + 16 q,
// element u: K index 4 q + u of vertex v): K 0..3 from acc1 of the vertex's
// lanes 0..3, K 4 (q = 1, u = 0) from acc4 of lane c / PE, the rest zero.
template <int PE, int NC>
__device__ __forceinline__ void narrow_to_mfma(const float (&acc1)[NC], const float (&acc4)[PE], float (&acc)[NC][4],
                                               int lane) {
    const int v = lane & 15, q = lane >> 4;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        float t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) t[u] = bperm(4 * (4 * v + u), acc1[c]);
        const float t4 = bperm(4 * (4 * v + c / PE), acc4[c % PE]);
        acc[c][0] = q == 0 ? t[0] : (q == 1 ? t4 : 0.0f);
#pragma unroll
        for (int u = 1; u < 4; u++) acc[c][u] = q == 0 ? t[u] : 0.0f;
    }
}

typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 b2v __attribute__((ext_vector_type(2)));

// x = hi + lo + O(2^-16 |x|): hi the BF16 rounding of x (hardware
// round-to-nearest-even), lo that of the remainder, packed as the A
// operand of a 16x16x16 BF16 MFMA (K index u = element u)
__device__ __forceinline__ void split_bf16x4(const float (&x)[4], s4v &hi, s4v &lo) {
    const uint32_t h01 = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){x[0], x[1]}, b2v));
    const uint32_t h23 = __builtin_bit_cast(uint32_t, __builtin_convertvector((f2v){x[2], x[3]}, b2v));
    const f2v r01 = {x[0] - __builtin_bit_cast(float, h01 << 16), x[1] - __builtin_bit_cast(float, h01 & 0xffff0000u)};
    const f2v r23 = {x[2] - __builtin_bit_cast(float, h23 << 16), x[3] - __builtin_bit_cast(float, h23 & 0xffff0000u)};
    const uint32_t l01 = __builtin_bit_cast(uint32_t, __builtin_convertvector(r01, b2v));
    const uint32_t l23 = __builtin_bit_cast(uint32_t, __builtin_convertvector(r23, b2v));
    hi = __builtin_bit_cast(s4v, (i32x2m){(int)h01, (int)h23});
    lo = __builtin_bit_cast(s4v, (i32x2m){(int)l01, (int)l23});
}

// The contraction of input block KB under the static schedule: D[t] +=
// A_c B_(KB,c,t) for every pair of the layout, NCH channels; rd(o) reads
// this lane's 8 bytes of the block at byte o of its section (hi parts,
// lo parts lo bytes further).  BND: the three boundary channels against
// the boundary blocks.  NS 1..3: BF16 products (1: A_hi B_hi, 2: + A_lo
// B_hi, 3: + A_hi B_lo); NS 4: FP32 products on the FP32 matrix cores,
// the blocks in FP32 (1 KB each, block index times 1024 bytes; rd(o) then
// reads this lane's 16 bytes): slice u of v_mfma_f32_16x16x4f32 takes
// element u of this lane's A values, K index 4 (lane / 16) + u, against
// B rows 4 (lane / 16) + u, so four products cover the 16 K indices with
// the A values as they are.
template <int SL, int NCH, int KB, bool BND, int NS, typename RD>
__device__ __forceinline__ void contract5s(const float (&acc)[NCH][4], RD &&rd, uint32_t lo, f4v (&D)[cl_nblk(SL)]) {
    if constexpr (NS == 4) {
        constexpr int NT = cl_nblk(SL);
        // the next channel's blocks load while this channel's products run
        // (a product right behind its LDS read would wait out the read)
        f4v bq[2][NT];
        auto load = [&](auto cc) __attribute__((always_inline)) {
            constexpr int c = decltype(cc)::value;
            static_for<NT>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                if constexpr (BND ? cl_bnd_pair(SL, KB, t) : cl_vol_pair(SL, KB, c, t))
                    bq[c & 1][t] = rd(1024u * (uint32_t)(BND ? cl_bnd_index(SL, KB, c, t)
                                                            : cl_vol_index(SL, NCH, KB, c, t)));
            });
        };
        load(std::integral_constant<int, 0>{});
        static_for<NCH>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if constexpr (c + 1 < NCH) load(std::integral_constant<int, c + 1>{});
            // slices outermost: consecutive products go to different output
            // blocks where the channel couples to more than one
#pragma unroll
            for (int u = 0; u < 4; u++)
                static_for<NT>([&](auto tt) {
                    constexpr int t = decltype(tt)::value;
                    if constexpr (BND ? cl_bnd_pair(SL, KB, t) : cl_vol_pair(SL, KB, c, t))
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(acc[c][u], bq[c & 1][t][u], D[t], 0, 0, 0);
                });
            __builtin_amdgcn_sched_barrier(0);
        });
    } else {
        static_for<NCH>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            s4v ah, al;
            split_bf16x4(acc[c], ah, al);
            static_for<cl_nblk(SL)>([&](auto tt) {
                constexpr int t = decltype(tt)::value;
                constexpr bool pair = BND ? cl_bnd_pair(SL, KB, t) : cl_vol_pair(SL, KB, c, t);
                if constexpr (pair) {
                    constexpr uint32_t off = 512u * (uint32_t)(BND ? cl_bnd_index(SL, KB, c, t)
                                                                   : cl_vol_index(SL, NCH, KB, c, t));
                    const s4v bh = rd(off);
                    D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, D[t], 0, 0, 0);
                    if constexpr (NS > 1) {
                        D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, D[t], 0, 0, 0);
                        if constexpr (NS > 2) {
                            const s4v bl = rd(off + lo);
                            D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, D[t], 0, 0, 0);
                        }
                    }
                }
            });
            // one channel's B reads at a time: hoisted above all the MFMAs they
            // would hold every block in registers at once
            __builtin_amdgcn_sched_barrier(0);
        });
    }
}

// Variant 5 under the static class layout SL (5 or 4; the vectors'
// harmonics in that class order, slm = cl_slm(SL)).  a.tab: the volume
// blocks' hi parts in cl_vol_index order, their lo parts, then the
// boundary blocks likewise (cl_bnd_index); with NS 4 the FP32 volume
// blocks, then the FP32 boundary blocks (the same bytes); the volume
// section is staged in LDS.  Lanes, tiles, grid and epilogues as k_apply5.
template <int NC, int WT, int SL, int NS, int EPI>
__global__ __launch_bounds__(512) void k_apply5s(OpArgs<float> a, const float *__restrict__ f,
                                                 float *__restrict__ r, double *__restrict__ part,
                                                 const double *__restrict__ skip, ChebStep<float> cs) {
    constexpr bool DOT = EPI == 1 || EPI == 3;
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int NT = cl_nblk(SL), RS = rec_stride<float>(NC), PE = (RS + 3) / 4;
    constexpr uint32_t BLK = 512, SLM = cl_slm(SL);
    constexpr uint32_t VLO = (uint32_t)cl_vol_count(SL, NC) * BLK, BLO = (uint32_t)cl_bnd_count(SL) * BLK;
    // slots of neighbour rows in flight, carried across the harmonic blocks
    constexpr int DEPTH = 6, NG = NT * WT, DEPTHB = 4;
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t nP = a.nP;
    {
        const i32x4m *src = reinterpret_cast<const i32x4m *>(a.tab);
        i32x4m *dst = reinterpret_cast<i32x4m *>(smem_raw);
        for (uint32_t i = threadIdx.x; i < 2u * VLO / 16u; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    // the boundary blocks, read through a buffer descriptor: the constant
    // block offsets ride in the instructions, not in per-block addresses
    const auto gbrs = rsrc4(reinterpret_cast<const unsigned char *>(a.tab) + 2u * VLO, 2u * BLO);
    const auto frs = rsrc4(f, (size_t)nP * SLM * sizeof(float));
    const auto brs = rsrc4(a.bof, (size_t)nP * 4);
    const auto nrs = rsrc4(a.nbr, (size_t)a.W * nP * 4);
    const auto bnrs = rsrc4(a.bnbr, (size_t)TR_WBMAX * a.nB * 4);
    const auto crs = rsrc4(a.crec, (size_t)a.W * nP * RS * sizeof(float));
    const auto bcrs = rsrc4(a.bcrec, (size_t)TR_WBMAX * a.nB * 4 * sizeof(float));
    const int vloc = lane & 15, q = lane >> 4;
    const int gv = lane >> 2, gq = lane & 3;

    const uint32_t ntiles = (nP + 16u * wpb - 1) / (16u * wpb);
    const uint32_t G = gridDim.x, X = 8;
    uint32_t t0 = blockIdx.x, t1 = ntiles, tstep = G;
    if (G % X == 0 && G <= ntiles) {
        const uint32_t xcd = blockIdx.x % X, per = ntiles / X, rem = ntiles % X;
        const uint32_t lo = xcd * per + (xcd < rem ? xcd : rem);
        t0 = lo + blockIdx.x / X;
        t1 = lo + per + (xcd < rem ? 1u : 0u);
        tstep = G / X;
    }
    const uint32_t nwt = (nP + 15u) / 16u;
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
    double d = 0;
    for (uint32_t tile = t0; tile < t1; tile += tstep) {
        const uint32_t e = tile * (uint32_t)wpb + wv;
        const uint32_t v0 = 16u * ((a.tiles && e < nwt) ? a.tiles[e] : e);
        uint32_t v = v0 + gv;
        asm volatile("" : "+v"(v));
        const bool valid = v < nP;
        // lanes past nP gather row 0 with zero coefficients
        const uint32_t vs = valid ? v : 0u;
        // byte offset of this lane's 16-byte piece of each slot's neighbour row
        uint32_t ro[WT];
        {
            uint32_t nbc[WT];
            load_idx<WT>(nbc, nrs, vs * 4u, nP * 4u, WT);
#pragma unroll
            for (int s2 = 0; s2 < WT; s2++) ro[s2] = nbc[s2] * (SLM * 4u) + 16u * (uint32_t)gq;
        }
        const int bl = (int)__builtin_amdgcn_raw_buffer_load_b32(brs, vs * 4u, 0u, 0);
        const int b = valid ? bl : -1;
        // this lane's quarter of the slots' coefficient records; lanes past
        // nP read past the record array, which returns zeros
        uint32_t cv = valid ? vs * (uint32_t)(RS * sizeof(float)) + (uint32_t)(gq * PE * (int)sizeof(float))
                            : 0x80000000u;
        const uint32_t cstride = opaque_uniform(nP * (uint32_t)(RS * sizeof(float)));
        asm volatile("" : "+v"(cv));
        auto ldc = [&](int s2, float *c) __attribute__((always_inline)) {
#pragma unroll
            for (int e2 = 0; e2 < PE; e2++)
                c[e2] = bload1<float, kNonTemporal>(crs, cv + (uint32_t)(e2 * (int)sizeof(float)),
                                                    (uint32_t)s2 * cstride);
        };
        // the records stream through non-temporal loads: read once per
        // application, they would otherwise push out of the caches the
        // vectors that the solver's next passes read (the mixed solve on
        // the 96^3 ball: 0.353 against 0.373 s).  The FP64 operator's
        // records load normally: there the non-temporal form was slower
        // (0.888 against 0.851 s for the plain FP64 solve).
        // records held in registers across the blocks (reloading them with
        // each block's rows frees registers for a third wavefront per SIMD,
        // but measured slower: 490 against 370 us on the 96^3 ball)
        float cf[WT][PE];
#pragma unroll
        for (int s2 = 0; s2 < WT; s2++) ldc(s2, cf[s2]);
        f4v D[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) D[t] = (f4v){0, 0, 0, 0};
        // a ring of DEPTH neighbour-row buffers runs through all the
        // harmonic blocks: the next block's first rows load while the
        // current one contracts
        // the fused dot's f at this lane's outputs (rows 4q + i, column
        // vloc of each output block), loaded now so the reads complete
        // during the gather instead of at the write
        float fo[EPI == 1 ? NT : 1][4];
        if constexpr (EPI == 1) {
            const auto fr2 = rsrc4(f, (size_t)nP * SLM * sizeof(float));
#pragma unroll
            for (int t = 0; t < NT; t++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t vr = v0 + (uint32_t)(4 * q + i);
                    const bool ok = vloc < cl_w(SL, t) && vr < nP;
                    // rows past nP read past the vector: zeros
                    fo[t][i] = __builtin_bit_cast(
                        float, __builtin_amdgcn_raw_buffer_load_b32(
                                   fr2, ok ? (vr * SLM + (uint32_t)cl_h0(SL, t) + (uint32_t)vloc) * 4u : 0xfffffff0u, 0u, 0));
                }
        }
        float xs[DEPTH][4];
        const uint32_t nq = 12u * (uint32_t)gq, nq4 = 16u * (uint32_t)gq;
        // ro[] holds this lane's 16-byte piece (+16 gq bytes); the five-wide
        // block reads harmonics h0 + gq and h0 + 4 instead
        auto ld = [&](int g, float *x) __attribute__((always_inline)) {
            const int kb = g / WT, s2 = g % WT;
            if (cl_w(SL, kb) == 5) {
                x[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                     frs, ro[s2] + 4u * (uint32_t)cl_h0(SL, kb) - nq, 0u, 0));
                x[1] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                     frs, ro[s2] + 4u * (uint32_t)cl_h0(SL, kb) + 16u - nq4, 0u, 0));
                x[2] = 0; x[3] = 0;
                return;
            }
            const f32x4m t4 = __builtin_bit_cast(
                f32x4m, __builtin_amdgcn_raw_buffer_load_b128(frs, ro[s2] + 4u * (uint32_t)cl_h0(SL, kb), 0u, 0));
            x[0] = t4.x;
            x[1] = t4.y;
            x[2] = t4.z;
            x[3] = t4.w;
        };
#pragma unroll
        for (int g = 0; g < DEPTH; g++) ld(g, xs[g]);
        static_for<NT>([&](auto kbc) {
            constexpr int kb = decltype(kbc)::value;
            float acc[NC][4];
            if constexpr (cl_w(SL, kb) == 5) {
                float acc1[NC], acc4[PE];
#pragma unroll
                for (int s2 = 0; s2 < WT; s2++) {
                    const int g = kb * WT + s2;
                    const float *cw = cf[s2];
                    if (s2 == 0)
                        slot_accum5<PE, NC, true>(acc1, acc4, cw, xs[g % DEPTH]);
                    else
                        slot_accum5<PE, NC, false>(acc1, acc4, cw, xs[g % DEPTH]);
                    if (g + DEPTH < NG) ld(g + DEPTH, xs[g % DEPTH]);
                }
                narrow_to_mfma<PE, NC>(acc1, acc4, acc, lane);
            } else {
#pragma unroll
                for (int s2 = 0; s2 < WT; s2++) {
                    const int g = kb * WT + s2;
                    const float *cw = cf[s2];
                    if (s2 == 0)
                        slot_accum<PE, NC, true>(acc, cw, xs[g % DEPTH]);
                    else
                        slot_accum<PE, NC, false>(acc, cw, xs[g % DEPTH]);
                    if (g + DEPTH < NG) ld(g + DEPTH, xs[g % DEPTH]);
                }
            }
            // harmonics past the block's width read the next block's values
            if constexpr (cl_w(SL, kb) == 5) {
                // the five-wide block is in the MFMA layout already (K index
                // 4 (lane / 16) + j), its entries past K 4 zero: this leaves
                // them so.  Without these selects the compiler's schedule of
                // the kernel needs 256 VGPRs and spills (measured); with
                // them 199
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (4 * (lane >> 4) + j >= 5)
#pragma unroll
                        for (int c = 0; c < NC; c++) acc[c][j] = 0;
            } else if constexpr (cl_w(SL, kb) < 16) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (4 * gq + j >= cl_w(SL, kb))
#pragma unroll
                        for (int c = 0; c < NC; c++) acc[c][j] = 0;
            }
            if constexpr (kb == 0)
                if (a.mf0 && valid && gq == 0) acc[0][0] *= a.mf0[v];
            if constexpr (cl_w(SL, kb) != 5) to_mfma_layout<float, 2, 4, NC>(acc, lane);
            if constexpr (NS == 4) {
                const unsigned char *lb = smem_raw + (uint32_t)lane * 16u;
                contract5s<SL, NC, kb, false, NS>(
                    acc,
                    [&](uint32_t o) __attribute__((always_inline)) { return *reinterpret_cast<const f4v *>(lb + o); },
                    0u, D);
            } else {
                const unsigned char *lb = smem_raw + (uint32_t)lane * 8u;
                contract5s<SL, NC, kb, false, NS>(
                    acc,
                    [&](uint32_t o) __attribute__((always_inline)) { return *reinterpret_cast<const s4v *>(lb + o); },
                    VLO, D);
            }
        });
        if (__any(b >= 0)) {
            uint32_t bn[TR_WBMAX];
            if (b >= 0) load_idx<TR_WBMAX>(bn, bnrs, (uint32_t)b * 4u, a.nB * 4u, TR_WBMAX);
            static_for<NT>([&](auto kbc) {
                constexpr int kb = decltype(kbc)::value;
                float bacc[3][4];
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int j = 0; j < 4; j++) bacc[c][j] = 0;
                if (b >= 0)
                    gather<float, 4, 3, 4, TR_WBMAX, DEPTHB>(bacc, bn, frs, bcrs, a.nB * (uint32_t)(4 * sizeof(float)),
                                                            (uint32_t)b * (uint32_t)(4 * sizeof(float)), TR_WBMAX,
                                                            SLM, gq, 4u * (uint32_t)cl_h0(SL, kb), 0u,
                                                            (uint32_t)cl_w(SL, kb));
                to_mfma_layout<float, 2, 4, 3>(bacc, lane);
                if constexpr (NS == 4) {
                    const uint32_t lo16 = (uint32_t)lane * 16u;
                    contract5s<SL, 3, kb, true, NS>(
                        bacc,
                        [&](uint32_t o) __attribute__((always_inline)) {
                            return __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(gbrs, lo16, o, 0));
                        },
                        0u, D);
                } else {
                    const uint32_t lo8 = (uint32_t)lane * 8u;
                    contract5s<SL, 3, kb, true, NS>(
                        bacc,
                        [&](uint32_t o) __attribute__((always_inline)) {
                            return __builtin_bit_cast(s4v, __builtin_amdgcn_raw_buffer_load_b64(gbrs, lo8, o, 0));
                        },
                        BLO, D);
                }
            });
        }
        // ---- write: D rows 4(l/16) + i, column l % 16 of output block t
        static_for<NT>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            const uint32_t n = (uint32_t)cl_h0(SL, t) + (uint32_t)vloc;
            const bool col_ok = vloc < cl_w(SL, t);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t vr = v0 + (uint32_t)(4 * q + i);
                if (col_ok && vr < nP) {
                    const float val = D[t][i];
                    const size_t o = (size_t)vr * SLM + n;
                    if constexpr (EPI >= 2) {
                        const float rh = cs.rhs[o];
                        float nv = cs.zc * f[o] + cs.ga * cs.dinv[o] * (rh - val);
                        if (cs.om != 0.0f) nv -= cs.om * r[o];
                        r[o] = nv;
                        if (EPI == 3) d = fma((double)rh, (double)nv, d);
                    } else {
                        r[o] = val;
                        if constexpr (EPI == 1) d = fma((double)fo[t][i], (double)val, d);
                    }
                }
            }
        });
    }
    if (DOT) {
        double *w = reinterpret_cast<double *>(smem_raw);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) d += __shfl_down(d, off, TR_WAVE);
        __syncthreads();
        if (lane == 0) w[wave] = d;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int i = 0; i < wpb; i++) t += w[i];
            part[blockIdx.x] = t;
        }
    }
}

// the static class-layout kernel for layout SL (Blk5::stat), BF16 split
// ns (1..3) or FP32 products (4) and epilogue EPI: ten or seven channels,
// 15 slots
template <int EPI>
static const void *apply5s_kernel(int NC, uint32_t W, uint32_t SL, int ns) {
#define TR_K5S(NCC, SLL, NSS) reinterpret_cast<const void *>(&k_apply5s<NCC, 15, SLL, NSS, EPI>)
#define TR_K5SN(NCC, SLL)                                                                                    \
    (ns == 1 ? TR_K5S(NCC, SLL, 1)                                                                          \
             : (ns == 2 ? TR_K5S(NCC, SLL, 2) : (ns == 3 ? TR_K5S(NCC, SLL, 3) : TR_K5S(NCC, SLL, 4))))
#define TR_K5SL(NCC) (SL == 5 ? TR_K5SN(NCC, 5) : TR_K5SN(NCC, 4))
    if (W != 15 || (NC != 7 && NC != 10) || (SL != 4 && SL != 5) || ns < 1 || ns > 4) return nullptr;
    return NC == 7 ? TR_K5SL(7) : TR_K5SL(10);
#undef TR_K5SL
#undef TR_K5SN
#undef TR_K5S
}

const void *apply5s_kernel_epi(int epi, int NC, uint32_t W, uint32_t SL, int ns) {
    switch (epi) {
        case 1: return apply5s_kernel<1>(NC, W, SL, ns);
        case 2: return apply5s_kernel<2>(NC, W, SL, ns);
        case 3: return apply5s_kernel<3>(NC, W, SL, ns);
        default: return apply5s_kernel<0>(NC, W, SL, ns);
    }
}

}  // namespace tr
