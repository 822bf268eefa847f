// kernels_mfma.hip -- the operator with its angular contraction on the matrix
// cores (schedule variant 4).
//
// A wavefront holds 16 vertices, four lanes (chunks q = 0..3) per vertex.
// Phase 1 gathers the ELL neighbours and accumulates the spatial channels
// T_c[v][m] in registers, register j of chunk q holding harmonic
// mfma_harm(E, q, j) (kernels.h).  In this phase the four lanes of a vertex
// are adjacent (lane 4v + q): each lane reads 16-byte pieces of the
// neighbour row, the four together 64 contiguous bytes, so the L1 cache
// serves a vertex's share of a load with one tag lookup rather than four
// (the row gathers, not HBM bandwidth, bound this kernel otherwise); each
// lane also loads only a quarter of the slot's coefficient record and the
// quad shares it through DPP broadcasts.  Phase 2 is
//   R[v][n] = sum_c sum_m T_c[v][m] Mat_c[n][m]
// (c = mass (Mat = diag(massfac)), the NC-1 stiffness channels and, for
// wavefronts holding boundary vertices, the 3 boundary channels), a
// [16 vertices x K] x [K x 16 outputs] product per output tile, on the
// matrix cores.  The accumulators move once per tile to the MFMA operand
// layout (lane v + 16q, v = lane % 16 the A row), where acc[c][j] of lane l
// is exactly the A operand of one MFMA step, so the channel sums never
// leave the registers.  B slices (4 x 16 couplings per step and tile) are
// staged in LDS once per workgroup.
//   FP64 operator: v_mfma_f64_16x16x4f64, K = 4 per step (one j).
//   FP32 operator: v_mfma_f32_16x16x4f32 (exact FP32), or BF16 MFMA
//   v_mfma_f32_16x16x16bf16_1k with FP32 accumulation, K = 16 per step
//   (4 consecutive j of every chunk), optionally with the hi/lo split
//   A_hi B_hi + A_lo B_hi + A_hi B_lo for ~16 significant bits.
#include "mfma_common.h"

namespace tr {

// MODE: 0 = FP64 (T = double), 1 = FP32 MFMA, 2 = BF16 MFMA (T = float).
// A persistent grid: every workgroup stages the B slices in LDS once
// and then walks vertex tiles (WPB wavefronts x 16 vertices).  With a grid
// that is a multiple of the 8 XCDs, the workgroups of one XCD (block ids
// congruent mod 8) share a contiguous range of tiles, so neighbouring
// vertices' data stay in that XCD's L2.
// EPI: 0 = r = A f; 1 = the same and the fused dot product f . r (one
// FP64 partial per workgroup in part); 2 = one Chebyshev step fused into the
// write, r = zc f - om r + ga dinv (rhs - A f) (r on entry holds the
// previous iterate, not read when om == 0); 3 = the Chebyshev step and the
// dot product rhs . r of its result
template <typename T, int MODE, int J, int NC, int WT, int EPI>
__global__ __launch_bounds__(512) void k_apply4(OpArgs<T> a, const T *__restrict__ f,
                                                T *__restrict__ r, double *__restrict__ part,
                                                const double *__restrict__ skip, ChebStep<T> cs) {
    constexpr bool DOT = EPI == 1 || EPI == 3;
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RS = rec_stride<T>(NC);
    constexpr int S = mfma_steps(MODE, J);
    constexpr int BW = (MODE == 2) ? 8 : (int)sizeof(T);
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t slm = a.slm, nP = a.nP;
    // output tiles of 16 harmonics: ceil(slm / 16), which the accepted
    // (J, slm) pairs make J / 4 rounded up
    constexpr int NT = (J * 4 + 15) / 16;
    const int nsplit = a.nsplit;
    const size_t vsec = (size_t)NC * S * NT * TR_WAVE * BW;  // one part of the volume B table
    const size_t bsec = (size_t)3 * S * NT * TR_WAVE * BW;   // one part of the boundary B table
    const int parts = (MODE == 2 && nsplit > 1) ? 2 : 1;
    {
        // volume then boundary slices, every part: the whole table
        const uint32_t nall = (uint32_t)((vsec + bsec) * parts);
        const i32x4m *src = reinterpret_cast<const i32x4m *>(a.tab);
        i32x4m *dst = reinterpret_cast<i32x4m *>(smem_raw);
        for (uint32_t i = threadIdx.x; i < nall / 16; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const auto frs = rsrc4(f, (size_t)nP * slm * sizeof(T));
    const auto brs = rsrc4(a.bof, (size_t)nP * 4);
    const auto nrs = rsrc4(a.nbr, (size_t)a.W * nP * 4);

    const auto bnrs = rsrc4(a.bnbr, (size_t)TR_WBMAX * a.nB * 4);
    const auto crs = rsrc4(a.crec, (size_t)a.W * nP * RS * sizeof(T));
    const auto bcrs = rsrc4(a.bcrec, (size_t)TR_WBMAX * a.nB * 4 * sizeof(T));

    // MFMA layout (contraction and write): vertex vloc, chunk q; gather
    // layout: vertex gv, chunk gq (the vertex's four lanes adjacent)
    const int vloc = lane & 15, q = lane >> 4;
    const int gv = lane >> 2, gq = lane & 3;
    typedef typename std::conditional<MODE == 0, d4v, f4v>::type acc4;
    // register budget per lane: 2 wavefronts per SIMD for FP64 (256 VGPRs),
    // 3 for FP32 (168); OVH covers the indices, the MFMA accumulators and
    // addresses (fitted to the compiler's allocation; 5 slots in flight at
    // NC = 7, J = 8 time best, 6 no better)
    constexpr int BUDGET = sizeof(T) == 8 ? 256 : 168;
    constexpr int OVH = WT + 8 * NT + (sizeof(T) == 8 ? 12 : (MODE == 2 ? 26 : 22));
    constexpr int PE = (RS + 3) / 4;
    constexpr int ACCR = NC * J * (int)sizeof(T) / 4, SLOTR = (J + PE) * (int)sizeof(T) / 4;
    constexpr int DEPTH = std::max(2, std::min(6, (BUDGET - ACCR - OVH) / SLOTR));
    constexpr int BACCR = 3 * J * (int)sizeof(T) / 4, BSLOTR = (J + 1) * (int)sizeof(T) / 4;
    constexpr int DEPTHB = std::max(2, std::min(4, (BUDGET - BACCR - OVH - 8) / BSLOTR));

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
    // first vertex of this wavefront's share of a tile (a.tiles: the order of
    // the 16-vertex wavefront tiles), the neighbour indices and the boundary
    // row of this lane's vertex (gather layout)
    const uint32_t nwt = (nP + 15u) / 16u;
    const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
    auto tile_idx = [&](uint32_t tile, uint32_t &vb, uint32_t (&nb)[WT], int &b) {
        const uint32_t e = tile * (uint32_t)wpb + wv;
        vb = 16u * ((a.tiles && e < nwt) ? a.tiles[e] : e);
        uint32_t v = vb + gv;
        asm volatile("" : "+v"(v));
        if (v < nP) {
            load_idx<WT>(nb, nrs, v * 4u, nP * 4u, WT);
            b = (int)__builtin_amdgcn_raw_buffer_load_b32(brs, v * 4u, 0u, 0);
        } else {
#pragma unroll
            for (int s = 0; s < WT; s++) nb[s] = 0;
            b = -1;
        }
    };
    uint32_t nbc[WT], vcur = nP;
    int bcur = -1;
    if (t0 < t1) tile_idx(t0, vcur, nbc, bcur);
    double d = 0;
    for (uint32_t tile = t0; tile < t1; tile += tstep) {
        const uint32_t v0 = vcur;
        const uint32_t v = v0 + gv;
        const bool valid = v < nP;
        // ---- volume: gather, then the contraction on the matrix cores
        T acc[NC][J];
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int j = 0; j < J; j++) acc[c][j] = 0;
        // one descriptor per slot: a slot's records stay below 4 GiB where
        // all W of them need not (unions of many problems)
        if (valid)
            gather<T, J, NC, RS, WT, DEPTH, true>(acc, nbc, frs, crs, nP * (uint32_t)(RS * sizeof(T)),
                                                   v * (uint32_t)(RS * sizeof(T)), WT, slm, gq, 0u, 0u, 0xffffffffu,
                                                   a.crec);
        // per-vertex factor of the mass channel at lm = 0 (isotropic albedo
        // per problem in a batch: 1 - albedo, with massfac[0] = 1); harmonic
        // 0 is register 0 of chunk 0
        if (a.mf0 && valid && gq == 0) acc[0][0] *= a.mf0[v];
        // the Chebyshev step's rhs and dinv at this lane's outputs load while
        // the tile contracts (the gather's registers are free by now).  FP32
        // EPI 2 only: there it fits the 168 registers of 3 wavefronts per
        // SIMD; FP64 and the other epilogues would lose occupancy
        constexpr int NPF = (sizeof(T) == 4 && EPI == 2) ? 2 : 0;
        T pf[NPF > 0 ? NPF : 1][NT][4];
        if constexpr (NPF > 0) {
            // not above the gather, whose buffers would then share registers with them
            asm volatile("" ::: "memory");
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const uint32_t n = 16u * t + (uint32_t)vloc;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t vr = v0 + (MODE == 0 ? (uint32_t)(q + 4 * i) : (uint32_t)(4 * q + i));
                    const bool ok = n < slm && vr < nP;
                    const size_t o = (size_t)vr * slm + n;
                    if constexpr (EPI >= 2) {
                        pf[0][t][i] = ok ? cs.rhs[o] : (T)0;
                        pf[1][t][i] = ok ? cs.dinv[o] : (T)0;
                    } else {
                        pf[0][t][i] = ok ? f[o] : (T)0;
                    }
                }
            }
        }
        to_mfma_layout<T, MODE, J, NC>(acc, lane);
        // the next tile's indices load while this tile contracts
        const int b = bcur;
        if (tile + tstep < t1) tile_idx(tile + tstep, vcur, nbc, bcur);
        acc4 D[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) D[t] = (acc4){0, 0, 0, 0};
        // the B slices are the same for every tile: an opaque zero offset
        // keeps the optimiser from hoisting all of their loads out of the
        // tile loop into registers
        uint32_t boff = 0;
        asm volatile("" : "+v"(boff));
        contract<T, MODE, J, NC>(acc, smem_raw + boff, vsec, lane, nsplit, D);
        // ---- vacuum boundary: only wavefronts that hold boundary vertices
        if (__any(b >= 0)) {
            T bacc[3][J];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int j = 0; j < J; j++) bacc[c][j] = 0;
            if (b >= 0) {
                uint32_t bn[TR_WBMAX];
                // all TR_WBMAX slots (the host pads the boundary stencil with
                // zero-coefficient slots): a fixed count keeps the pipelined
                // gather free of branches, whose merges would make the
                // compiler wait for every load in flight
                load_idx<TR_WBMAX>(bn, bnrs, (uint32_t)b * 4u, a.nB * 4u, TR_WBMAX);
                gather<T, J, 3, 4, TR_WBMAX, DEPTHB>(bacc, bn, frs, bcrs, a.nB * (uint32_t)(4 * sizeof(T)),
                                                     (uint32_t)b * (uint32_t)(4 * sizeof(T)), TR_WBMAX, slm, gq);
            }
            to_mfma_layout<T, MODE, J, 3>(bacc, lane);
            contract<T, MODE, J, 3>(bacc, smem_raw + vsec * parts + boff, bsec, lane, nsplit, D);
        }
        // ---- write: FP64 D holds rows (l/16) + 4i, FP32 rows 4(l/16) + i; column l % 16
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const uint32_t n = 16u * t + (uint32_t)vloc;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t vr = v0 + (MODE == 0 ? (uint32_t)(q + 4 * i) : (uint32_t)(4 * q + i));
                if (n < slm && vr < nP) {
                    const T val = (T)D[t][i];
                    const size_t o = (size_t)vr * slm + n;
                    if constexpr (EPI >= 2) {
                        const T rh = NPF ? pf[0][t][i] : cs.rhs[o];
                        T nv = cs.zc * f[o] + cs.ga * (NPF ? pf[1][t][i] : cs.dinv[o]) * (rh - val);
                        if (cs.om != (T)0) nv -= cs.om * r[o];
                        r[o] = nv;
                        if (EPI == 3) d = fma((double)rh, (double)nv, d);
                    } else {
                        r[o] = val;
                        if (DOT) d = fma((double)(NPF ? pf[0][t][i] : f[o]), (double)val, d);
                    }
                }
            }
        }
    }
    if (DOT) {
        // the B tables are done with: their LDS takes the wavefront sums (a
        // separate array would push two FP64 workgroups past the LDS of a CU)
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

// Variant 5: the FP32 operator for 33 to 80 harmonics (K = 4, 5) with the
// angular contraction on the matrix cores in BF16.  At these orders one
// lane cannot hold a quarter of every channel's row (10 channels x 20
// harmonics), so the harmonics go through the kernel in NT blocks of at
// most 16 (a.b5: block kb is harmonics h0 .. h0 + w - 1 of the vectors'
// order), the K dimension of v_mfma_f32_16x16x16bf16: per block the ELL
// gather of one 16-byte piece per lane (a vertex's four lanes read a
// 64-byte run of each neighbour row), the move to the MFMA layout, and
// D[t] += A_c B_c[kb][t] for every channel c and output block t whose B
// block is not all zero, the D accumulators covering all output blocks
// throughout.
//
// When the vectors' harmonics are in reflection-parity classes (cos or sin
// in m, times the parity of m: 21/15/15/15 harmonics at K = 5), every block
// lies in one class and each channel couples a class to one class only
// (the mirror parities of Omega_a Omega_b), so only 62 of the 250 B blocks
// at K = 5 with ten channels are nonzero: they are all that is stored and
// multiplied.  The stored volume blocks are staged in LDS once per
// workgroup of the persistent grid, BF16 hi parts then (nsplitB 2) lo
// parts; the boundary blocks are read from global memory by the wavefronts
// that hold boundary vertices.  a.nsplit: 1 = A_hi B_hi, 2 = + A_lo B_hi,
// 3 = + A_hi B_lo (~16 significant bits).  EPI as k_apply4.
template <int NC, int NT>
__device__ __forceinline__ void contract5(const float (&acc)[NC][4], const unsigned char *B, uint32_t lo,
                                          uint64_t mask, uint32_t off, int lane, int nsplit, f4v (&D)[NT]) {
    constexpr size_t BLK = (size_t)TR_WAVE * 8;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (!((mask >> (c * NT)) & ((1ull << NT) - 1))) continue;
        f4v x;
#pragma unroll
        for (int u = 0; u < 4; u++) x[u] = acc[c][u];
        const bf4v hb = __builtin_convertvector(x, bf4v);
        const s4v ah = __builtin_bit_cast(s4v, hb);
        const s4v al = __builtin_bit_cast(s4v, __builtin_convertvector(x - __builtin_convertvector(hb, f4v), bf4v));
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const uint32_t bit = (uint32_t)(c * NT + t);
            if (!((mask >> bit) & 1ull)) continue;
            const uint32_t idx = off + (uint32_t)__builtin_popcountll(mask & ((1ull << bit) - 1ull));
            const unsigned char *bp = B + (size_t)idx * BLK + (size_t)lane * 8;
            const s4v bh = *reinterpret_cast<const s4v *>(bp);
            D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, D[t], 0, 0, 0);
            if (nsplit > 1) {
                D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, D[t], 0, 0, 0);
                if (nsplit > 2) {
                    const s4v bl = *reinterpret_cast<const s4v *>(bp + lo);
                    D[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, D[t], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <int NC, int WT, int NT, int EPI>
__global__ __launch_bounds__(512) void k_apply5(OpArgs<float> a, const float *__restrict__ f,
                                                float *__restrict__ r, double *__restrict__ part,
                                                const double *__restrict__ skip, ChebStep<float> cs) {
    constexpr bool DOT = EPI == 1 || EPI == 3;
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RS = rec_stride<float>(NC);
    constexpr int J = 4;
    constexpr size_t BLK = (size_t)TR_WAVE * 8;               // one 16 x 16 BF16 B block
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t slm = a.slm, nP = a.nP;
    const int nsplit = a.nsplit;
    const uint32_t vlo = a.b5.nvol * (uint32_t)BLK;            // volume lo parts after the hi parts
    const uint32_t vsec = vlo * a.b5.nsplitB;                  // bytes of the volume section
    const uint32_t blo = a.b5.nbnd * (uint32_t)BLK;
    {
        const i32x4m *src = reinterpret_cast<const i32x4m *>(a.tab);
        i32x4m *dst = reinterpret_cast<i32x4m *>(smem_raw);
        for (uint32_t i = threadIdx.x; i < vsec / 16u; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const unsigned char *gB = reinterpret_cast<const unsigned char *>(a.tab) + vsec;
    const auto frs = rsrc4(f, (size_t)nP * slm * sizeof(float));
    const auto brs = rsrc4(a.bof, (size_t)nP * 4);
    const auto nrs = rsrc4(a.nbr, (size_t)a.W * nP * 4);
    const auto bnrs = rsrc4(a.bnbr, (size_t)TR_WBMAX * a.nB * 4);
    const auto crs = rsrc4(a.crec, (size_t)a.W * nP * RS * sizeof(float));
    const auto bcrs = rsrc4(a.bcrec, (size_t)TR_WBMAX * a.nB * 4 * sizeof(float));
    const int vloc = lane & 15, q = lane >> 4;
    const int gv = lane >> 2, gq = lane & 3;
    // the tile's coefficient records (WT x PE per lane) stay in registers
    // across the harmonic blocks (two wavefronts per SIMD: capping the
    // registers at 168 for a third spills and measured slower, 523 against
    // 480 us on the 96^3 ball)
    constexpr int BUDGET = 256;
    constexpr int PE = (RS + 3) / 4;
    constexpr int OVH = WT + 4 * NT + 30 + WT * PE;
    constexpr int ACCR = NC * J, SLOTR = J;
    constexpr int DEPTH = std::max(2, std::min(8, (BUDGET - ACCR - OVH) / SLOTR));
    constexpr int DEPTHB = 4;

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
        uint32_t nbc[WT];
        int b = -1;
        if (valid) {
            load_idx<WT>(nbc, nrs, v * 4u, nP * 4u, WT);
            b = (int)__builtin_amdgcn_raw_buffer_load_b32(brs, v * 4u, 0u, 0);
        } else {
#pragma unroll
            for (int s2 = 0; s2 < WT; s2++) nbc[s2] = 0;
        }
        f4v D[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) D[t] = (f4v){0, 0, 0, 0};
        // lane q's quarter of every slot's coefficient record (reads past
        // the record's RS elements fall outside the descriptor: zeros)
        float cf[WT][PE];
        {
            uint32_t cv = v * (uint32_t)(RS * sizeof(float)) + (uint32_t)(gq * PE * (int)sizeof(float));
            const uint32_t cstride = opaque_uniform(nP * (uint32_t)(RS * sizeof(float)));
            asm volatile("" : "+v"(cv));
#pragma unroll
            for (int s2 = 0; s2 < WT; s2++)
#pragma unroll
                for (int e2 = 0; e2 < PE; e2++)
                    cf[s2][e2] = valid ? bload1<float>(crs, cv + (uint32_t)(e2 * (int)sizeof(float)), (uint32_t)s2 * cstride)
                                       : 0.0f;
        }
#pragma unroll 1
        for (int kb = 0; kb < NT; kb++) {
            float acc[NC][J];
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < J; j++) acc[c][j] = 0;
            const uint32_t h0 = a.b5.h0[kb], hw = a.b5.w[kb];
            if (valid) gather_pre<float, J, NC, PE, WT, DEPTH>(acc, nbc, frs, cf, slm, gq, 4u * h0, hw);
            if (a.mf0 && valid && gq == 0 && kb == 0) acc[0][0] *= a.mf0[v];
            to_mfma_layout<float, 2, J, NC>(acc, lane);
            contract5<NC, NT>(acc, smem_raw, vlo, a.b5.vmask[kb], a.b5.voff[kb], lane, nsplit, D);
        }
        if (__any(b >= 0)) {
            uint32_t bn[TR_WBMAX];
            if (b >= 0) load_idx<TR_WBMAX>(bn, bnrs, (uint32_t)b * 4u, a.nB * 4u, TR_WBMAX);
#pragma unroll 1
            for (int kb = 0; kb < NT; kb++) {
                if (!a.b5.bmask[kb]) continue;
                float bacc[3][J];
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int j = 0; j < J; j++) bacc[c][j] = 0;
                if (b >= 0)
                    gather<float, J, 3, 4, TR_WBMAX, DEPTHB>(bacc, bn, frs, bcrs, a.nB * (uint32_t)(4 * sizeof(float)),
                                                            (uint32_t)b * (uint32_t)(4 * sizeof(float)), TR_WBMAX,
                                                            slm, gq, 4u * a.b5.h0[kb], 0u, a.b5.w[kb]);
                to_mfma_layout<float, 2, J, 3>(bacc, lane);
                contract5<3, NT>(bacc, gB, blo, a.b5.bmask[kb], a.b5.boff[kb], lane, nsplit, D);
            }
        }
        // ---- write: D rows 4(l/16) + i, column l % 16 of output block t
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const uint32_t n = a.b5.h0[t] + (uint32_t)vloc;
            const bool col_ok = (uint32_t)vloc < a.b5.w[t];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t vr = v0 + (uint32_t)(4 * q + i);
                if (col_ok && vr < nP) {
                    const float val = D[t][i];
                    const size_t o = (size_t)vr * slm + n;
                    if constexpr (EPI >= 2) {
                        const float rh = cs.rhs[o];
                        float nv = cs.zc * f[o] + cs.ga * cs.dinv[o] * (rh - val);
                        if (cs.om != 0.0f) nv -= cs.om * r[o];
                        r[o] = nv;
                        if (EPI == 3) d = fma((double)rh, (double)nv, d);
                    } else {
                        r[o] = val;
                        if (DOT) d = fma((double)f[o], (double)val, d);
                    }
                }
            }
        }
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

// The FP64 contraction of one channel group (channels c0 .. c0 + CG - 1 of
// the operator) against the stored FP64 B blocks: block idx holds, for K
// step st and lane l, the coupling of output column h0_t + l % 16 with input
// harmonic h0_kb + mfma_harm(2, l / 16, st), 2 KB per block.
template <int CG, int NT>
__device__ __forceinline__ void contract5d(const double (&acc)[CG][4], const unsigned char *B, int c0,
                                           uint64_t mask, uint32_t off, int lane, d4v (&D)[NT]) {
    constexpr size_t BLK = (size_t)TR_WAVE * 8 * 4;
#pragma unroll
    for (int cc = 0; cc < CG; cc++) {
        const int c = c0 + cc;
        if (!((mask >> (c * NT)) & ((1ull << NT) - 1))) continue;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const uint32_t bit = (uint32_t)(c * NT + t);
            if (!((mask >> bit) & 1ull)) continue;
            const uint32_t idx = off + (uint32_t)__builtin_popcountll(mask & ((1ull << bit) - 1ull));
            const double *bp = reinterpret_cast<const double *>(B + (size_t)idx * BLK) + lane;
#pragma unroll
            for (int st = 0; st < 4; st++)
                D[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[cc][st], bp[st * TR_WAVE], D[t], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Variant 5 in FP64: the operator for 33 to 80 harmonics with its angular
// contraction on the FP64 matrix cores (v_mfma_f64_16x16x4f64, four K steps
// per 16-harmonic block), in the reflection-parity classes only (the stored
// blocks, 2 KB each, must fit the LDS: 56 at K = 5 with ten channels).
// The gather reads the vector fg, whose harmonics are in the class order;
// the results go to position hpo[n] of the output row for class position
// n (hpo null: the class order too).  The epilogue's reads (f, and the
// Chebyshev step's vectors) are in the output order.  FP64 records and
// accumulators need twice the registers of FP32, so the ten channels run
// in groups of CG (two groups of five; seven channels in one), each group
// with its quarter-records in registers across the harmonic blocks and the
// neighbour rows gathered once per group.  EPI as k_apply4.
template <int NC, int WT, int NT, int CG, int EPI>
__global__ __launch_bounds__(512) void k_apply5d(OpArgs<double> a, const double *__restrict__ fg,
                                                 const double *__restrict__ f, double *__restrict__ r,
                                                 double *__restrict__ part, const double *__restrict__ skip,
                                                 ChebStep<double> cs) {
    constexpr bool DOT = EPI == 1 || EPI == 3;
    static_assert(NC % CG == 0, "channel groups");
    constexpr int NG = NC / CG;
    if (skip && *skip != 0.0) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int RS = rec_stride<double>(NC);
    constexpr int J = 4;
    constexpr int PE = (CG + 3) / 4;
    const int lane = threadIdx.x & (TR_WAVE - 1), wave = threadIdx.x / TR_WAVE;
    const int wpb = blockDim.x / TR_WAVE;
    const uint32_t slm = a.slm, nP = a.nP;
    const uint32_t vsec = a.b5.nvol * (uint32_t)(TR_WAVE * 8 * 4);
    {
        const i32x4m *src = reinterpret_cast<const i32x4m *>(a.tab);
        i32x4m *dst = reinterpret_cast<i32x4m *>(smem_raw);
        for (uint32_t i = threadIdx.x; i < vsec / 16u; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const unsigned char *gB = reinterpret_cast<const unsigned char *>(a.tab) + vsec;
    const auto frs = rsrc4(fg, (size_t)nP * slm * sizeof(double));
    const auto brs = rsrc4(a.bof, (size_t)nP * 4);
    const auto nrs = rsrc4(a.nbr, (size_t)a.W * nP * 4);
    const auto bnrs = rsrc4(a.bnbr, (size_t)TR_WBMAX * a.nB * 4);
    const auto crs = rsrc4(a.crec, (size_t)a.W * nP * RS * sizeof(double));
    const auto bcrs = rsrc4(a.bcrec, (size_t)TR_WBMAX * a.nB * 4 * sizeof(double));
    const int vloc = lane & 15, q = lane >> 4;
    const int gv = lane >> 2, gq = lane & 3;
    // two wavefronts per SIMD (the LDS holds one 512-lane workgroup per CU):
    // 256 registers.  Three slots in flight is the most that compiles
    // without spills (the compiler's resource report: 4 spills 2 to 8
    // registers of the ten-channel kernels)
    constexpr int DEPTH = 3;
    constexpr int DEPTHB = 4;

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
        uint32_t nbc[WT];
        int b = -1;
        if (valid) {
            load_idx<WT>(nbc, nrs, v * 4u, nP * 4u, WT);
            b = (int)__builtin_amdgcn_raw_buffer_load_b32(brs, v * 4u, 0u, 0);
        } else {
#pragma unroll
            for (int s2 = 0; s2 < WT; s2++) nbc[s2] = 0;
        }
        d4v D[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) D[t] = (d4v){0, 0, 0, 0};
        // the groups unrolled: with g a run-time value the channels' mask
        // bits and block indices are too, and the compiler keeps them in
        // SGPRs (about a hundred spilled)
#pragma unroll
        for (int g = 0; g < NG; g++) {
            // lane q's share of the group's channels in every slot's record
            // (elements g CG + q PE .. + PE - 1; those past the group belong
            // to no channel the group uses)
            double cf[WT][PE];
            {
                uint32_t cv = v * (uint32_t)(RS * sizeof(double)) +
                              (uint32_t)((g * CG + gq * PE) * (int)sizeof(double));
                const uint32_t cstride = opaque_uniform(nP * (uint32_t)(RS * sizeof(double)));
                asm volatile("" : "+v"(cv));
#pragma unroll
                for (int s2 = 0; s2 < WT; s2++)
#pragma unroll
                    for (int e2 = 0; e2 < PE; e2++)
                        cf[s2][e2] = (valid && gq * PE + e2 < CG)
                                         ? bload1<double>(crs, cv + (uint32_t)(e2 * (int)sizeof(double)),
                                                          (uint32_t)s2 * cstride)
                                         : 0.0;
            }
#pragma unroll 1
            for (int kb = 0; kb < NT; kb++) {
                double acc[CG][J];
#pragma unroll
                for (int c = 0; c < CG; c++)
#pragma unroll
                    for (int j = 0; j < J; j++) acc[c][j] = 0;
                const uint32_t h0 = a.b5.h0[kb], hw = a.b5.w[kb];
                if (valid) gather_pre<double, J, CG, PE, WT, DEPTH>(acc, nbc, frs, cf, slm, gq, 8u * h0, hw);
                // per-vertex factor of the mass channel (channel 0) at
                // harmonic 0 (register 0 of chunk 0 in block 0)
                if (a.mf0 && valid && gq == 0 && kb == 0 && g == 0) acc[0][0] *= a.mf0[v];
                to_mfma_layout<double, 0, J, CG>(acc, lane);
                contract5d<CG, NT>(acc, smem_raw, g * CG, a.b5.vmask[kb], a.b5.voff[kb], lane, D);
            }
        }
        if (__any(b >= 0)) {
            uint32_t bn[TR_WBMAX];
            if (b >= 0) load_idx<TR_WBMAX>(bn, bnrs, (uint32_t)b * 4u, a.nB * 4u, TR_WBMAX);
#pragma unroll 1
            for (int kb = 0; kb < NT; kb++) {
                if (!a.b5.bmask[kb]) continue;
                double bacc[3][J];
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int j = 0; j < J; j++) bacc[c][j] = 0;
                if (b >= 0)
                    gather<double, J, 3, 4, TR_WBMAX, DEPTHB>(bacc, bn, frs, bcrs,
                                                              a.nB * (uint32_t)(4 * sizeof(double)),
                                                              (uint32_t)b * (uint32_t)(4 * sizeof(double)), TR_WBMAX,
                                                              slm, gq, 8u * a.b5.h0[kb], 0u, a.b5.w[kb]);
                to_mfma_layout<double, 0, J, 3>(bacc, lane);
                contract5d<3, NT>(bacc, gB, 0, a.b5.bmask[kb], a.b5.boff[kb], lane, D);
            }
        }
        // ---- write: D rows (l/16) + 4i, column l % 16 of output block t
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const uint32_t n = a.b5.h0[t] + (uint32_t)vloc;
            const bool col_ok = (uint32_t)vloc < a.b5.w[t];
            const uint32_t no = (col_ok && a.hpo) ? (uint32_t)a.hpo[n] : n;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t vr = v0 + (uint32_t)(q + 4 * i);
                if (col_ok && vr < nP) {
                    const double val = D[t][i];
                    const size_t o = (size_t)vr * slm + no;
                    if constexpr (EPI >= 2) {
                        const double rh = cs.rhs[o];
                        double nv = cs.zc * f[o] + cs.ga * cs.dinv[o] * (rh - val);
                        if (cs.om != 0.0) nv -= cs.om * r[o];
                        r[o] = nv;
                        if (EPI == 3) d = fma(rh, nv, d);
                    } else {
                        r[o] = val;
                        if (DOT) d = fma(f[o], val, d);
                    }
                }
            }
        }
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

// dst[v][k] = src[v][hp[k]] (inverse: dst[v][hp[k]] = src[v][k]), the
// harmonics of every row of n = nP slm entries
__global__ __launch_bounds__(256) void k_perm_harm(const double *__restrict__ src, double *__restrict__ dst,
                                                   const uint16_t *__restrict__ hp, uint32_t slm, size_t n,
                                                   int inverse, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t v = i / slm, k = i - v * slm;
        const size_t j = v * slm + hp[k];
        if (inverse)
            dst[j] = src[i];
        else
            dst[i] = src[j];
    }
}

hipError_t launch_perm_harm(const double *src, double *dst, const uint16_t *hp, uint32_t slm, size_t n,
                            bool inverse, const double *skip, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!src || !dst || !hp || src == dst || slm == 0) return hipErrorInvalidValue;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_perm_harm, dim3(g), dim3(256), 0, s, src, dst, hp, slm, n, inverse ? 1 : 0, skip);
    return hipGetLastError();
}

// LDS of the FP64 variant 5: the stored volume blocks, 2 KB each
size_t apply5d_lds_bytes(const Blk5 &b5) { return std::max<size_t>((size_t)b5.nvol * 2048, 64); }

template <int EPI>
static const void *apply5d_kernel(int NC, uint32_t W, uint32_t NT) {
#define TR_K5D(NCC, WW, NTT) reinterpret_cast<const void *>(&k_apply5d<NCC, WW, NTT, (NCC == 10 ? 5 : 7), EPI>)
#define TR_K5DT(NCC, WW) (NT == 3 ? TR_K5D(NCC, WW, 3) : (NT == 4 ? TR_K5D(NCC, WW, 4) : TR_K5D(NCC, WW, 5)))
#define TR_K5DW(NCC) (W == 15 ? TR_K5DT(NCC, 15) : TR_K5DT(NCC, 16))
    if ((W != 15 && W != 16) || (NC != 7 && NC != 10) || NT < 3 || NT > 5) return nullptr;
    return NC == 7 ? TR_K5DW(7) : TR_K5DW(10);
#undef TR_K5DW
#undef TR_K5DT
#undef TR_K5D
}

static const void *apply5d_kernel_epi(int epi, int NC, uint32_t W, uint32_t NT) {
    switch (epi) {
        case 1: return apply5d_kernel<1>(NC, W, NT);
        case 2: return apply5d_kernel<2>(NC, W, NT);
        case 3: return apply5d_kernel<3>(NC, W, NT);
        default: return apply5d_kernel<0>(NC, W, NT);
    }
}

uint32_t apply5d_grid(const OpArgs<double> &a, int NC, int wpb, int epi) {
    const uint32_t per = 16u * (uint32_t)wpb, ntiles = (a.nP + per - 1) / per;
    const void *k = apply5d_kernel_epi(epi, NC, a.W, a.b5.nb);
    if (!k || ntiles == 0) return ntiles;
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, TR_WAVE * wpb, apply5d_lds_bytes(a.b5)) != hipSuccess ||
        nb < 1 || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        return ntiles;
    }
    uint32_t g = (uint32_t)nb * (uint32_t)cus;
    if (g >= 8) g -= g % 8;
    return std::min(g, ntiles);
}

// The FP64 variant 5 on vectors in the internal (lm) order when a.cls is
// set: f is first permuted into the class order in a.cls (a scratch vector
// of the context) and the results written back through a.hpo; with a.cls
// null the vectors are in the class order already.
hipError_t launch_apply5d(const OpArgs<double> &a, int NC, int wpb, uint32_t grid, const double *f, double *r,
                          double *part, const double *skip, hipStream_t s, const ChebStep<double> *cheb) {
    if (a.nP == 0) return hipSuccess;
    const uint32_t per = 16u * (uint32_t)wpb, ntiles = (a.nP + per - 1) / per;
    bool blocks_ok = a.b5.nb >= 3 && a.b5.nb <= 5;
    for (uint32_t k = 0; blocks_ok && k < a.b5.nb; k++)
        blocks_ok = a.b5.w[k] >= 1 && a.b5.w[k] <= 16 && a.b5.h0[k] + a.b5.w[k] <= a.slm;
    if ((a.W != 15 && a.W != 16) || a.WB > TR_WBMAX || !a.crec || !a.tab || (a.nB && !a.bcrec) || wpb < 1 ||
        wpb > 8 || !mfma5_supported(a.slm) || !blocks_ok || (a.cls && !a.hpo) || (!a.cls && a.hpo) ||
        (a.cls && (a.cls == f || a.cls == r)))
        return hipErrorInvalidValue;
    if (cheb && (!cheb->rhs || !cheb->dinv)) return hipErrorInvalidValue;
    const void *k = apply5d_kernel_epi(cheb ? (part ? 3 : 2) : (part ? 1 : 0), NC, a.W, a.b5.nb);
    if (!k) return hipErrorInvalidValue;
    const double *fg = f;
    if (a.cls) {
        hipError_t e = launch_perm_harm(f, a.cls, a.hpo, a.slm, (size_t)a.nP * a.slm, false, skip, s);
        if (e != hipSuccess) return e;
        fg = a.cls;
    }
    dim3 g(grid ? std::min(grid, ntiles) : ntiles), b(TR_WAVE * wpb);
    OpArgs<double> aa = a;
    ChebStep<double> cs = cheb ? *cheb : ChebStep<double>{nullptr, nullptr, 0.0, 0.0, 0.0};
    void *args[] = {&aa, &fg, &f, &r, &part, &skip, &cs};
    return hipLaunchKernel(k, g, b, args, apply5d_lds_bytes(a.b5), s);
}

bool mfma5_supported(uint32_t slm) { return slm > 32 && slm <= 80; }

// LDS of variant 5: the stored volume B blocks (hi, and lo with nsplitB 2)
size_t apply5_lds_bytes(const Blk5 &b5) { return (size_t)b5.nvol * 512 * b5.nsplitB; }

template <int EPI>
static const void *apply5_kernel(int NC, uint32_t W, uint32_t NT) {
#define TR_K5(NCC, WW, NTT) reinterpret_cast<const void *>(&k_apply5<NCC, WW, NTT, EPI>)
#define TR_K5T(NCC, WW) (NT == 3 ? TR_K5(NCC, WW, 3) : (NT == 4 ? TR_K5(NCC, WW, 4) : TR_K5(NCC, WW, 5)))
#define TR_K5W(NCC) (W == 15 ? TR_K5T(NCC, 15) : TR_K5T(NCC, 16))
    if ((W != 15 && W != 16) || (NC != 7 && NC != 10) || NT < 3 || NT > 5) return nullptr;
    return NC == 7 ? TR_K5W(7) : TR_K5W(10);
#undef TR_K5W
#undef TR_K5T
#undef TR_K5
}

static const void *apply5_kernel_epi(int epi, int NC, uint32_t W, uint32_t NT, uint32_t stat = 0, int ns = 3) {
    if (stat) return apply5s_kernel_epi(epi, NC, W, stat, ns);
    switch (epi) {
        case 1: return apply5_kernel<1>(NC, W, NT);
        case 2: return apply5_kernel<2>(NC, W, NT);
        case 3: return apply5_kernel<3>(NC, W, NT);
        default: return apply5_kernel<0>(NC, W, NT);
    }
}

uint32_t apply5_grid(const OpArgs<float> &a, int NC, int wpb, int epi) {
    const uint32_t per = 16u * (uint32_t)wpb, ntiles = (a.nP + per - 1) / per;
    const void *k = apply5_kernel_epi(epi, NC, a.W, a.b5.nb, a.b5.stat, a.nsplit);
    if (!k || ntiles == 0) return ntiles;
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, TR_WAVE * wpb, apply5_lds_bytes(a.b5)) != hipSuccess ||
        nb < 1 || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        return ntiles;
    }
    uint32_t g = (uint32_t)nb * (uint32_t)cus;
    if (g >= 8) g -= g % 8;
    return std::min(g, ntiles);
}

hipError_t launch_apply5(const OpArgs<float> &a, int NC, int wpb, uint32_t grid, const float *f, float *r,
                         double *part, const double *skip, hipStream_t s, const ChebStep<float> *cheb) {
    if (a.nP == 0) return hipSuccess;
    const uint32_t per = 16u * (uint32_t)wpb, ntiles = (a.nP + per - 1) / per;
    // blocks: 3..5, widths <= 16 inside slm; the lo parts of B for nsplit 3
    bool blocks_ok = a.b5.nb >= 3 && a.b5.nb <= 5;
    for (uint32_t k = 0; blocks_ok && k < a.b5.nb; k++)
        blocks_ok = a.b5.w[k] >= 1 && a.b5.w[k] <= 16 && a.b5.h0[k] + a.b5.w[k] <= a.slm;
    if ((a.W != 15 && a.W != 16) || a.WB > TR_WBMAX || !a.crec || !a.tab || (a.nB && !a.bcrec) || wpb < 1 ||
        wpb > 8 || !mfma5_supported(a.slm) || a.nsplit < 1 || a.nsplit > 4 || !blocks_ok ||
        (a.nsplit >= 3 && a.b5.nsplitB != 2) || ((a.nsplit == 4) != (a.b5.f32 != 0)) ||
        (a.b5.f32 && !a.b5.stat))
        return hipErrorInvalidValue;
    if (cheb && (!cheb->rhs || !cheb->dinv)) return hipErrorInvalidValue;
    // the static layout: its blocks, slm and both B parts as the kernel assumes
    if (a.b5.stat) {
        const uint32_t sl = a.b5.stat;
        bool ok = (sl == 4 || sl == 5) && a.slm == cl_slm((int)sl) && a.b5.nb == (uint32_t)cl_nblk((int)sl) &&
                  a.b5.nsplitB == 2 && a.b5.nvol == (uint32_t)cl_vol_count((int)sl, NC) &&
                  a.b5.nbnd == (uint32_t)cl_bnd_count((int)sl);
        for (uint32_t k = 0; ok && k < a.b5.nb; k++)
            ok = a.b5.h0[k] == (uint32_t)cl_h0((int)sl, (int)k) && a.b5.w[k] == (uint32_t)cl_w((int)sl, (int)k);
        if (!ok) return hipErrorInvalidValue;
    }
    const void *k = apply5_kernel_epi(cheb ? (part ? 3 : 2) : (part ? 1 : 0), NC, a.W, a.b5.nb, a.b5.stat, a.nsplit);
    if (!k) return hipErrorInvalidValue;
    dim3 g(grid ? std::min(grid, ntiles) : ntiles), b(TR_WAVE * wpb);
    OpArgs<float> aa = a;
    ChebStep<float> cs = cheb ? *cheb : ChebStep<float>{nullptr, nullptr, 0.0f, 0.0f, 0.0f};
    void *args[] = {&aa, &f, &r, &part, &skip, &cs};
    return hipLaunchKernel(k, g, b, args, apply5_lds_bytes(a.b5), s);
}

int mfma_J(uint32_t slm, int elem_bytes) {
    const int E = 16 / elem_bytes;
    return E * (int)((slm + 4 * E - 1) / (4 * E));
}

bool mfma_supported(uint32_t slm, int mode) {
    const int J = mfma_J(slm, mode == 0 ? 8 : 4);
    // the kernels take ceil(slm / 16) output tiles to be ceil(4J / 16); the
    // FP64 contraction at J = 8 takes slm <= 28 (mfma_steps)
    return (J == 2 || J == 4 || J == 8) && (slm + 15) / 16 == (uint32_t)(4 * J + 15) / 16 &&
           !(mode == 0 && J == 8 && slm > 28);
}

// bytes of the B tables (volume and boundary slices) staged in LDS
size_t apply4_lds_bytes(int mode, int NC, uint32_t slm, int nsplit) {
    const int J = mfma_J(slm, mode == 0 ? 8 : 4), S = mfma_steps(mode, J), BW = mode == 2 ? 8 : (mode == 0 ? 8 : 4);
    const int NT = (int)((slm + 15) / 16);
    return (size_t)(NC + 3) * S * NT * 64 * BW * (mode == 2 && nsplit > 1 ? 2 : 1);
}

// The kernel instantiation for (J, NC, W) with epilogue D, or null.
template <typename T, int MODE, int D>
static const void *apply4_kernel(int J, int NC, uint32_t W) {
#define TR_K4(JJ, NCC, WW) reinterpret_cast<const void *>(&k_apply4<T, MODE, JJ, NCC, WW, D>)
#define TR_K4W(JJ, NCC) (W == 15 ? TR_K4(JJ, NCC, 15) : TR_K4(JJ, NCC, 16))
#define TR_K4N(JJ) (NC == 7 ? TR_K4W(JJ, 7) : TR_K4W(JJ, 10))
    if ((W != 15 && W != 16) || (NC != 7 && NC != 10)) return nullptr;
    if constexpr (sizeof(T) == 8)
        if (J == 2) return TR_K4N(2);
    if (J == 4) return TR_K4N(4);
    if (J == 8) return TR_K4N(8);
    return nullptr;
#undef TR_K4N
#undef TR_K4W
#undef TR_K4
}

// Persistent grid of the variant-4 operator (the kernel without or with the
// fused dot product): the workgroups resident on the device at once
// (occupancy x CUs), rounded down to a multiple of the 8 XCDs, at most one
// per tile.
template <typename T, int MODE>
static const void *apply4_kernel_epi(int epi, uint32_t slm, int NC, uint32_t W) {
    const int J = mfma_J(slm, (int)sizeof(T));
    switch (epi) {
        case 1: return apply4_kernel<T, MODE, 1>(J, NC, W);
        case 2: return apply4_kernel<T, MODE, 2>(J, NC, W);
        case 3: return apply4_kernel<T, MODE, 3>(J, NC, W);
        default: return apply4_kernel<T, MODE, 0>(J, NC, W);
    }
}

template <typename T, int MODE>
uint32_t apply4_grid(const OpArgs<T> &a, int NC, int wpb, int epi) {
    const uint32_t per = 16u * (uint32_t)wpb, ntiles = (a.nP + per - 1) / per;
    const void *k = apply4_kernel_epi<T, MODE>(epi, a.slm, NC, a.W);
    if (!k || ntiles == 0) return ntiles;
    const size_t sh = apply4_lds_bytes(MODE, NC, a.slm, a.nsplit);
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, TR_WAVE * wpb, sh) != hipSuccess || nb < 1 ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        return ntiles;
    }
    uint32_t g = (uint32_t)nb * (uint32_t)cus;
    if (g >= 8) g -= g % 8;
    return std::min(g, ntiles);
}

template <typename T, int MODE>
hipError_t launch_apply4(const OpArgs<T> &a, int NC, int wpb, uint32_t grid, const T *f, T *r, double *part,
                         const double *skip, hipStream_t s, const ChebStep<T> *cheb) {
    if (a.nP == 0) return hipSuccess;
    const uint32_t per = 16u * (uint32_t)wpb, ntiles = (a.nP + per - 1) / per;
    // the unrolled slot loops read exactly W = 15 or 16 volume slots and at
    // most TR_WBMAX boundary slots; coefficient records and B slices must be
    // present
    if ((a.W != 15 && a.W != 16) || a.WB > TR_WBMAX || !a.crec || !a.tab || (a.nB && !a.bcrec) || wpb < 1 ||
        wpb > 8)
        return hipErrorInvalidValue;
    dim3 g(grid ? std::min(grid, ntiles) : ntiles), b(TR_WAVE * wpb);
    const size_t sh = apply4_lds_bytes(MODE, NC, a.slm, a.nsplit);
    if (cheb && (!cheb->rhs || !cheb->dinv)) return hipErrorInvalidValue;
    const void *k = apply4_kernel_epi<T, MODE>(cheb ? (part ? 3 : 2) : (part ? 1 : 0), a.slm, NC, a.W);
    if (!k) return hipErrorInvalidValue;
    OpArgs<T> aa = a;
    ChebStep<T> cs = cheb ? *cheb : ChebStep<T>{nullptr, nullptr, (T)0, (T)0, (T)0};
    void *args[] = {&aa, &f, &r, &part, &skip, &cs};
    return hipLaunchKernel(k, g, b, args, sh, s);
}

template hipError_t launch_apply4<double, 0>(const OpArgs<double> &, int, int, uint32_t, const double *, double *,
                                             double *, const double *, hipStream_t, const ChebStep<double> *);
template hipError_t launch_apply4<float, 1>(const OpArgs<float> &, int, int, uint32_t, const float *, float *,
                                            double *, const double *, hipStream_t, const ChebStep<float> *);
template hipError_t launch_apply4<float, 2>(const OpArgs<float> &, int, int, uint32_t, const float *, float *,
                                            double *, const double *, hipStream_t, const ChebStep<float> *);
template uint32_t apply4_grid<double, 0>(const OpArgs<double> &, int, int, int);
template uint32_t apply4_grid<float, 1>(const OpArgs<float> &, int, int, int);
template uint32_t apply4_grid<float, 2>(const OpArgs<float> &, int, int, int);

}  // namespace tr
