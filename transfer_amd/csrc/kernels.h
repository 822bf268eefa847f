// kernels.h -- launch interface of the gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "transfer.h"

// most ELL slots the batched schedules (variants 2-4) unroll
#define TR_WMAX 16
// most boundary-stencil slots the matrix-core schedule (variant 4) unrolls
#define TR_WBMAX 12

namespace tr {

// Lane mapping of the vertex kernels: a wavefront of 64 lanes holds VPW
// vertices, L = slm / J lanes each, every lane owning J consecutive
// harmonics.  WPB wavefronts per workgroup.
struct LaneMap {
    int J = 1, L = 1, VPW = 1, WPB = 4;
    int variant = 3;  // 1: stage every vertex of a wavefront at once; 2: batched (k_apply2);
                      // 3: batched with buffer loads and coefficient records (k_apply3);
                      // 4: angular contraction on the matrix cores (k_apply4);
                      // 5: the same in blocks of 16 harmonics, FP32/BF16 (k_apply5)
    int mfma_mode = 1;  // FP32 operator under variant 4: 1 = FP32 MFMA, 2 = BF16 MFMA
    uint32_t grid = 0;      // variant 4: workgroups of the persistent grid (0 = one per tile)
    uint32_t grid_dot = 0;  // the same for the launch with the fused dot product
    uint32_t grid_cheb = 0; // the same for the launch with a fused Chebyshev step
};

// Harmonic blocks of the K-blocked matrix-core schedule (variant 5): block
// kb covers the w[kb] harmonics from h0[kb] of the FP32 vectors' order,
// and B block (kb, c, t) (input block kb, channel c, output block t) is
// stored only when bit c*nb + t of vmask[kb] is set, at position voff[kb]
// + (the set bits below it); the boundary channels likewise with bmask and
// boff.  nvol / nbnd: stored volume / boundary blocks; nsplitB 2 when each
// has a BF16 lo part (after all hi parts of its section).
struct Blk5 {
    uint32_t nb = 0, nvol = 0, nbnd = 0, nsplitB = 1;
    uint32_t h0[8] = {}, w[8] = {};
    uint64_t vmask[8] = {}, bmask[8] = {};
    uint32_t voff[8] = {}, boff[8] = {};
    // static class layout (cl_* below) whose stored order the blocks follow,
    // 0 = none (vmask / voff describe them)
    uint32_t stat = 0;
    // 1: FP32 blocks in the operand layout of v_mfma_f32_16x16x4f32 (static
    // layouts only; 1 KB each, the bytes of both BF16 parts: nsplitB 2)
    uint32_t f32 = 0;
};

// Static parity-class layouts of the K-blocked schedule.  The harmonics of
// the even P_K expansion in the four reflection-parity classes (cos with m
// even, cos with m odd, sin with |m| even, sin with |m| odd), each class in
// blocks of at most 16: layout 5 is K = 5 (66 harmonics: 21, 15, 15, 15 ->
// blocks 16, 5 | 15 | 15 | 15), layout 4 is K = 4 (45 harmonics: 15, 10,
// 10, 10, one block per class).
__host__ __device__ constexpr int cl_nblk(int sl) { return sl; }
__host__ __device__ constexpr uint32_t cl_slm(int sl) { return sl == 5 ? 66u : 45u; }
__host__ __device__ constexpr int cl_h0(int sl, int b) {
    return sl == 5 ? (b == 0 ? 0 : (b == 1 ? 16 : 21 + 15 * (b - 2))) : (b == 0 ? 0 : 15 + 10 * (b - 1));
}
__host__ __device__ constexpr int cl_w(int sl, int b) {
    return sl == 5 ? (b == 0 ? 16 : (b == 1 ? 5 : 15)) : (b == 0 ? 15 : 10);
}
__host__ __device__ constexpr int cl_class(int sl, int b) { return sl == 5 ? (b < 2 ? 0 : b - 1) : b; }
// The class that volume channel c maps class k onto: the mass channel and
// K^xx, K^yy, K^zz keep it; the xy channels (symmetric part 4, antisymmetric
// 7) swap 0 <-> 2 and 1 <-> 3, xz (5, 8) 0 <-> 1 and 2 <-> 3, yz (6, 9)
// 0 <-> 3 and 1 <-> 2 (the mirror parities of Omega_a Omega_b).
__host__ __device__ constexpr int cl_chan_map(int c, int k) {
    return c < 4 ? k : ((c - 4) % 3 == 0 ? (k ^ 2) : ((c - 4) % 3 == 1 ? (k ^ 1) : 3 - k));
}
// Whether output block t can receive input block kb through volume channel
// c (the mass channel is diagonal in the harmonics), or through any of the
// three boundary channels (|Omega_n| keeps every class).
__host__ __device__ constexpr bool cl_vol_pair(int sl, int kb, int c, int t) {
    return c == 0 ? t == kb : cl_class(sl, t) == cl_chan_map(c, cl_class(sl, kb));
}
__host__ __device__ constexpr bool cl_bnd_pair(int sl, int kb, int t) { return cl_class(sl, t) == cl_class(sl, kb); }
// Stored position of block (kb, c, t) among the pairs in (kb, c, t) order,
// and the number of pairs; boundary blocks likewise over 3 channels.
__host__ __device__ constexpr int cl_vol_index(int sl, int nc, int kb, int c, int t) {
    int n = 0;
    for (int k = 0; k < cl_nblk(sl); k++)
        for (int cc = 0; cc < nc; cc++)
            for (int tt = 0; tt < cl_nblk(sl); tt++) {
                if (k == kb && cc == c && tt == t) return n;
                if (cl_vol_pair(sl, k, cc, tt)) n++;
            }
    return n;
}
__host__ __device__ constexpr int cl_vol_count(int sl, int nc) { return cl_vol_index(sl, nc, cl_nblk(sl), 0, 0); }
__host__ __device__ constexpr int cl_bnd_index(int sl, int kb, int c, int t) {
    int n = 0;
    for (int k = 0; k < cl_nblk(sl); k++)
        for (int cc = 0; cc < 3; cc++)
            for (int tt = 0; tt < cl_nblk(sl); tt++) {
                if (k == kb && cc == c && tt == t) return n;
                if (cl_bnd_pair(sl, k, tt)) n++;
            }
    return n;
}
__host__ __device__ constexpr int cl_bnd_count(int sl) { return cl_bnd_index(sl, cl_nblk(sl), 0, 0); }

// Device operator data (see stencil.h), precision T.
template <typename T>
struct OpArgs {
    uint32_t nP, slm, W, arow, nB, WB, brow;
    int L, VPW;
    const uint32_t *nbr;   // [W][nP]
    const T *coef;         // [W][NC][nP]
    const uint16_t *aidx;  // [arow][slm]
    const T *aval;         // [arow][slm]
    const T *massfac;      // [slm]
    const T *ddiag;        // [3][slm]
    const int32_t *bof;    // [nP] boundary-stencil row of each vertex, -1 inside
    const uint32_t *bnbr;  // [WB][nB]
    const T *bcoef;        // [WB][3][nB]
    const uint16_t *bidx;  // [brow][slm]
    const T *bval;         // [brow][slm]
    const T *bdiag;        // [3][slm]
    // record layout for the buffer-load schedule (variant 3), may be null
    const T *crec;         // [W][nP][RS] coefficients of slot s of vertex v, RS = rec_stride<T>(NC)
    const void *tab;       // variant 3: [(arow+brow)][slm] {T value; uint32 byte offset};
                           // variant 4: MFMA B slices (volume, then boundary)
    int nsplit;            // variant 4, BF16: 1 = plain, 3 = hi/lo split
    const T *bcrec;        // variant 4: [WB][nB][4] boundary coefficient records (3 channels + pad)
    const uint32_t *mask;  // specular reflection: entries of r that vanish, may be null
    uint32_t nmask;
    const T *mf0;          // [nP] per-vertex factor of the lm = 0 mass term (batches), may be null
    const uint32_t *tiles; // variant 4: [ceil(nP/16)] a permutation of the 16-vertex wavefront
                           // tiles, the order of the walk; null = natural order
    // vacuum faces off the axes with their exact angular matrices (see
    // stencil.h), applied by k_bgen after the operator kernel; ngv = 0: none
    uint32_t ngv = 0;
    const uint32_t *gvert = nullptr, *gstart = nullptr, *gface = nullptr, *gpos = nullptr, *gfv = nullptr;
    const T *garea = nullptr, *gR = nullptr, *gOd = nullptr, *gOz = nullptr;
    Blk5 b5;               // variant 5 only
    // FP64 variant 5 on vectors in the internal order: cls is a scratch
    // vector for the input in the class order and hpo[k] the internal
    // harmonic of class position k; both null when the vectors are in the
    // class order already
    const uint16_t *hpo = nullptr;
    T *cls = nullptr;
    // batches: vprob[v] = problem of vertex v and pstop[p * pstride] != 0
    // when problem p has stopped; the buffer-load VALU schedule (variant 3)
    // skips workgroups whose vertices all belong to stopped problems (their
    // outputs keep their old values).  Null: every workgroup runs.
    const uint32_t *vprob = nullptr;
    const double *pstop = nullptr;
    uint32_t pstride = 0;
};

// One step of the Chebyshev iteration on D^-1 A z = D^-1 rhs, written as
// out = zc z - om out_prev + ga dinv (rhs - A z), out_prev the iterate
// before z (held in out on entry; not read when om == 0).
template <typename T>
struct ChebStep {
    const T *rhs, *dinv;
    T zc, om, ga;
};

// coefficient record stride (elements): NC rounded up to a 16-byte multiple
template <typename T>
constexpr int rec_stride(int NC) {
    return (NC * (int)sizeof(T) + 15) / 16 * (16 / (int)sizeof(T));
}
// bytes of one contraction-table record: {T value, uint32 offset} padded to 8 or 16
template <typename T>
constexpr int tab_rec_bytes() {
    return sizeof(T) == 8 ? 16 : 8;
}

// r = A f (volume and vacuum boundary in one pass).  With part != null each
// workgroup also writes sum_v f_v . r_v (FP64) to part[blockIdx]; with
// skip != null the launch does nothing when *skip != 0 (solver loops).
template <typename T>
hipError_t launch_apply(const OpArgs<T> &a, int NC, const LaneMap &m, const T *f, T *r,
                        double *part, const double *skip, hipStream_t s);
// One Chebyshev step (ChebStep): fused into the operator's write on the
// matrix-core schedules and the buffer-load VALU schedule (variant 3), else
// the operator into tmp and an elementwise pass.  Faces off the axes add
// their share afterwards (k_bgen), so with them part must be null.
// With part != null (matrix-core schedules only) each workgroup of the
// persistent grid (cheb_parts of them) also writes sum rhs . out (FP64).
template <typename T>
hipError_t launch_cheb_step(const OpArgs<T> &a, int NC, const LaneMap &m, const T *z, T *out,
                            const ChebStep<T> &cs, T *tmp, const double *skip, hipStream_t s,
                            double *part = nullptr);
// partials written by launch_cheb_step with part, 0 where it cannot fuse them
uint32_t cheb_parts(uint32_t nP, const LaneMap &m);
// r[mask[i]] = 0 for the entries that specular reflection removes (no-op
// without reflecting planes)
template <typename T>
hipError_t launch_zero_mask(const OpArgs<T> &a, T *r, const double *skip, hipStream_t s);
// r = D f, the angular-diagonal operator (kernels.cu:290-357, 440-493)
template <typename T>
hipError_t launch_apply_diag(const OpArgs<T> &a, int NC, const LaneMap &m, const T *f, T *r,
                             double *part, const double *skip, hipStream_t s);
size_t apply_lds_bytes(const LaneMap &m, int NC, uint32_t slm, size_t elem);
size_t apply2_lds_bytes(const LaneMap &m, int NC, uint32_t slm, uint32_t arow, uint32_t brow,
                        size_t elem);
size_t apply3_lds_bytes(const LaneMap &m, int NC, uint32_t slm, uint32_t arow, uint32_t brow,
                        size_t elem);
// workgroups (= dot-product partials) of an operator launch with the fused
// dot product
uint32_t apply_blocks(uint32_t nP, const LaneMap &m);
// vertex tiles (VPW x WPB vertices) of a lane map: the grid of the
// non-persistent kernels (k_apply_diag) and an upper bound of apply_blocks
uint32_t apply_tiles(uint32_t nP, const LaneMap &m);

// variant 4 (kernels_mfma.hip): J = ceil(slm/4) harmonics per lane
// Variant 4: the four chunk lanes q = 0..3 of a vertex each hold J harmonics
// of its row; register j of lane q holds harmonic mfma_harm(E, q, j), E the
// elements of a 16-byte piece (2 FP64, 4 FP32): the lane's pieces are the
// row's pieces 4k + q, so a vertex's four lanes read 64 contiguous bytes
// per load.  J = E * ceil(slm / 4E).
__host__ __device__ constexpr uint32_t mfma_harm(int E, int q, int j) {
    return (uint32_t)(4 * E * (j / E) + E * q + j % E);
}
// K steps of the contraction (mode 0 FP64, 1 FP32, 2 BF16): one register
// per step, four for BF16; FP64 at J = 8 (slm <= 28) folds registers 6 and 7
// of chunks 0 and 1 (harmonics 24..27; the rest are past slm) into one step.
__host__ __device__ constexpr int mfma_steps(int mode, int J) {
    return mode == 2 ? (J + 3) / 4 : (mode == 0 && J == 8 ? 7 : J);
}
// harmonic of K index q in step st (modes 0 and 1)
__host__ __device__ constexpr uint32_t mfma_kharm(int mode, int J, int q, int st) {
    return (mode == 0 && J == 8 && st == 6) ? (uint32_t)(24 + q) : mfma_harm(mode == 0 ? 2 : 4, q, st);
}
int mfma_J(uint32_t slm, int elem_bytes);
// mode 0 FP64, 1 FP32, 2 BF16
bool mfma_supported(uint32_t slm, int mode);
size_t apply4_lds_bytes(int mode, int NC, uint32_t slm, int nsplit);
template <typename T, int MODE>
hipError_t launch_apply4(const OpArgs<T> &a, int NC, int wpb, uint32_t grid, const T *f, T *r, double *part,
                         const double *skip, hipStream_t s, const ChebStep<T> *cheb = nullptr);
// variant 5 (kernels_mfma.hip): FP32 operator for 33..80 harmonics, BF16
// matrix-core projection in blocks of 16 harmonics; tab = B blocks
// [block][channel][tile][lane] (volume, then boundary), nsplit 1 or 2
bool mfma5_supported(uint32_t slm);
// the static class-layout kernel of variant 5 (kernels_mfma5s.hip) for
// layout SL (Blk5::stat), BF16 split ns (1..3) and epilogue epi (0..3):
// 15 volume slots; null otherwise
const void *apply5s_kernel_epi(int epi, int NC, uint32_t W, uint32_t SL, int ns);
size_t apply5_lds_bytes(const Blk5 &b5);
uint32_t apply5_grid(const OpArgs<float> &a, int NC, int wpb, int epi);
hipError_t launch_apply5(const OpArgs<float> &a, int NC, int wpb, uint32_t grid, const float *f, float *r,
                         double *part, const double *skip, hipStream_t s, const ChebStep<float> *cheb = nullptr);
// variant 5 in FP64: the same blocks in the parity classes, contracted on
// v_mfma_f64_16x16x4f64; tab = FP64 B blocks (2 KB each, volume then
// boundary); a.cls / a.hpo as in OpArgs
size_t apply5d_lds_bytes(const Blk5 &b5);
uint32_t apply5d_grid(const OpArgs<double> &a, int NC, int wpb, int epi);
hipError_t launch_apply5d(const OpArgs<double> &a, int NC, int wpb, uint32_t grid, const double *f, double *r,
                          double *part, const double *skip, hipStream_t s, const ChebStep<double> *cheb = nullptr);
// dst[v][k] = src[v][hp[k]] over the n = nP slm entries (inverse: dst[v][hp[k]]
// = src[v][k]); nothing when *skip != 0
hipError_t launch_perm_harm(const double *src, double *dst, const uint16_t *hp, uint32_t slm, size_t n,
                            bool inverse, const double *skip, hipStream_t s);
// workgroups of the persistent variant-4 grid for this operator and WPB,
// for the launch with epilogue epi (0 plain, 1 fused dot product, 2
// Chebyshev step)
template <typename T, int MODE>
uint32_t apply4_grid(const OpArgs<T> &a, int NC, int wpb, int epi);

// moments U, T and flux W of a solution (postproc.hip)
hipError_t launch_moments(const double *f, const uint32_t *perm, const double *mval, const uint32_t *mpos,
                          uint32_t nP, uint32_t slm, double *U, double *T, hipStream_t s);
// outward flux density per boundary face (rec: uint4 per face, col: rows of slm)
hipError_t launch_boundary_flux(const double *f, const void *rec, const double *col, uint32_t nF, uint32_t slm,
                                double *F, hipStream_t s);
hipError_t launch_flux(const tr_tetrahedron *mesh, uint32_t nT, uint32_t nP, const double *T, double *W,
                       hipStream_t s);

// b[v][lm] = lm == 0 ? rhs[v] : 0
hipError_t launch_rhs(const double *rhs, uint32_t nP, uint32_t slm, double *b,
                      hipStream_t s);

// elementwise x = wx*x + wy*y over n entries (y may be null when wy == 0)
hipError_t launch_axpby(double *x, double wx, const double *y, double wy, size_t n,
                        hipStream_t s);

// The learned coarse correction z = dinv r + g (g.r)/(gAg) (coarse.hip);
// part holds learned_coarse_partials() doubles, gAg one device double;
// nothing is done when *skip != 0.
template <typename T>
hipError_t launch_learned_coarse(const T *r, T *z, const T *dinv, const T *g, size_t n, double *part,
                                 const double *gAg, const double *skip, hipStream_t s);
int learned_coarse_partials();
// out[0] = sum_i x[i]*y[i] in FP64, deterministic two-pass reduction;
// part needs red_partials() doubles.
int red_partials();
hipError_t launch_dot(const double *x, const double *y, size_t n, double *part,
                      double *out, hipStream_t s);

// layout conversions between the caller's [vertex][aslm] and the internal
// renumbered [vertex][slm]
hipError_t launch_to_internal(const double *src, double *dst, const uint32_t *perm,
                              uint32_t nP, uint32_t slm, uint32_t aslm, hipStream_t s);
hipError_t launch_to_external(const double *src, double *dst, const uint32_t *iperm,
                              uint32_t nP, uint32_t slm, uint32_t aslm, hipStream_t s);

}  // namespace tr
