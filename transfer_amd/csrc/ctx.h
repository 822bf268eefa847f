// ctx.h -- tr_ctx, the device context behind the C ABI (context.hip), and
// the helpers the batch entry points (batch.hip) share with it.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <set>
#include <string>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "stencil.h"

namespace tr {
struct Multilevel;
}

#define HIPCHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return tr::fail(TR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct tr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    tr::HostOperator op;
    tr::LaneMap map;
    // device operator data (FP64)
    uint32_t *d_nbr = nullptr, *d_bnbr = nullptr;
    int32_t *d_bof = nullptr;
    uint32_t *d_perm = nullptr, *d_iperm = nullptr;
    uint16_t *d_aidx = nullptr, *d_bidx = nullptr;
    double *d_coef = nullptr, *d_bcoef = nullptr, *d_aval = nullptr, *d_bval = nullptr;
    double *d_massfac = nullptr, *d_ddiag = nullptr, *d_bdiag = nullptr, *d_rhs = nullptr;
    double *d_part = nullptr, *d_red = nullptr, *d_stage = nullptr;
    double *d_crec = nullptr;   // coefficient records for the buffer-load schedule
    double *d_bcrec = nullptr;  // boundary coefficient records (variant 4)
    uint32_t *d_mask = nullptr; // specular reflection: vanishing entries
    uint32_t *d_tiles = nullptr; // variant 4: order of the 16-vertex wavefront tiles
    double *d_mf0 = nullptr;       // [nP] lm = 0 mass factor per vertex (batches), or null
    float *f_mf0 = nullptr;
    double *d_vsum = nullptr, *d_xyz = nullptr;  // learned-guess feature inputs [3][nP] each, or null
    // vacuum faces off the axes (general normals, see stencil.h)
    uint32_t *d_gvert = nullptr, *d_gstart = nullptr, *d_gface = nullptr, *d_gpos = nullptr, *d_gfv = nullptr;
    double *d_garea = nullptr, *d_gR = nullptr, *d_gOd = nullptr, *d_gOz = nullptr;
    float *f_garea = nullptr, *f_gR = nullptr, *f_gOd = nullptr, *f_gOz = nullptr;
    double *d_feat_tmp = nullptr;  // [2][nP] kappa and optical depth + 8 flags (guess.hip), on first use
    float *d_guess_tmp = nullptr;  // [9][nP] features and prediction (guess.hip), on first use
    double *d_mom_val = nullptr;   // moments table [3][slm]
    uint32_t *d_mom_pos = nullptr;
    void *d_tab = nullptr;      // contraction-table records
    void *d_tab4 = nullptr;     // MFMA B slices (variant 4, FP64)
    // FP32 copy of the operator data (mixed-precision solves), built lazily;
    // index arrays are shared with the FP64 set
    bool have32 = false;
    tr::LaneMap map32;
    float *f_coef = nullptr, *f_bcoef = nullptr, *f_aval = nullptr, *f_bval = nullptr;
    float *f_massfac = nullptr, *f_ddiag = nullptr, *f_bdiag = nullptr;
    float *f_crec = nullptr, *f_bcrec = nullptr;
    void *f_tab = nullptr;
    void *f_tab4 = nullptr, *f_tab4b1 = nullptr, *f_tab4b3 = nullptr;  // variant 4: FP32, BF16, BF16 hi/lo
    void *f_tab5 = nullptr;  // variant 5: BF16 B blocks (33..80 harmonics)
    bool tab5_static = true; // variant 5: the static class layout where it applies (Blk5::stat)
    // variant 5: the harmonic blocks and stored B blocks (kernels.h, Blk5).
    // Under variant 5 the FP32 vectors hold harmonic hp32[k] of the
    // internal order at position k (reflection-parity classes, so that
    // every block lies in one class; empty = the internal order itself);
    // d_hp32 is its device copy.  FP64 vectors never change order.
    tr::Blk5 b5;
    std::vector<uint16_t> hp32;
    uint16_t *d_hp32 = nullptr;
    // FP64 variant 5 (the parity classes on the FP64 matrix cores), built
    // at context creation where it applies (v5d_ok): its blocks, B tables,
    // class order hp64 (device copy d_hp64) and d_cls, the scratch vector
    // that holds an internal-order input in the class order; vmap is the
    // VALU schedule the angular-diagonal operator keeps
    tr::Blk5 b5d;
    void *d_tab5d = nullptr;
    std::vector<uint16_t> hp64;
    uint16_t *d_hp64 = nullptr;
    double *d_cls = nullptr;
    tr::LaneMap vmap;
    int ang_prec = 3;  // FP32 operator's angular contraction: 0 FP32, 1 BF16, 3 BF16 hi/lo split
    std::set<void *> vectors, raw;
    // batches: internal vertex range of problem p = voff[p] .. voff[p+1]
    // (empty for a single problem)
    std::vector<uint32_t> voff;
    uint32_t *d_vprob = nullptr;  // batches: problem of every internal vertex (built on first solve)
    bool own_stream = true;        // false for the levels of a multilevel hierarchy
    tr::Multilevel *mg = nullptr;  // multilevel preconditioner (use_prec 3), built on first use
    uint32_t mg_coarsest = 3000, mg_agg = 16; // hierarchy shape: coarsest-level size, aggregate size
    int mg_nu = 2;                            // damped-Jacobi steps on the coarsest level
    // Chebyshev polynomial preconditioner (use_prec 4): degree, the ratio
    // lambda_max / lambda_min of the interval it targets, and lambda_max of
    // D^-1 A (power iteration on first use, 0 before)
    int cheb_deg = 4;
    double cheb_ratio = 100.0, cheb_lmax = 0.0;
    int cheb_split = 2;  // BF16 split of the polynomial's steps: 0 = the operator's, else 1..3
    // mixed solver: fall back to the FP64 solver when a refinement step
    // cuts the true residual by less than this factor; stagnation window of
    // the inner CG (0 = none)
    double ir_stagnation = 0.5;
    int ir_stall_window = 100;
    // mixed solver: 1 = each refinement step continues the inner CG from
    // its last direction (residual replacement); 0 = restarts it
    int ir_reliable = 1;
    std::vector<void *> owned;

    tr::OpArgs<double> args() const { return args(map); }
    // the FP64 operator on vectors already in its class order (variant 5;
    // else the same as args())
    tr::OpArgs<double> args_cls() const {
        tr::OpArgs<double> a = args(map);
        a.hpo = nullptr;
        a.cls = nullptr;
        return a;
    }
    // the FP64 vectors' harmonic order inside class-order solves (null:
    // the internal order)
    const uint16_t *perm64() const { return map.variant == 5 && !hp64.empty() ? d_hp64 : nullptr; }
    tr::OpArgs<double> args(const tr::LaneMap &map) const {
        tr::OpArgs<double> a;
        a.nP = op.nP;
        a.slm = op.slm;
        a.W = op.W;
        a.arow = op.arow;
        a.nB = op.nB;
        a.WB = op.WB;
        a.brow = op.brow;
        a.L = map.L;
        a.VPW = map.VPW;
        a.nbr = d_nbr;
        a.coef = d_coef;
        a.aidx = d_aidx;
        a.aval = d_aval;
        a.massfac = d_massfac;
        a.ddiag = d_ddiag;
        a.bof = d_bof;
        a.bnbr = d_bnbr;
        a.bcoef = d_bcoef;
        a.bidx = d_bidx;
        a.bval = d_bval;
        a.bdiag = d_bdiag;
        a.crec = d_crec;
        a.tab = map.variant == 4 ? d_tab4 : (map.variant == 5 ? d_tab5d : d_tab);
        a.nsplit = 1;
        if (map.variant == 5) {
            a.b5 = b5d;
            a.hpo = d_hp64;
            a.cls = d_cls;
        }
        a.bcrec = d_bcrec;
        a.mask = d_mask;
        a.nmask = (uint32_t)op.mask.size();
        a.mf0 = d_mf0;
        a.tiles = d_tiles;
        a.ngv = (uint32_t)op.gvert.size();
        a.gvert = d_gvert;
        a.gstart = d_gstart;
        a.gface = d_gface;
        a.gpos = d_gpos;
        a.gfv = d_gfv;
        a.garea = d_garea;
        a.gR = d_gR;
        a.gOd = d_gOd;
        a.gOz = d_gOz;
        return a;
    }
    tr::OpArgs<float> args32() const { return args32(map32); }
    tr::OpArgs<float> args32(const tr::LaneMap &map32) const {
        tr::OpArgs<double> d = args(map);
        tr::OpArgs<float> a;
        a.nP = d.nP;
        a.slm = d.slm;
        a.W = d.W;
        a.arow = d.arow;
        a.nB = d.nB;
        a.WB = d.WB;
        a.brow = d.brow;
        a.L = map32.L;
        a.VPW = map32.VPW;
        a.nbr = d.nbr;
        a.aidx = d.aidx;
        a.bof = d.bof;
        a.bnbr = d.bnbr;
        a.bidx = d.bidx;
        a.coef = f_coef;
        a.aval = f_aval;
        a.massfac = f_massfac;
        a.ddiag = f_ddiag;
        a.bcoef = f_bcoef;
        a.bval = f_bval;
        a.bdiag = f_bdiag;
        a.crec = f_crec;
        a.tab = map32.variant == 5 ? f_tab5
                : map32.variant != 4 ? f_tab
                                     : (map32.mfma_mode == 1 ? f_tab4 : (ang_prec == 1 ? f_tab4b1 : f_tab4b3));
        // variant 5 has B lo parts only when its stored blocks fit the LDS
        // twice: else the A operand split at most; FP32 blocks: 4
        a.nsplit = (map32.variant == 5 && b5.f32) ? 4
                   : ang_prec == 1                ? 1
                                                  : (map32.variant == 5 ? (b5.nsplitB == 2 ? 3 : 2) : 3);
        a.b5 = b5;
        a.bcrec = f_bcrec;
        a.mask = d.mask;
        a.nmask = d.nmask;
        a.mf0 = f_mf0;
        a.tiles = d.tiles;
        a.ngv = d.ngv;
        a.gvert = d.gvert;
        a.gstart = d.gstart;
        a.gface = d.gface;
        a.gpos = d.gpos;
        a.gfv = d.gfv;
        a.garea = f_garea;
        a.gR = f_gR;
        a.gOd = f_gOd;
        a.gOz = f_gOz;
        return a;
    }
    // the FP32 vectors' harmonic order under the current FP32 schedule
    // (null: the internal order)
    const uint16_t *perm32() const { return map32.variant == 5 && !hp32.empty() ? d_hp32 : nullptr; }
    size_t nint() const { return (size_t)op.nP * op.slm; }
    size_t next() const { return (size_t)op.nP * op.aslm; }
};

namespace tr {

// Uploads a host vector into a new device allocation owned by the context.
// The copy is synchronous: callers pass temporaries, which end with the
// call's full expression.
template <typename T>
int upload(tr_ctx *c, T **dst, const std::vector<T> &src) {
    size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
    HIPCHECK(hipMalloc((void **)dst, bytes));
    c->owned.push_back(*dst);
    if (!src.empty()) {
        HIPCHECK(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    }
    return TR_OK;
}

// Sets up the allocation-free dot's scratch on `device` (current): once per
// device (noalloc.hip).
int noalloc_init(int device);
// Second half of context creation: schedules and device data for the
// operator already in c->op (the context is destroyed on failure).
int ctx_init(tr_ctx *c, int device);
// Lazily builds the FP32 operator data of the mixed solver.
int ensure_fp32(tr_ctx *c);
// diag(A) per entry of the internal [v][slm] layout (Jacobi preconditioner)
std::vector<double> operator_diagonal(const HostOperator &op);

// multilevel preconditioner (mg.hip)
int build_multilevel(tr_ctx *c, bool f32, uint32_t min_coarse, uint32_t size);
void destroy_multilevel(Multilevel *mg);
hipError_t multilevel_apply64(tr_ctx *c, const double *r, double *z, const double *skip);
hipError_t multilevel_apply32(tr_ctx *c, const float *r, float *z, const double *skip);
int multilevel_levels(const tr_ctx *c);
// largest eigenvalue of D^-1 A (FP64 power iteration; dinv, x, y device [nint])
int lambda_max(tr_ctx *c, const double *dinv, double *x, double *y, double &lam);
// lambda_max of D^-1 A for the Chebyshev preconditioner (1.1 x the power
// iteration's; dinv = 1 / diag(A) on the host), computed once per context
int cheb_lambda(tr_ctx *c, const std::vector<double> &dinv);

}  // namespace tr
