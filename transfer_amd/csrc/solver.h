// solver.h -- device-resident CG (solver.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

#include "kernels.h"

namespace tr {

// slots of the device scalar array
enum {
    SC_RZ = 0,
    SC_PAP = 1,
    SC_ALPHA = 2,
    SC_RR = 3,
    SC_BETA = 4,
    SC_BB = 5,
    SC_THR = 6,
    SC_CONV = 7,  // 0 running, 1 converged, 2 iteration cap, 3 breakdown
                  // (NaN, or p.Ap <= 0), 4 stagnation (see CgParams)
    SC_ITER = 8,
    SC_RZ2 = 9,   // r.z of odd iterations (SC_RZ holds it for even ones)
    SC_RRMIN = 10, // smallest r.r so far
    SC_RRWIN = 11, // SC_RRMIN at the last stagnation check
    SC_RUN = 12,   // 1 when the iteration's update pass ran (its p update
                   // pass follows it), 0 when the solve had stopped
    SC_N = 16
};

// Work space of one CG solve of A x = b in precision T.  All device
// pointers; dinv may be null (no preconditioner).
template <typename T>
struct CgSpace {
    OpArgs<T> a;
    int NC;
    LaneMap map;
    size_t n;          // nP * slm
    T *x, *r, *p, *Ap;
    const T *b, *dinv;
    double *part;      // >= max(apply_blocks, 3*1024) doubles
    double *part2;     // >= 2*1024 doubles
    double *sc;        // SC_N doubles
    // general preconditioner z = M r (launches on the solve's stream, no
    // host synchronisation: it is captured into the iteration graph; its
    // kernels return at once when *skip != 0); when set, dinv is ignored
    // and z (n entries) holds M r
    std::function<hipError_t(const T *r, T *z, const double *skip)> prec;
    T *z = nullptr;
    // optional split of a polynomial preconditioner inside the iteration
    // loop: the update pass writes pre_out = pre_scale * pre_dinv * r (its
    // first iterate), prec_loop(r, z, skip, rzpart) does the rest and, when
    // prec_nparts > 0, writes that many r.z partials to rzpart itself
    std::function<hipError_t(const T *r, T *z, const double *skip, double *rzpart)> prec_loop;
    T *pre_out = nullptr;
    const T *pre_dinv = nullptr;
    T pre_scale = 0;
    int prec_nparts = 0;
};

struct CgParams {
    int maxit = 1000;
    double tol_abs = 0, tol_rel = 1e-8;
    int chunk = 16;
    int diag_op = 0;   // 1: iterate with the angular-diagonal operator
    // stagnation test (0 = off): every stall_window iterations the smallest
    // r.r so far must have dropped below 0.81 of its value at the previous
    // check (||r|| by 10%), else the solve stops with flag 4
    int stall_window = 0;
    // 1: continue a stopped solve: skip r = b - A x and p = z; the caller
    // has set r, p and the scalars (cg_replace_residual)
    int resume = 0;
    // optional: the iteration graph kept across calls on the same space and
    // parameters (captured on the first call; the caller destroys it with
    // hipGraphExecDestroy); null: captured and destroyed per call
    hipGraphExec_t *graph = nullptr;
};

struct CgResult {
    int iters = 0, flag = 0;
    double rr = 0, bb = 0, seconds = 0;
    double rz_dir = 0;  // r.z that built the last direction (0 if none)
};

// Solves from the x held on entry.
template <typename T>
int pcg_device(const CgSpace<T> &S, const CgParams &prm, hipStream_t st, CgResult &res);

// A refinement step continues the previous step's CG (residual
// replacement) only while the true residual, in the stopped solve's scale,
// is within this factor of the recursive one; else it restarts (measured:
// replacements at ratios up to 1.004 saved 5-12% of the inner iterations,
// at 1.05 to 1.9 they cost up to 60%).
constexpr double kReplaceDrift = 1.02;

// Residual replacement (a "reliable update", Sleijpen and van der Vorst
// 1996): a solve stopped converged (flag 1) with its last direction p_k in
// S.p and rz_dir = r_k.z_k; the caller has overwritten S.r with a new
// residual whose units are 1/s of the stopped solve's (a vector q of that
// solve is s q in the new units) and reset S.x.  The direction becomes
// p = z + beta s p_k' with beta = r.z / (s^2 rz_dir) and p_k' = p_k less
// its component along the new r (the recursive residual is orthogonal to
// p_k), and the scalars are set for a solve with prm.resume = 1 to
// tolerance prm.tol_rel relative to |r|.  ok = false (nothing changed but
// z) when r.r, r.z, rz_dir or s is not positive and finite.
template <typename T>
int cg_replace_residual(const CgSpace<T> &S, const CgParams &prm, double s, double rz_dir, hipStream_t st,
                        bool &ok);

// dst = (float)(src * s);  x += s * (double)d; hp (with slm): the FP32
// vectors' harmonic order, position k of a vertex's FP32 row holding
// harmonic hp[k] of its FP64 row (null: the same order)
hipError_t launch_to_f32_scaled(const double *src, float *dst, double s, size_t n, hipStream_t st,
                                const uint16_t *hp = nullptr, uint32_t slm = 0);
hipError_t launch_add_f32_scaled(double *x, const float *d, double s, size_t n, hipStream_t st,
                                 const uint16_t *hp = nullptr, uint32_t slm = 0);

}  // namespace tr
