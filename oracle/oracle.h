/*
 * oracle.h -- CPU FP64 restatement of uranix/Transfer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (transfer_amd/, include/)
 * may include, link or call this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it, and only as the checker.
 *
 * Every function cites the reference file:line it restates.  Layouts are the
 * reference's USE_DOUBLE layouts (REAL = double, idx = uint32_t).
 */
#ifndef TRANSFER_ORACLE_H
#define TRANSFER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint32_t or_idx;

/* tetrahedron.h:7-20 (USE_DOUBLE: 128 bytes) and face.h:8-18 (64 bytes) */
#pragma pack(push, 1)
typedef struct or_tet {
    or_idx p[4];
    double kappa_volume;
    double I_p;
    double s[4][3];
} or_tet;
typedef struct or_face {
    or_idx p[4];
    double s[3];
    double _padd[3];
} or_face;
#pragma pack(pop)

/* ---- angular data (AngularData.cpp:9-68, Spherical.cpp:39-82) ---- */
int or_slm(int maxk);
/* Host (unpadded) arrays: omega/omega_pos [3][slm][slm], Ox/Oy/Oz [slm][slm].
 * albedo/g/hg_lmax: scattering extension (albedo = 0 reproduces the reference).
 * massfac[slm] = 1 - albedo * g_l (g_l = g^l for l <= hg_lmax, else 0;
 * hg_lmax = 0 means isotropic scattering). */
int or_angular(int maxk, double albedo, double g, int hg_lmax,
               double *omega, or_idx *omega_pos, double *Ox, double *Oy,
               double *Oz, double *massfac);
/* Raw exact moments, for tests:
 * <Y_a Y_b> (a,b over even+odd harmonics up to degree lmax, index l*l+l+m) */
int or_sh_gram(int lmax, double *gram);
/* Y_lm at a point (same convention as Spherical.cpp:62-82), all l<=lmax */
int or_sh_eval(int lmax, double x, double y, double z, double *out);

/* DeviceAngularData.cpp:7-67 padding: aslm = align(slm, 64/sizeof(double)) */
int or_aslm(int slm);
void or_pad_angular(int slm, const double *omega, const or_idx *omega_pos,
                    const double *Ox, const double *Oy, const double *Oz,
                    double *omega_d, or_idx *omega_pos_d, double *Ox_d,
                    double *Oy_d, double *Oz_d);

/* ---- the operator (kernels.cu), all vectors [nP][aslm] ---- */
typedef struct or_problem {
    or_idx nP, nT, nF;
    const or_idx *tetstart, *tetidx, *tetpos;
    const or_tet *mesh;
    const or_idx *facestart, *faceidx, *facepos;
    const or_face *bnd;
    or_idx slm, aslm;
    const double *omega;     /* [3][aslm][aslm] padded */
    const or_idx *omega_pos; /* [3][aslm][aslm] padded */
    const double *Ox, *Oy, *Oz; /* [aslm][aslm] padded */
    const double *massfac;   /* [slm] extension, NULL = 1 */
    int nthreads;            /* OpenMP threads, <=0: default */
    const unsigned char *face_kind; /* [nF] extension: 0 vacuum (the reference's
                                       only kind), 1 specular reflection, which
                                       is the natural condition of the
                                       even-parity form (no boundary term);
                                       NULL = all vacuum */
} or_problem;

void or_volume_part(const or_problem *pr, const double *f, double *r);
void or_surface_part(const or_problem *pr, const double *f, double *r);
void or_volume_part_diag(const or_problem *pr, const double *f, double *r);
void or_surface_part_diag(const or_problem *pr, const double *f, double *r);
void or_compute_lhs(const or_problem *pr, const double *f, double *r);
void or_compute_lhs_diag(const or_problem *pr, const double *f, double *r);
void or_compute_rhs(const or_problem *pr, double *b);

/* BLAS-1 (kernels.cu:12-50) over all nP*aslm entries, padding included */
void or_scale(const or_problem *pr, double *x, double wx);
void or_add_prod(const or_problem *pr, double *x, const double *y, double wy);
void or_mul_add(const or_problem *pr, double *x, double wx, const double *y);
void or_mul_add_prod(const or_problem *pr, double *x, double wx,
                     const double *y, double wy);
/* reductions (kernels.cu:55-115, CudaContext.cpp:175-198): padding excluded */
double or_norm(const or_problem *pr, const double *x);
double or_dot(const or_problem *pr, const double *x, const double *y);

/* ---- PCG (main.cpp:125-162) with invprec (main.cpp:223-248) ---- */
typedef struct or_pcg_params {
    int use_prec;        /* cfg precondition */
    int maxit;           /* main.cpp:142 uses 1000 */
    double tol_abs;      /* main.cpp:149 uses 1e-4 (absolute ||r||) */
    double tol_rel;      /* extension: stop when ||r|| <= tol_rel*||b|| (0 = off) */
    int inner_maxit;     /* main.cpp:232 uses 1000 */
    double inner_tol;    /* main.cpp:239 uses 1e-8 (absolute) */
    int record;          /* number of leading (alpha, beta, ||r||) triples to record */
} or_pcg_params;
typedef struct or_pcg_stats {
    int iters;           /* k at exit (main.cpp:142-160) */
    int preck;           /* total inner iterations (main.cpp:247) */
    double rnorm;        /* final recursive ||r|| */
    double bnorm;
} or_pcg_stats;
/* f: output solution [nP][aslm]; hist: 3*record doubles (alpha, beta, ||r||) */
int or_pcg(const or_problem *pr, const or_pcg_params *prm, double *f,
           or_pcg_stats *st, double *hist);

/* ---- post-processing (MeshData.cpp:153-193) ---- */
/* I: [nP][aslm] solution; U[nP]; T[9][nP] (row-major i*3+j); W[3][nT] */
void or_compute_moments(const or_problem *pr, const double *omega_host,
                        const or_idx *omega_pos_host, const double *f,
                        double *U, double *T);
void or_compute_flux(const or_problem *pr, const double *T, double *W);
/* Outward flux density F[nF] per boundary face (mean over the face), the
 * axis rule's O_n (kernels.cu:414-425); 0 on reflecting faces. */
void or_boundary_flux(const or_problem *pr, const double *f, double *F);

/* ---- analytic step problem (main.cpp:498-515, E1.cpp:136-159) ---- */
double or_expn(int n, double x);
/* infinite-medium step (hot x<0, cold x>0, kappa=1): U, W, Txx */
double or_step_U(double x);
double or_step_W(double x);
double or_step_T(double x);
/* finite purely absorbing slab [-L, L] (kappa=1), source I_p=1 on [-L,0] */
double or_slab_U(double x, double L);

#ifdef __cplusplus
}
#endif
#endif
