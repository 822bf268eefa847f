/*
 * transfer.h -- C ABI of transfer_amd, an MI355X-native drop-in for the
 * CudaContext + kernels.cu path of uranix/Transfer.
 *
 * Plain C types only: pointers, sizes, doubles.  Every call returns 0 on
 * success and a negative code on failure (tr_last_error() has the text),
 * except the few that return a value; those document their failure value.
 *
 * Vector convention (the reference's): a "vector" is nP*aslm REAL values laid
 * out [vertex][aslm] in the caller's vertex order, aslm = slm rounded up to
 * 64 bytes (DeviceAngularData.cpp:10, common.cuh:5-6), padding lanes zero.
 * REAL is double (the reference's USE_DOUBLE build, util.h:24-28).
 * Device vectors returned by tr_alloc_vector() are stored internally in a
 * renumbered, unpadded layout; tr_copy_to_dev/tr_copy_to_host convert when
 * the device side is such a vector, so callers only ever see the reference
 * layout.  Raw buffers from tr_device_alloc() are copied byte for byte.
 */
#ifndef TRANSFER_AMD_H
#define TRANSFER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TR_OK 0
#define TR_ERR_ARG (-1)
#define TR_ERR_HIP (-2)
#define TR_ERR_STATE (-3)
#define TR_ERR_NOCONV (-4)

/* ------------------------------------------------------------------------ */
/* Host data structures                                                      */
/* ------------------------------------------------------------------------ */

/* Same bytes as the reference's tetrahedron (tetrahedron.h:7-20) and face
 * (face.h:8-18) in the USE_DOUBLE build: 128 and 64 bytes. */
#pragma pack(push, 1)
typedef struct tr_tetrahedron {
    uint32_t p[4];
    double kappa_volume; /* kappa * V */
    double I_p;
    double s[4][3];      /* s_j = 3 V grad(e_j) = -S_j n_j (MeshData.cpp:65-73) */
} tr_tetrahedron;

typedef struct tr_face {
    uint32_t p[4];       /* p[3] unused */
    double s[3];         /* S * n (MeshData.cpp:79-91) */
    double _padd[3];
} tr_face;
#pragma pack(pop)

/* Kernel-argument view of MeshData (MeshData.h:13-32, DeviceMeshDataRaw
 * MeshData.h:34-47).  Arrays are borrowed for the duration of the call that
 * takes the view. */
typedef struct tr_mesh_view {
    uint32_t nP, nF, nT;
    const uint32_t *tetstart, *tetidx, *tetpos;
    const tr_tetrahedron *mesh;
    const uint32_t *facestart, *faceidx, *facepos;
    const tr_face *bnd;
    /* Optional, may be NULL.  coords[nP][3] drives the Morton renumbering
     * (without it the device numbering comes from a reverse Cuthill-McKee
     * pass over the vertex graph).  face_kind[nF]: 0 = vacuum (the
     * reference's only kind), 1 = specular reflection (no boundary term). */
    const double *coords;
    const uint8_t *face_kind;
} tr_mesh_view;

/* View of AngularData (AngularData.h:8-18): unpadded host arrays
 * omega/omega_pos [3][slm][slm], Ox/Oy/Oz [slm][slm]; massfac[slm] is an
 * extension (collision factor 1 - albedo*g_l per harmonic; NULL = 1, the
 * reference's scattering-free operator). */
typedef struct tr_angular_view {
    uint32_t slm;
    const double *omega;
    const uint32_t *omega_pos;
    const double *Ox, *Oy, *Oz;
    const double *massfac;
} tr_angular_view;

/* ------------------------------------------------------------------------ */
/* Front end: AngularData, MeshData, Config                                  */
/* ------------------------------------------------------------------------ */

typedef struct tr_angular tr_angular;
typedef struct tr_mesh tr_mesh;
typedef struct tr_config tr_config;

/* Replaces AngularData(int maxk) (AngularData.cpp:9-68).  albedo/g/hg_lmax
 * add scattering: albedo = 0 gives the reference's matrices; hg_lmax = 0
 * with albedo > 0 is isotropic scattering; hg_lmax = L is Henyey-Greenstein
 * with asymmetry g truncated after P_L. */
int tr_angular_create(int maxk, double albedo, double g, int hg_lmax,
                      tr_angular **out);
int tr_angular_destroy(tr_angular *a);
int tr_angular_get_view(const tr_angular *a, tr_angular_view *v);

/* Replaces MeshData(const Config&) (MeshData.cpp:9-94), whose mesh came from
 * meshProcessor's Netgen reader: here meshes come from generators.
 * A box [lo,hi] split into nx*ny*nz cubes of 6 Kuhn tetrahedra each.
 * jitter > 0 moves every vertex not on the boundary surface by
 * U[-jitter, jitter] cell widths per axis (seeded).  ball_radius > 0 keeps
 * only cubes whose centroid lies within ball_radius of ball_center. */
typedef struct tr_mesh_spec {
    int nx, ny, nz;
    double lo[3], hi[3];
    double jitter;
    double ball_radius;
    double ball_center[3];
    uint64_t seed;
} tr_mesh_spec;
int tr_mesh_create_box(const tr_mesh_spec *spec, tr_mesh **out);
/* General meshes: vertices coords[nP][3], tets tets[nT][4]; per-tet
 * kappa/Ip may be NULL (defaults 1 and 0). */
int tr_mesh_create(uint32_t nP, const double *coords, uint32_t nT,
                   const uint32_t *tets, const double *kappa, const double *Ip,
                   tr_mesh **out);
int tr_mesh_destroy(tr_mesh *m);
int tr_mesh_get_view(const tr_mesh *m, tr_mesh_view *v);
/* coords [nP][3], tets [nT][4], centroids [nT][3], volumes [nT] */
int tr_mesh_get_geometry(const tr_mesh *m, double *coords, uint32_t *tets,
                         double *centroids, double *volumes);
/* Per-tet fields (the reference's [region] kappa / Ip, Config.cpp:88-101). */
int tr_mesh_set_fields(tr_mesh *m, const double *kappa, const double *Ip);
/* Boundary-face kinds (0 vacuum, 1 reflective) by an axis-aligned plane
 * test: faces lying on x_axis = value (|x - value| < tol) get `kind`. */
int tr_mesh_set_face_kind_plane(tr_mesh *m, int axis, double value, double tol,
                                int kind);

/* Netgen neutral volume mesh (.vol), the reference's mesh input (read by
 * meshProcessor in MeshData.cpp:10): `points` and `volumeelements` (1-based
 * point numbers; material number = region).  With a config, kappa and I_p
 * of every tet come from its region's [region N] section (MeshData.cpp:
 * 62-63), else kappa = 1, I_p = 0.  Other sections are ignored; boundary
 * faces are the faces with one tet. */
int tr_mesh_load_vol(const char *path, const tr_config *cfg, tr_mesh **out);
/* region[nT]: the material numbers of a loaded mesh (0 for generated ones). */
int tr_mesh_get_regions(const tr_mesh *m, uint32_t *region);
/* Legacy VTK unstructured grid of the fields main.cpp:209-216 saves (through
 * meshProcessor's saveVtk there): W[3][nT] per tet (vector), U[nP],
 * T[9][nP] (tensor, the layout of tr_compute_moments), aU, aW, aT [nP] per
 * vertex.  Host arrays; any may be NULL. */
int tr_write_vtk(const tr_mesh *m, const char *path, const double *U, const double *T,
                 const double *W, const double *aU, const double *aW, const double *aT);

/* Per-tet field generators (seeded; SplitMix64 counter-based RNG). */
enum {
    TR_FIELD_CONST = 0,      /* p[0]                                          */
    TR_FIELD_RADIAL_EXP = 1, /* p[0] * exp(-r / p[1]), r = |centroid - p[2..4]| */
    TR_FIELD_UNIFORM = 2,    /* p[0] + p[1] * U[0,1)                          */
    TR_FIELD_LOGNORMAL = 3,  /* exp(p[0] + p[1] * N(0,1))                     */
    TR_FIELD_BLOBS = 4,      /* p[0] + sum of n=p[1] Gaussians: amplitude
                                log-uniform [p[2],p[3]], width U[p[4],p[5]],
                                centre U[lo,hi]^3 of the mesh bounding box     */
    TR_FIELD_PLANCK = 5,     /* (1 - p[0] * r)^4, r as RADIAL_EXP (p[2..4])   */
    TR_FIELD_STEP = 6        /* centroid[axis=p[0]] < p[1] ? p[2] : p[3]       */
};
int tr_field_eval(const tr_mesh *m, int kind, const double *params,
                  uint64_t seed, uint64_t stream, double *out /* [nT] */);
/* SplitMix64 uniforms in [0,1): out[i] = U(seed, stream, i) */
int tr_rng_uniform(uint64_t seed, uint64_t stream, uint64_t n, double *out);

/* Config (Config.h:4-48, Config.cpp:8-131): same cfg grammar and keys. */
int tr_config_load(const char *path, tr_config **out);
int tr_config_destroy(tr_config *c);
/* ints: maxk, device, dump, precondition, soltype, numregs */
int tr_config_ints(const tr_config *c, int *ints6);
int tr_config_region(const tr_config *c, int region, double *kappa, double *Ip);
int tr_config_strings(const tr_config *c, char *meshfile, char *outfile,
                      size_t cap);

/* ------------------------------------------------------------------------ */
/* Device context: the CudaContext of CudaContext.h:9-50                     */
/* ------------------------------------------------------------------------ */

typedef struct tr_ctx tr_ctx;

/* CudaContext(dev, MeshData, AngularData) (CudaContext.cpp:24-50):
 * uploads mesh and angular data (DeviceMeshData.cpp:8-52,
 * DeviceAngularData.cpp:7-67).  flags: 0 (the reference's operator) or
 * TR_CTX_GENERAL_NORMALS. */
/* Vacuum faces whose normal is not along an axis get the exact boundary
 * term <|Omega.n| Y Y'> of their own normal and their true area (extension;
 * the reference's axis rule, kernels.cu:359-361 "WORKS ONLY IF NORMAL IS
 * (+/-1,0,0)", Issue #15, stays the default). Supported by the operator,
 * the Diag operator, the FP64 and mixed device solvers with use_prec 0, 2
 * and 4; not by the multilevel preconditioner or batches. */
#define TR_CTX_GENERAL_NORMALS 1
int tr_ctx_create(int device, const tr_mesh_view *mesh,
                  const tr_angular_view *ang, int flags, tr_ctx **out);
int tr_ctx_destroy(tr_ctx *ctx);               /* ~CudaContext CudaContext.cpp:52-58 */
uint32_t tr_ctx_N(const tr_ctx *ctx);          /* N() = nP*aslm CudaContext.cpp:67-69 */
uint32_t tr_ctx_slm(const tr_ctx *ctx);
uint32_t tr_ctx_aslm(const tr_ctx *ctx);
uint32_t tr_ctx_nP(const tr_ctx *ctx);
/* Autotuner: times every candidate schedule of the operator kernel
 * (harmonics per lane, wavefronts per workgroup, LDS staging scheme) on
 * this context's mesh and keeps the fastest.  precision 0 = FP64 operator,
 * 1 = FP32 operator (mixed solves).  report (may be NULL) gets one line per
 * candidate.  No reference counterpart (the reference launches fixed
 * geometry, CudaContext.cpp:108-120). */
int tr_ctx_autotune(tr_ctx *ctx, int precision, char *report, size_t cap);
/* The schedule in use: out5 = {J, L, VPW, WPB, variant} */
int tr_ctx_schedule(tr_ctx *ctx, int precision, int *out5);
/* Forces a schedule (J harmonics per lane, WPB wavefronts per workgroup,
 * variant 1 = whole-wavefront LDS staging, 2 = batched staging, 3 = batched
 * staging with buffer loads and per-slot coefficient records, 4 = angular
 * contraction on the matrix cores, J ignored; 5 = the same for 33..80
 * harmonics in blocks of 16: FP32 with BF16 matrix-core products, or FP64
 * (precision 0, WPB 8) in the reflection-parity classes where they apply:
 * a single problem without reflecting planes or faces off the axes). */
int tr_ctx_set_schedule(tr_ctx *ctx, int precision, int J, int WPB, int variant);
/* (variant 5: the matrix-core schedule for 33..80 harmonics, FP32 operator
 * with a BF16 angular precision only, J ignored.) */
/* Multilevel preconditioner (use_prec 3 of the device solvers, extension):
 * Galerkin coarse operators of the same Kronecker form under aggregation of
 * `agg` consecutive vertices, coarsened until at most `coarsest` vertices
 * remain; symmetric V-cycle with one damped-Jacobi step per side and `nu`
 * on the coarsest level (defaults 3000, 16, 2). */
int tr_ctx_set_multilevel(tr_ctx *ctx, uint32_t coarsest, uint32_t agg, int nu);
/* Chebyshev polynomial preconditioner (use_prec 4 of the device solvers,
 * extension): `degree` steps of the Chebyshev iteration on D^-1 A z = D^-1 r
 * over [lambda_max / ratio, lambda_max] (lambda_max of D^-1 A from a power
 * iteration), each step one operator launch with the update fused into its
 * write.  step_split (TR_SOLVER_MIXED with a BF16 matrix-core operator):
 * the BF16 split of the steps, 0 = the operator's own, 2 = the gathered
 * values split and the couplings BF16 only, 1 = plain BF16.  Defaults 4,
 * 100, 2. */
int tr_ctx_set_chebyshev(tr_ctx *ctx, int degree, double ratio, int step_split);
/* Angular precision of the FP32 operator used by TR_SOLVER_MIXED:
 * 0 = FP32 throughout, 1 = BF16 matrix-core projection with FP32
 * accumulation, 3 (default) = BF16 with the hi/lo split (three BF16
 * products per term, ~16 significant bits; two, the gathered values alone
 * split, where the couplings' lo parts do not fit the LDS), 4 = for 33..80
 * harmonics under the parity-class layout (one problem, K = 4 or 5): FP32
 * couplings and products on the FP32 matrix cores (error ~2e-7 of
 * max|A f| against ~6e-6 for 3); elsewhere as 3. */
int tr_ctx_set_angular_precision(tr_ctx *ctx, int prec);
/* Robustness of TR_SOLVER_MIXED (extension; the reference has only its
 * iteration caps, main.cpp:142, 232): when a refinement step cuts the true
 * FP64 residual by less than `stagnation` (it must reach stagnation times
 * the previous one), or makes it larger or non-finite (that update is
 * taken back), or the inner CG breaks down, the solve continues with
 * TR_SOLVER_PCG_FP64 (same use_prec) from the last good iterate and
 * reports stats.fallback = 2.  The inner CG stops early when its residual
 * norm has not dropped by 10% over `inner_stall_window` iterations (0 =
 * never).  Defaults 0.5 and 100. */
int tr_ctx_set_refinement(tr_ctx *ctx, double stagnation, int inner_stall_window);
/* TR_SOLVER_MIXED's refinement steps, in tr_solve and (per problem) in
 * tr_batch_solve (extension; the reference has no mixed solver):
 * 1 (default) = each step after the first continues the inner CG of the
 * previous one, its residual replaced by the new true FP64 residual and its
 * last search direction kept (a "reliable update", Sleijpen and van der
 * Vorst 1996), while the true residual stays within 2% of the inner CG's
 * recursive one (else that step restarts); 0 = each step restarts the
 * inner CG from a zero correction.  inner_tol is then the inner reduction
 * between two replacements: 1e-2 measured best. */
int tr_ctx_set_residual_replacement(tr_ctx *ctx, int on);
/* Returns the internal numbering: perm[new] = caller vertex id. */
int tr_ctx_permutation(const tr_ctx *ctx, uint32_t *perm);

void *tr_device_alloc(tr_ctx *ctx, size_t bytes);   /* zero-filled, CudaContext.cpp:60-65; NULL on failure */
double *tr_alloc_vector(tr_ctx *ctx);               /* CudaContext.cpp:71-73; NULL on failure */
int tr_device_free(tr_ctx *ctx, void *mem);         /* CudaContext.cpp:75-77 */
int tr_copy_to_dev(tr_ctx *ctx, void *dst, const void *src, size_t bytes);   /* CudaContext.cpp:79-81 */
int tr_copy_to_host(tr_ctx *ctx, void *dst, const void *src, size_t bytes);  /* CudaContext.cpp:83-85 */
/* getHostSmall (CudaContext.cpp:87-98): dst[nP][slm] (caller-allocated, not malloc'd) */
int tr_get_host_small(tr_ctx *ctx, const double *src, double *dst);

int tr_compute_lhs(tr_ctx *ctx, const double *f, double *Af);      /* CudaContext.cpp:108-120 */
int tr_compute_lhs_diag(tr_ctx *ctx, const double *f, double *Af); /* CudaContext.cpp:122-134 */
int tr_compute_rhs(tr_ctx *ctx, double *b);                        /* CudaContext.cpp:99-106 */
/* computeLhs with the operator in FP32 (precision 1; the inner operator of
 * the mixed solver) or FP64 (precision 0); f and Af stay FP64 vectors. */
int tr_compute_lhs_prec(tr_ctx *ctx, int precision, const double *f, double *Af);
/* MeshData::ComputeMoments (MeshData.cpp:96-116) of the device vector f:
 * device arrays U[nP] = 4 pi f[00] and T[3i+j][nP], caller's vertex order. */
int tr_compute_moments(tr_ctx *ctx, const double *f, double *U, double *T);
/* MeshData::ComputeFlux (MeshData.cpp:118-136): device W[j][nT] from the
 * device T of tr_compute_moments and the mesh view the context was built
 * from. */
int tr_compute_flux(tr_ctx *ctx, const tr_mesh_view *mesh, const double *T, double *W);
/* Boundary flux (the "boundary flux" output of the reference workflow, which
 * no reference function computes; derived from the vacuum boundary term of
 * kernels.cu:414-437): device F[nF] in the view's face order, the outward
 * flux density through each boundary face averaged over the face,
 * 4 pi sum_lm O_n[lm][00] f_lm with the face's O_n as the operator takes it
 * (axis rule, kernels.cu:419-425, or its own normal under
 * TR_CTX_GENERAL_NORMALS); 0 on reflecting faces.  Times the face area
 * (|s_x + s_y + s_z| under the axis rule) it sums to 4 pi times the lm = 0
 * rows of the boundary term.  Single-problem contexts only. */
int tr_compute_boundary_flux(tr_ctx *ctx, const tr_mesh_view *mesh, const double *f, double *F);
int tr_scale(tr_ctx *ctx, double *x, double wx);                   /* CudaContext.cpp:137-144 */
int tr_add_prod(tr_ctx *ctx, double *x, const double *y, double wy);          /* CudaContext.cpp:147-154 */
int tr_mul_add(tr_ctx *ctx, double *x, double wx, const double *y);           /* CudaContext.cpp:157-164 */
int tr_mul_add_prod(tr_ctx *ctx, double *x, double wx, const double *y, double wy); /* CudaContext.cpp:166-173 */
/* norm/dot (CudaContext.cpp:175-198) over the slm real lanes; NaN on failure */
double tr_norm(tr_ctx *ctx, const double *x);
double tr_dot(tr_ctx *ctx, const double *x, const double *y);
int tr_synchronize(tr_ctx *ctx);
/* No counterpart in the reference: tr_dot for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  *result
 * (a device double) = x . y over the slm real lanes, computed on the
 * context's stream with no allocation, no host copy and no
 * synchronisation; the partials go to a static per-GPU scratch array of
 * 1024 doubles (8 KB, held until the process ends), so vectors of any
 * length fit.  Bit-identical to tr_dot.  Calls that overlap on one GPU
 * share the scratch and race. */
int tr_dot_noalloc(tr_ctx *ctx, const double *x, const double *y, double *result);

const char *tr_last_error(void);

/* ------------------------------------------------------------------------ */
/* Solvers (extensions of main.cpp:125-162 / 223-248)                        */
/* ------------------------------------------------------------------------ */

enum {
    TR_SOLVER_REFERENCE = 0, /* main.cpp's PCG + invprec, host-driven, FP64 */
    TR_SOLVER_PCG_FP64 = 1,  /* CG with device-resident scalars and graph-
                                captured iterations; use_prec 0 = none,
                                2 = Jacobi (exact diagonal of A),
                                3 = multilevel, 4 = Chebyshev polynomial,
                                5 = Jacobi plus the learned coarse
                                correction g (g.r)/(g.Ag), g the learned
                                guess passed in f (x0 = 1; a rejected guess
                                leaves Jacobi)                            */
    TR_SOLVER_MIXED = 2      /* FP64 iterative refinement (true residual
                                b - A x in FP64) around the FP32 operator's
                                device CG run to inner_tol (relative), or
                                to the reduction still needed when that is
                                less; maxit caps the refinement steps       */
};

typedef struct tr_solve_params {
    int solver;          /* TR_SOLVER_* */
    int use_prec;        /* 1: invprec inner CG on the Diag operator
                            (TR_SOLVER_REFERENCE); 2: Jacobi           */
    int maxit;           /* outer cap (main.cpp:142: 1000) */
    double tol_abs;      /* stop when ||r|| < tol_abs (main.cpp:149: 1e-4) */
    double tol_rel;      /* and/or ||r|| <= tol_rel*||b|| (0 = off) */
    int inner_maxit;     /* main.cpp:232: 1000 */
    double inner_tol;    /* main.cpp:239: 1e-8 absolute */
    int record;          /* leading iterations whose (alpha, beta, ||r||) to record */
    int x0;              /* TR_SOLVER_PCG_FP64 / MIXED: start from the contents of f
                            (e.g. the learned guess), scaled by (f.b)/(f.Af)
                            (its energy-norm best multiple, never farther from
                            the solution than 0), instead of 0; the solve falls
                            back to 0 when that scaling is undefined (NaN, Inf,
                            f.Af <= 0) or f.b <= 0 */
} tr_solve_params;

typedef struct tr_solve_stats {
    int iters;           /* outer iterations (main.cpp's k at exit) */
    int preck;           /* total inner iterations */
    double rnorm;        /* final recursive ||r|| */
    double bnorm;        /* ||b|| */
    double true_rel;     /* ||b - A f|| / ||b|| with the FP64 operator */
    double seconds;      /* wall time of the solve, device-synchronised */
    int status;          /* 0 converged, TR_ERR_NOCONV otherwise */
    int fallback;        /* 1: the x0 guess was rejected and the solve started
                            from 0; 2: TR_SOLVER_MIXED stagnated and finished
                            with TR_SOLVER_PCG_FP64 (tr_ctx_set_refinement) */
} tr_solve_stats;

/* Solves A f = b, b = computeRhs, from f = 0.  f is a device vector.
 * hist: 3*record doubles (alpha, beta, ||r||) or NULL. */
int tr_solve(tr_ctx *ctx, const tr_solve_params *prm, double *f,
             tr_solve_stats *st, double *hist);

/* ------------------------------------------------------------------------ */
/* Batches (extension: the reference solves one mesh per run).               */
/* ------------------------------------------------------------------------ */

/* A context over nprob independent problems with the same angular data:
 * the block-diagonal union of their operators, problem p with isotropic
 * scattering albedo[p] (NULL = none; the angular view must carry no
 * scattering of its own).  Every tr_ctx call works on the union: vectors
 * are [union vertex][aslm] with problem p's vertices at voff[p] ..
 * voff[p+1] (tr_batch_info) in the order of `meshes`.  flags: 0 or
 * TR_CTX_GENERAL_NORMALS (as tr_ctx_create, for every problem). */
int tr_batch_create(int device, int nprob, const tr_mesh_view *meshes,
                    const tr_angular_view *ang, const double *albedo, int flags,
                    tr_ctx **out);
/* Number of problems and their vertex offsets voff[nprob+1] (may be NULL). */
int tr_batch_info(const tr_ctx *ctx, uint32_t *nprob, uint32_t *voff);

typedef struct tr_batch_stats {
    int iters;           /* CG iterations (FP64 solver) or refinement steps */
    int inner_iters;     /* FP32 CG iterations (mixed solver) */
    double true_rel;     /* ||b - A f|| / ||b|| of the problem, FP64 operator */
    int status;          /* 0 converged, TR_ERR_NOCONV otherwise */
} tr_batch_stats;

/* Solves every problem of a batch from f = 0 to its own stopping test
 * (TR_SOLVER_PCG_FP64 or TR_SOLVER_MIXED; use_prec 0, 2 = Jacobi or 4 =
 * Chebyshev with one lambda_max for the whole union) with per-problem CG
 * scalars; one operator launch per iteration covers all problems.  total: aggregate (iters = max, preck = sum of inner, true_rel =
 * worst problem); per: [nprob] or NULL. */
int tr_batch_solve(tr_ctx *ctx, const tr_solve_params *prm, double *f,
                   tr_solve_stats *total, tr_batch_stats *per);

/* ------------------------------------------------------------------------ */
/* Learned initial guess (extension): a per-vertex MLP, 8 features, two     */
/* hidden layers of 64 (SiLU), sigmoid output = U / U_inf with              */
/* U_inf = 4 pi I_p / (1 - albedo); BF16 weights, FP32 accumulation.        */
/* ------------------------------------------------------------------------ */
#define TR_MLP_FEATURES 8
#define TR_MLP_PARAMS 4801   /* W1[64][8] b1[64] W2[64][64] b2[64] w3[64] b3 */

typedef struct tr_mlp tr_mlp;
/* Per-vertex features (caller's order) of a mesh with coordinates and its
 * albedo: feat[nP][8] (see features.cpp for the list). */
int tr_mlp_features(const tr_mesh_view *mesh, double albedo, float *feat);
/* params: TR_MLP_PARAMS FP32 values, or NULL for a seeded initialisation. */
int tr_mlp_create(int device, const float *params, uint64_t seed, tr_mlp **out);
int tr_mlp_destroy(tr_mlp *m);
/* Device buffers owned by the network (freed with it) and copies between
 * them and the host (to_host 0: host -> device). */
int tr_mlp_alloc(tr_mlp *m, size_t bytes, void **out);
int tr_mlp_copy(tr_mlp *m, void *dst, const void *src, size_t bytes, int to_host);
/* Feature normalisation xh = (x - mu) / sigma (8 values each). */
int tr_mlp_set_norm(tr_mlp *m, const float *mu, const float *sigma);
int tr_mlp_get_params(tr_mlp *m, float *params);
/* y[n] (device) from device features feat[n][8]. */
int tr_mlp_predict(tr_mlp *m, const float *feat, uint32_t n, float *y);
/* Mean squared error of y against target[n] (device arrays) and its
 * gradient (host, TR_MLP_PARAMS floats) with respect to the BF16 weights. */
int tr_mlp_grad(tr_mlp *m, const float *feat, const float *target, uint32_t n,
                float *loss, float *grad);
/* `steps` Adam steps on minibatches of `batch` samples drawn from the n
 * device samples; loss_hist[steps] (host, may be NULL). */
int tr_mlp_train(tr_mlp *m, const float *feat, const float *target, uint32_t n,
                 int steps, uint32_t batch, float lr, uint64_t seed, float *loss_hist);
/* f (device vector) = the guess with scalar radiation U = 4 pi scale y,
 * i.e. f[v][00] = scale * y[v], other harmonics 0; y[nP] is a device array
 * in the caller's vertex order (scale = I_p / (1 - albedo) for the MLP). */
int tr_ctx_scalar_guess(tr_ctx *ctx, const float *y, double scale, double *f);
/* The features of tr_mlp_features computed on the device from the
 * context's data (contexts built from meshes with coordinates; batches use
 * each problem's albedo): feat[nP][8] device, caller's vertex order.  The
 * optical depth is the same shortest-path value, by parallel Bellman-Ford
 * sweeps; *sweeps (may be NULL) receives their number. */
int tr_ctx_mlp_features(tr_ctx *ctx, double albedo, float *feat, int *sweeps);
/* f (device vector) = the learned guess: device features, the network's
 * prediction and f[v][00] = y I_p / (1 - albedo), with I_p = ip_scale or,
 * when ip_scale <= 0, each vertex's own source (exp of feature 7). */
int tr_ctx_learned_guess(tr_ctx *ctx, tr_mlp *m, double albedo, double ip_scale,
                         double *f);

/* The dense system dump of main.cpp:73-114 (small meshes, N*N <= 2^28):
 * one text line per unknown k outside the padding lanes with the nP*slm
 * entries of A e_k and b[k]; then the symmetry check |Z_ij - Z_ji| >
 * 100 eps (|Z_ij| + |Z_ji|), eps the unit roundoff (Util.cpp:4-14).
 * *n_asym receives the number of pairs i < j that fail it, *max_ratio the
 * largest |Z_ij - Z_ji| / (eps (|Z_ij| + |Z_ji|)); both may be NULL. */
int tr_dump_slae(tr_ctx *ctx, const char *path, uint64_t *n_asym, double *max_ratio);

/* Lanczos estimate of the operator's extreme eigenvalues (Eigenvalue.cpp:
 * 5-88, eigest; disabled in main.cpp:117-123): up to `steps` Lanczos steps
 * from the right-hand side, on A (jacobi 0) or on D^-1/2 A D^-1/2, the
 * spectrum of D^-1 A (jacobi 1), stopping when the largest Ritz value moves
 * by less than tol of itself (eigest: 1e-6; 0 = all steps).  out4 =
 * {smallest Ritz value, largest, the largest extrapolated as eigest does,
 * steps taken}. */
int tr_eig_estimate(tr_ctx *ctx, int steps, int jacobi, double tol, double *out4);

/* Seconds per operator application (median over reps), for benchmarking. */
double tr_bench_lhs(tr_ctx *ctx, int reps);
/* Times `steps` operator applications (precision 0 FP64, 1 the FP32
 * operator of the mixed solver, 2 the FP64 operator on vectors already in
 * its schedule's harmonic order: under the FP64 K-blocked schedule the
 * kernel alone, without the permutation of an internal-order input) after
 * `warmup` untimed ones, with HIP events on the context's stream around
 * each launch.  Precision 3 times the FP32 operator with its fused f.Af
 * dot (as a solver iteration launches it), 4 the same with a 1 GiB copy
 * before each application (the cache state after a solve's vector passes;
 * the copy is outside the timed events).  out[0] wall seconds of the timed
 * loop (device-synchronised on both sides), out[1] mean, out[2] min, out[3]
 * max ms per application. */
int tr_bench_operator(tr_ctx *ctx, int precision, int warmup, int steps, double *out4);
/* Name of the HIP kernel of the current operator schedule, as profilers
 * print it (without the namespace and argument list). */
int tr_ctx_kernel_name(tr_ctx *ctx, int precision, char *buf, size_t cap);
/* A plain device copy of `bytes` timed with HIP events (out[0] = ms per
 * copy): the calibration kernel for memory-traffic counters. */
int tr_bench_copy(tr_ctx *ctx, size_t bytes, int reps, double *out1);

/* ------------------------------------------------------------------------ */
/* Introspection of the assembled operator data (host only, no device work). */
/* No reference counterpart: the reference rebuilds the couplings inside     */
/* kernels.cu on every application.  Tests use it to check the assembly.    */
/* ------------------------------------------------------------------------ */
typedef struct tr_stencil tr_stencil;
int tr_stencil_create(const tr_mesh_view *mesh, const tr_angular_view *ang,
                      tr_stencil **out);
int tr_stencil_destroy(tr_stencil *s);
/* info: nP, slm, NC, W, nB, WB, arow, brow */
int tr_stencil_info(const tr_stencil *s, uint32_t *info8);
/* The Galerkin coarse operator of the multilevel preconditioner: vertices
 * aggregated in runs of `size` consecutive internal ids (aggregate I =
 * agg[I] .. agg[I+1]-1, agg has ceil(nP/size)+1 entries, may be NULL).
 * The result's perm is the identity over aggregates. */
int tr_stencil_coarsen(const tr_stencil *s, uint32_t size, tr_stencil **out, uint32_t *agg);
/* Specular reflection (extension; faces of kind 1 on axis-aligned planes):
 * the entries v*slm + lm of the internal [v][slm] numbering that vanish,
 * the harmonics odd under a reflecting plane's mirror at its vertices.
 * *count receives the number; list (may be NULL) receives the entries. */
int tr_stencil_mask(const tr_stencil *s, uint32_t *count, uint32_t *list);
/* Any pointer may be NULL.  perm[nP]; nbr[W][nP]; coef[W][NC][nP];
 * bvert[nB]; bnbr[WB][nB]; bcoef[WB][3][nB]; omc[NC-1][slm][slm];
 * onc[3][slm][slm]; massfac[slm]; rhs[nP] (all in the internal numbering) */
int tr_stencil_export(const tr_stencil *s, uint32_t *perm, uint32_t *nbr,
                      double *coef, uint32_t *bvert, uint32_t *bnbr,
                      double *bcoef, double *omc, double *onc, double *massfac,
                      double *rhs);

/* The exact boundary matrix of a vacuum face with normal n (general
 * normals): O(n) = <|Omega.n| Y Y'> [slm][slm] by a Gauss-Legendre x
 * trapezoid rule on the hemisphere around n; its column lm' = 00 by the
 * Funk-Hecke formula; and the block-diagonal rotation R of the even
 * harmonics (one (2l+1)^2 block per even l, *size values) with
 * O(n) = R O(z) R^T, the form the face kernel applies.  Host only. */
int tr_abs_normal_matrix(int maxk, const double *n, double *O);
int tr_abs_normal_column(int maxk, const double *n, double *col);
int tr_normal_rotation(int maxk, const double *n, double *R, uint32_t *size);

#ifdef __cplusplus
}
#endif
#endif
