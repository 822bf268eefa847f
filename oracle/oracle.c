/*
 * oracle.c -- CPU FP64 restatement of uranix/Transfer's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Not part of the product; the
 * product must never load it.
 *
 * What each part restates:
 *   angular matrices   AngularData.cpp:9-68, Spherical.cpp:39-82 (values by
 *                      exact monomial integration instead of quadrature)
 *   padding            DeviceAngularData.cpp:7-67, common.cuh:5-7
 *   operator           kernels.cu:135-165 (rhs), 200-287 (volumePart),
 *                      290-357 (volumePartDiag), 383-438 (surfacePart),
 *                      440-493 (surfacePartDiag)
 *   BLAS-1/reductions  kernels.cu:12-115, CudaContext.cpp:188-250
 *   PCG / invprec      main.cpp:125-162, 223-248
 *   moments / flux     MeshData.cpp:96-136
 *   step analytic      main.cpp:498-515, E1.cpp:136-159 (with a correct E_n)
 *
 * Extensions beyond the reference (scattering, relative stopping rule) are
 * marked "extension" and reduce to the reference when switched off.
 */
#include "oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif


/* The angular integrals expand products of harmonics of degree up to 2K + 2
 * into monomials whose coefficients alternate in sign; at K = 7 that sum
 * cancels about nine digits, so it is carried in binary128. */
typedef __float128 qreal;
/* ===================================================================== */
/*  Spherical harmonics as polynomials in (x, y, z)                       */
/* ===================================================================== */

typedef struct {
    int a, b, c;
    qreal v;
} term;

typedef struct {
    int n;
    term t[64];
} poly;

#define GMAX 48
static qreal g_tab[GMAX][GMAX][GMAX];
static int g_ready = 0;

/* (1/4pi) * integral over the unit sphere of |x|^a |y|^b |z|^c */
static qreal avg_abs(int a, int b, int c) {
    return tgammaq((a + 1) / 2.0L) * tgammaq((b + 1) / 2.0L) *
           tgammaq((c + 1) / 2.0L) /
           (2.0Q * M_PIq * tgammaq((a + b + c + 3) / 2.0L));
}

static void g_init(void) {
    if (g_ready) return;
    for (int a = 0; a < GMAX; a++)
        for (int b = 0; b < GMAX; b++)
            for (int c = 0; c < GMAX; c++)
                g_tab[a][b][c] = (a + b + c < GMAX + 20) ? avg_abs(a, b, c) : 0;
    g_ready = 1;
}

/* <x^a y^b z^c>: zero unless every exponent is even */
static qreal mono_avg(int a, int b, int c) {
    if ((a | b | c) & 1) return 0;
    return g_tab[a][b][c];
}

/* <|x_axis| x^a y^b z^c> (axis 0,1,2) */
static qreal mono_avg_absaxis(int axis, int a, int b, int c) {
    if ((a | b | c) & 1) return 0;
    if (axis == 0) return g_tab[a + 1][b][c];
    if (axis == 1) return g_tab[a][b + 1][c];
    return g_tab[a][b][c + 1];
}

static qreal fact(int n) {
    qreal r = 1;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

static qreal binom(int n, int k) {
    return fact(n) / (fact(k) * fact(n - k));
}

/* Legendre polynomial coefficients P_l(z) = sum_k pc[k] z^k */
static void legendre_coeffs(int l, qreal *pc) {
    qreal p0[GMAX] = {0}, p1[GMAX] = {0}, p2[GMAX];
    p0[0] = 1;
    if (l == 0) {
        memcpy(pc, p0, sizeof(qreal) * (l + 1));
        return;
    }
    p1[1] = 1;
    for (int n = 1; n < l; n++) {
        /* (n+1) P_{n+1} = (2n+1) z P_n - n P_{n-1} */
        memset(p2, 0, sizeof(p2));
        for (int k = 0; k <= n; k++) p2[k + 1] += (2 * n + 1) * p1[k];
        for (int k = 0; k <= n - 1; k++) p2[k] -= n * p0[k];
        for (int k = 0; k <= n + 1; k++) p2[k] /= (n + 1);
        memcpy(p0, p1, sizeof(p0));
        memcpy(p1, p2, sizeof(p1));
    }
    memcpy(pc, p1, sizeof(qreal) * (l + 1));
}

/* Y_lm = c_lm * (d^|m| P_l/dz^|m|)(z) * {Re,Im}(x+iy)^|m|.
 * Same convention as Spherical.cpp:39-82: no Condon-Shortley phase,
 * <Y^2> = 1, cos(m phi) for m >= 0, sin(|m| phi) for m < 0. */
static void sh_poly(int l, int m, poly *Y) {
    int am = m < 0 ? -m : m;
    qreal pc[GMAX], dc[GMAX];
    legendre_coeffs(l, pc);
    for (int k = 0; k <= l - am; k++)
        dc[k] = pc[k + am] * fact(k + am) / fact(k);
    qreal c = sqrtq((2.0L * l + 1) * (am == 0 ? 1.0L : 2.0L) *
                          fact(l - am) / fact(l + am));
    Y->n = 0;
    for (int j = 0; j <= am; j++) {
        qreal rv;
        if (m >= 0) {
            if (j & 1) continue;
            rv = binom(am, j) * ((j / 2) & 1 ? -1.0L : 1.0L);
        } else {
            if (!(j & 1)) continue;
            rv = binom(am, j) * (((j - 1) / 2) & 1 ? -1.0L : 1.0L);
        }
        for (int k = 0; k <= l - am; k++) {
            if (dc[k] == 0) continue;
            term *t = &Y->t[Y->n++];
            t->a = am - j;
            t->b = j;
            t->c = k;
            t->v = c * rv * dc[k];
        }
    }
}

/* <P Q x^ea y^eb z^ec> */
static qreal pair_avg(const poly *P, const poly *Q, int ea, int eb,
                            int ec) {
    qreal s = 0;
    for (int i = 0; i < P->n; i++)
        for (int j = 0; j < Q->n; j++)
            s += P->t[i].v * Q->t[j].v *
                 mono_avg(P->t[i].a + Q->t[j].a + ea, P->t[i].b + Q->t[j].b + eb,
                          P->t[i].c + Q->t[j].c + ec);
    return s;
}

static qreal pair_avg_abs(const poly *P, const poly *Q, int axis) {
    qreal s = 0;
    for (int i = 0; i < P->n; i++)
        for (int j = 0; j < Q->n; j++)
            s += P->t[i].v * Q->t[j].v *
                 mono_avg_absaxis(axis, P->t[i].a + Q->t[j].a,
                                  P->t[i].b + Q->t[j].b, P->t[i].c + Q->t[j].c);
    return s;
}

int or_slm(int maxk) { return (1 + maxk) * (1 + 2 * maxk); }

int or_sh_gram(int lmax, double *gram) {
    g_init();
    int n = (lmax + 1) * (lmax + 1);
    poly *Y = (poly *)malloc(sizeof(poly) * n);
    for (int l = 0; l <= lmax; l++)
        for (int m = -l; m <= l; m++) sh_poly(l, m, &Y[l * l + l + m]);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            gram[i * n + j] = (double)pair_avg(&Y[i], &Y[j], 0, 0, 0);
    free(Y);
    return n;
}

int or_sh_eval(int lmax, double x, double y, double z, double *out) {
    int n = (lmax + 1) * (lmax + 1);
    for (int l = 0; l <= lmax; l++)
        for (int m = -l; m <= l; m++) {
            poly Y;
            sh_poly(l, m, &Y);
            qreal s = 0;
            for (int i = 0; i < Y.n; i++)
                s += Y.t[i].v * powq(x, Y.t[i].a) * powq(y, Y.t[i].b) *
                     powq(z, Y.t[i].c);
            out[l * l + l + m] = (double)s;
        }
    return n;
}

/* AngularData.cpp:9-68.  omega[i][lm1][lm2] = <Y1 Y2 Omega_i Omega_j> for the
 * first j whose value exceeds 1e-10 in magnitude, else j = 2 (AngularData.cpp:40);
 * omega_pos records that j.  Ox/Oy/Oz[lm1][lm2] = <|Omega_axis| Y1 Y2>.
 * Extension: with albedo > 0 the stiffness couplings become
 *   <Y2 Omega_i W Omega_j Y1>, W = sum_{odd o} |Y_o>(1 - albedo g_o)^{-1}<Y_o|
 * (exact over odd l_o <= 2K+1) and massfac[lm] = 1 - albedo g_l. */
int or_angular(int maxk, double albedo, double g, int hg_lmax, double *omega,
               or_idx *omega_pos, double *Ox, double *Oy, double *Oz,
               double *massfac) {
    g_init();
    int slm = or_slm(maxk);
    poly *Y = (poly *)malloc(sizeof(poly) * slm);
    int *lof = (int *)malloc(sizeof(int) * slm);
    for (int l = 0, lm = 0; l <= 2 * maxk; l += 2)
        for (int m = -l; m <= l; m++, lm++) {
            sh_poly(l, m, &Y[lm]);
            lof[lm] = l;
        }
    /* scattering moments g_l (HG: g^l, truncated at hg_lmax) */
    qreal gl[GMAX];
    for (int l = 0; l < GMAX; l++) gl[l] = (l <= hg_lmax) ? powq(g, l) : 0;
    gl[0] = 1;
    if (massfac)
        for (int lm = 0; lm < slm; lm++)
            massfac[lm] = (double)(1.0L - albedo * gl[lof[lm]]);

    qreal *val = (qreal *)malloc(sizeof(qreal) * 9 * slm * slm);
    if (albedo == 0) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int e[3] = {0, 0, 0};
                e[i]++;
                e[j]++;
                for (int a = 0; a < slm; a++)
                    for (int b = 0; b < slm; b++)
                        val[((i * 3 + j) * slm + a) * slm + b] =
                            pair_avg(&Y[a], &Y[b], e[0], e[1], e[2]);
            }
    } else {
        int lo = 2 * maxk + 1, no = 0;
        for (int l = 1; l <= lo; l += 2) no += 2 * l + 1;
        poly *Yo = (poly *)malloc(sizeof(poly) * no);
        qreal *w = (qreal *)malloc(sizeof(qreal) * no);
        for (int l = 1, o = 0; l <= lo; l += 2)
            for (int m = -l; m <= l; m++, o++) {
                sh_poly(l, m, &Yo[o]);
                w[o] = 1.0L / (1.0L - albedo * gl[l]);
            }
        qreal *A = (qreal *)malloc(sizeof(qreal) * 3 * slm * no);
        for (int i = 0; i < 3; i++) {
            int e[3] = {0, 0, 0};
            e[i]++;
            for (int a = 0; a < slm; a++)
                for (int o = 0; o < no; o++)
                    A[(i * slm + a) * no + o] =
                        pair_avg(&Y[a], &Yo[o], e[0], e[1], e[2]);
        }
        /* kernels.cu:276-282 multiplies omega[i][lms][lm] into output lm from
         * input lms, so the stored [i][a][b] entry is the operator's
         * (output b, input a) coupling <Y_b Omega_i W Omega_j Y_a>.  Without
         * scattering that equals its transpose; with it, it does not. */
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int a = 0; a < slm; a++)
                    for (int b = 0; b < slm; b++) {
                        qreal s = 0;
                        for (int o = 0; o < no; o++)
                            s += A[(i * slm + b) * no + o] * w[o] *
                                 A[(j * slm + a) * no + o];
                        val[((i * 3 + j) * slm + a) * slm + b] = s;
                    }
        free(A);
        free(w);
        free(Yo);
    }
    for (int i = 0; i < 3; i++)
        for (int a = 0; a < slm; a++)
            for (int b = 0; b < slm; b++) {
                for (int j = 0; j < 3; j++) {
                    double v = (double)val[((i * 3 + j) * slm + a) * slm + b];
                    if (fabs(v) > 1e-10 || j == 2) {
                        omega[i * slm * slm + a * slm + b] = v;
                        omega_pos[i * slm * slm + a * slm + b] = (or_idx)j;
                        break;
                    }
                }
            }
    for (int a = 0; a < slm; a++)
        for (int b = 0; b < slm; b++) {
            Ox[a * slm + b] = (double)pair_avg_abs(&Y[a], &Y[b], 0);
            Oy[a * slm + b] = (double)pair_avg_abs(&Y[a], &Y[b], 1);
            Oz[a * slm + b] = (double)pair_avg_abs(&Y[a], &Y[b], 2);
        }
    free(val);
    free(lof);
    free(Y);
    return slm;
}

/* DeviceAngularData.cpp:10 with COALESCED_NUM(double) = 64/8 (common.cuh:5-6) */
int or_aslm(int slm) { return (slm + 7) & ~7; }

/* DeviceAngularData.cpp:40-54: copy into [.][aslm][aslm], zero elsewhere */
void or_pad_angular(int slm, const double *omega, const or_idx *omega_pos,
                    const double *Ox, const double *Oy, const double *Oz,
                    double *omega_d, or_idx *omega_pos_d, double *Ox_d,
                    double *Oy_d, double *Oz_d) {
    int aslm = or_aslm(slm);
    for (int i = 0; i < aslm; i++)
        for (int j = 0; j < aslm; j++) {
            int in = i < slm && j < slm;
            for (int a = 0; a < 3; a++) {
                omega_d[a * aslm * aslm + i * aslm + j] =
                    in ? omega[a * slm * slm + i * slm + j] : 0;
                omega_pos_d[a * aslm * aslm + i * aslm + j] =
                    in ? omega_pos[a * slm * slm + i * slm + j] : 0;
            }
            Ox_d[i * aslm + j] = in ? Ox[i * slm + j] : 0;
            Oy_d[i * aslm + j] = in ? Oy[i * slm + j] : 0;
            Oz_d[i * aslm + j] = in ? Oz[i * slm + j] : 0;
        }
}

/* ===================================================================== */
/*  Operator (kernels.cu), vertex by vertex                               */
/* ===================================================================== */

static void set_threads(const or_problem *pr) {
#ifdef _OPENMP
    if (pr->nthreads > 0) omp_set_num_threads(pr->nthreads);
#else
    (void)pr;
#endif
}

static double mfac(const or_problem *pr, or_idx lm) {
    return (pr->massfac && lm < pr->slm) ? pr->massfac[lm] : 1.0;
}

/* kernels.cu:200-287.  One "block" per vertex, one "lane" per lm < aslm. */
void or_volume_part(const or_problem *pr, const double *f, double *r) {
    set_threads(pr);
    const or_idx aslm = pr->aslm, slm = pr->slm;
#pragma omp parallel
    {
        double *sums = (double *)malloc(sizeof(double) * 9 * aslm);
        double *sum = (double *)malloc(sizeof(double) * aslm);
#pragma omp for schedule(static)
        for (long vertex = 0; vertex < (long)pr->nP; vertex++) {
            or_idx lo = pr->tetstart[vertex], hi = pr->tetstart[vertex + 1];
            for (or_idx q = 0; q < 9 * aslm; q++) sums[q] = 0;
            for (or_idx lm = 0; lm < aslm; lm++) sum[lm] = 0;
            for (or_idx j = lo; j < hi; j++) {
                or_idx local = pr->tetpos[j];
                const or_tet *tet = &pr->mesh[pr->tetidx[j]];
                for (or_idx lm = 0; lm < aslm; lm++) {
                    double fl[4], fsum = 0, sum_i[3];
                    for (int s = 0; s < 4; s++) {
                        fl[s] = f[(size_t)aslm * tet->p[s] + lm];
                        fsum += fl[s];
                    }
                    sum[lm] += mfac(pr, lm) * tet->kappa_volume *
                               (fl[local] + fsum) * (1. / 20.);
                    for (int si = 0; si < 3; si++)
                        sum_i[si] = tet->s[local][si] / tet->kappa_volume * (1. / 9.);
                    for (int si = 0, v = 0; si < 3; si++)
                        for (int sj = 0; sj < 3; sj++, v++)
                            for (int k = 0; k < 4; k++)
                                sums[v * aslm + lm] += sum_i[si] * fl[k] * tet->s[k][sj];
                }
            }
            /* kernels.cu:276-284: rowsum over a and lms with the omega_pos selector */
            for (or_idx lm = 0; lm < aslm; lm++) {
                double rowsum = 0;
                for (or_idx a = 0; a < 3; a++)
                    for (or_idx lms = 0; lms < slm; lms++) {
                        size_t v = (size_t)a * aslm * aslm + (size_t)lms * aslm + lm;
                        rowsum += pr->omega[v] *
                                  sums[(3 * a + pr->omega_pos[v]) * aslm + lms];
                    }
                r[(size_t)aslm * vertex + lm] = sum[lm] + rowsum;
            }
        }
        free(sums);
        free(sum);
    }
}

/* kernels.cu:290-357: a = b and lm = lm' only */
void or_volume_part_diag(const or_problem *pr, const double *f, double *r) {
    set_threads(pr);
    const or_idx aslm = pr->aslm;
#pragma omp parallel for schedule(static)
    for (long vertex = 0; vertex < (long)pr->nP; vertex++) {
        or_idx lo = pr->tetstart[vertex], hi = pr->tetstart[vertex + 1];
        for (or_idx lm = 0; lm < aslm; lm++) {
            double sum = 0, sum_ij[3] = {0, 0, 0};
            for (or_idx j = lo; j < hi; j++) {
                or_idx local = pr->tetpos[j];
                const or_tet *tet = &pr->mesh[pr->tetidx[j]];
                double fl[4], fsum = 0;
                for (int s = 0; s < 4; s++) {
                    fl[s] = f[(size_t)aslm * tet->p[s] + lm];
                    fsum += fl[s];
                }
                sum += mfac(pr, lm) * tet->kappa_volume * (fl[local] + fsum) * (1. / 20.);
                for (int si = 0; si < 3; si++) {
                    double sum_i = tet->s[local][si] / tet->kappa_volume * (1. / 9.);
                    for (int k = 0; k < 4; k++)
                        sum_ij[si] += sum_i * fl[k] * tet->s[k][si];
                }
            }
            double rowsum = 0;
            for (int a = 0; a < 3; a++)
                rowsum += pr->omega[(size_t)a * aslm * aslm + (size_t)lm * aslm + lm] *
                          sum_ij[a];
            r[(size_t)aslm * vertex + lm] = sum + rowsum;
        }
    }
}

/* kernels.cu:419-425: axis chosen by 2|s_a| > |s_x + s_y + s_z| */
static const double *pick_axis(const or_problem *pr, const or_face *tr,
                               double *surf) {
    *surf = fabs(tr->s[0] + tr->s[1] + tr->s[2]);
    if (2 * fabs(tr->s[0]) > *surf) return pr->Ox;
    if (2 * fabs(tr->s[1]) > *surf) return pr->Oy;
    return pr->Oz;
}

/* kernels.cu:383-438 (r +=) */
void or_surface_part(const or_problem *pr, const double *f, double *r) {
    set_threads(pr);
    const or_idx aslm = pr->aslm, slm = pr->slm;
#pragma omp parallel
    {
        double *fv = (double *)malloc(sizeof(double) * 4 * aslm);
        double *sum = (double *)malloc(sizeof(double) * aslm);
#pragma omp for schedule(static)
        for (long vertex = 0; vertex < (long)pr->nP; vertex++) {
            or_idx lo = pr->facestart[vertex], hi = pr->facestart[vertex + 1];
            for (or_idx lm = 0; lm < aslm; lm++) sum[lm] = 0;
            for (or_idx j = lo; j < hi; j++) {
                or_idx local = pr->facepos[j];
                if (pr->face_kind && pr->face_kind[pr->faceidx[j]] != 0) continue;  /* extension */
                const or_face *tr = &pr->bnd[pr->faceidx[j]];
                double surf;
                const double *On = pick_axis(pr, tr, &surf);
                for (or_idx lm = 0; lm < aslm; lm++) {
                    fv[3 * aslm + lm] = 0;
                    for (int s = 0; s < 3; s++) {
                        double t = f[(size_t)aslm * tr->p[s] + lm];
                        fv[s * aslm + lm] = t;
                        fv[3 * aslm + lm] += t;
                    }
                }
                for (or_idx lm = 0; lm < aslm; lm++)
                    for (or_idx lms = 0; lms < slm; lms++)
                        sum[lm] += surf * On[(size_t)aslm * lms + lm] *
                                   (fv[local * aslm + lms] + fv[3 * aslm + lms]) *
                                   (1. / 12.);
            }
            for (or_idx lm = 0; lm < aslm; lm++) r[(size_t)aslm * vertex + lm] += sum[lm];
        }
        free(fv);
        free(sum);
    }
}

/* kernels.cu:440-493 (r +=), diagonal of O_n only */
void or_surface_part_diag(const or_problem *pr, const double *f, double *r) {
    set_threads(pr);
    const or_idx aslm = pr->aslm;
#pragma omp parallel for schedule(static)
    for (long vertex = 0; vertex < (long)pr->nP; vertex++) {
        or_idx lo = pr->facestart[vertex], hi = pr->facestart[vertex + 1];
        for (or_idx lm = 0; lm < aslm; lm++) {
            double sum = 0;
            for (or_idx j = lo; j < hi; j++) {
                or_idx local = pr->facepos[j];
                if (pr->face_kind && pr->face_kind[pr->faceidx[j]] != 0) continue;  /* extension */
                const or_face *tr = &pr->bnd[pr->faceidx[j]];
                double surf;
                const double *On = pick_axis(pr, tr, &surf);
                double fv[4];
                fv[3] = 0;
                for (int s = 0; s < 3; s++) {
                    fv[s] = f[(size_t)aslm * tr->p[s] + lm];
                    fv[3] += fv[s];
                }
                sum += surf * On[(size_t)aslm * lm + lm] * (fv[local] + fv[3]) * (1. / 12.);
            }
            r[(size_t)aslm * vertex + lm] += sum;
        }
    }
}

/* CudaContext.cpp:160-172 */
void or_compute_lhs(const or_problem *pr, const double *f, double *r) {
    or_volume_part(pr, f, r);
    or_surface_part(pr, f, r);
}

/* CudaContext.cpp:174-186 */
void or_compute_lhs_diag(const or_problem *pr, const double *f, double *r) {
    or_volume_part_diag(pr, f, r);
    or_surface_part_diag(pr, f, r);
}

/* kernels.cu:135-165: lane 0 gets sum kappa*V*I_p/4, other lanes 0 */
void or_compute_rhs(const or_problem *pr, double *b) {
    set_threads(pr);
    const or_idx aslm = pr->aslm;
#pragma omp parallel for schedule(static)
    for (long vertex = 0; vertex < (long)pr->nP; vertex++) {
        double sum = 0;
        for (or_idx j = pr->tetstart[vertex]; j < pr->tetstart[vertex + 1]; j++) {
            const or_tet *tet = &pr->mesh[pr->tetidx[j]];
            sum += tet->kappa_volume * tet->I_p * (1. / 4.);
        }
        for (or_idx lm = 0; lm < aslm; lm++)
            b[(size_t)aslm * vertex + lm] = lm == 0 ? sum : 0;
    }
}

/* ===================================================================== */
/*  BLAS-1 and reductions                                                 */
/* ===================================================================== */

#define NALL ((size_t)pr->nP * pr->aslm)

void or_scale(const or_problem *pr, double *x, double wx) {
    for (size_t i = 0; i < NALL; i++) x[i] *= wx;
}
void or_add_prod(const or_problem *pr, double *x, const double *y, double wy) {
    for (size_t i = 0; i < NALL; i++) x[i] += wy * y[i];
}
void or_mul_add(const or_problem *pr, double *x, double wx, const double *y) {
    for (size_t i = 0; i < NALL; i++) x[i] = wx * x[i] + y[i];
}
void or_mul_add_prod(const or_problem *pr, double *x, double wx, const double *y,
                     double wy) {
    for (size_t i = 0; i < NALL; i++) x[i] = wx * x[i] + wy * y[i];
}

/* Sums over lm < slm only (kernels.cu:60, 93), accumulated in long double */
double or_dot(const or_problem *pr, const double *x, const double *y) {
    long double s = 0;
    for (size_t v = 0; v < pr->nP; v++)
        for (or_idx lm = 0; lm < pr->slm; lm++)
            s += (long double)x[v * pr->aslm + lm] * y[v * pr->aslm + lm];
    return (double)s;
}
double or_norm(const or_problem *pr, const double *x) {
    return sqrt(or_dot(pr, x, x));
}

/* ===================================================================== */
/*  PCG (main.cpp:125-162) and invprec (main.cpp:223-248)                 */
/* ===================================================================== */

static int invprec(const or_problem *pr, const or_pcg_params *prm, double *f,
                   const double *b, double *Ap, double *p, double *r) {
    or_mul_add_prod(pr, f, 0, b, 0);
    or_compute_lhs_diag(pr, f, Ap);
    or_mul_add(pr, r, 0, b);
    or_add_prod(pr, r, Ap, -1);
    or_mul_add(pr, p, 0, r);
    int k = 0;
    double nrz = or_norm(pr, r);
    nrz *= nrz;
    while (k < prm->inner_maxit) {
        or_compute_lhs_diag(pr, p, Ap);
        double alpha = nrz / or_dot(pr, p, Ap);
        or_add_prod(pr, f, p, alpha);
        or_add_prod(pr, r, Ap, -alpha);
        double nr = or_norm(pr, r);
        if (nr < prm->inner_tol) break;
        double nrz2 = nr * nr;
        double beta = nrz2 / nrz;
        nrz = nrz2;
        or_mul_add(pr, p, beta, r);
        k++;
    }
    return k;
}

int or_pcg(const or_problem *pr, const or_pcg_params *prm, double *f,
           or_pcg_stats *st, double *hist) {
    size_t n = NALL;
    double *b = (double *)calloc(n, sizeof(double));
    double *p = (double *)calloc(n, sizeof(double));
    double *Ap = (double *)calloc(n, sizeof(double));
    double *z = (double *)calloc(n, sizeof(double));
    double *r = (double *)calloc(n, sizeof(double));
    double *pp = (double *)calloc(n, sizeof(double));
    double *pAp = (double *)calloc(n, sizeof(double));
    double *pr_ = (double *)calloc(n, sizeof(double));
    int preck = 0;
    memset(f, 0, n * sizeof(double));
    or_compute_rhs(pr, b);
    double bnorm = or_norm(pr, b);
    or_compute_lhs(pr, f, Ap);
    or_mul_add(pr, r, 0, b);
    or_add_prod(pr, r, Ap, -1);
    if (prm->use_prec)
        preck += invprec(pr, prm, z, r, pAp, pp, pr_);
    else
        or_mul_add(pr, z, 0, r);
    or_mul_add(pr, p, 0, z);
    int k = 0;
    double nrz = or_dot(pr, r, z), nr = or_norm(pr, r);
    while (k < prm->maxit) {
        or_compute_lhs(pr, p, Ap);
        double alpha = nrz / or_dot(pr, p, Ap);
        or_add_prod(pr, f, p, alpha);
        or_add_prod(pr, r, Ap, -alpha);
        nr = or_norm(pr, r);
        double beta = 0;
        int stop = nr < prm->tol_abs || (prm->tol_rel > 0 && nr <= prm->tol_rel * bnorm);
        if (!stop) {
            if (prm->use_prec)
                preck += invprec(pr, prm, z, r, pAp, pp, pr_);
            else
                or_mul_add(pr, z, 0, r);
            double nrz2 = or_dot(pr, z, r);
            beta = nrz2 / nrz;
            nrz = nrz2;
            or_mul_add(pr, p, beta, z);
        }
        if (hist && k < prm->record) {
            hist[3 * k + 0] = alpha;
            hist[3 * k + 1] = beta;
            hist[3 * k + 2] = nr;
        }
        if (stop) break;
        k++;
    }
    if (st) {
        st->iters = k;
        st->preck = preck;
        st->rnorm = nr;
        st->bnorm = bnorm;
    }
    free(b); free(p); free(Ap); free(z); free(r); free(pp); free(pAp); free(pr_);
    return 0;
}

/* ===================================================================== */
/*  Post-processing (MeshData.cpp:96-136)                                 */
/* ===================================================================== */

void or_compute_moments(const or_problem *pr, const double *omega_host,
                        const or_idx *omega_pos_host, const double *f, double *U,
                        double *T) {
    const double fourpi = 12.5663706143591729538505735331;
    const or_idx slm = pr->slm, aslm = pr->aslm;
    for (size_t k = 0; k < pr->nP; k++) {
        U[k] = fourpi * f[k * aslm + 0];
        for (int q = 0; q < 9; q++) T[(size_t)q * pr->nP + k] = 0;
        for (or_idx lm = 0; lm < slm; lm++)
            for (int i = 0; i < 3; i++) {
                or_idx j = omega_pos_host[(size_t)i * slm * slm + (size_t)lm * slm];
                double v = omega_host[(size_t)i * slm * slm + (size_t)lm * slm];
                T[(size_t)(3 * i + j) * pr->nP + k] += fourpi * v * f[k * aslm + lm];
            }
    }
}

void or_compute_flux(const or_problem *pr, const double *T, double *W) {
    for (size_t i = 0; i < pr->nT; i++) {
        const or_tet *t = &pr->mesh[i];
        for (int j = 0; j < 3; j++) {
            double w = 0;
            for (int k = 0; k < 4; k++)
                for (int a = 0; a < 3; a++)
                    w += t->s[k][a] * T[(size_t)(3 * a + j) * pr->nP + t->p[k]];
            W[(size_t)j * pr->nT + i] = (1.0 / 3.0) * w / t->kappa_volume;
        }
    }
}

/* Outward flux density through each vacuum boundary face (no reference
 * function outputs it; derived from the vacuum boundary term of
 * kernels.cu:414-437).  At a vacuum face the even-parity form has
 * psi- = sign(Omega.n) psi+, so the outward flux is the integral of
 * |Omega.n| psi+ over the sphere, 4 pi sum_lm O_n[lm][00] f_lm, with O_n
 * chosen by the axis rule of kernels.cu:419-425.  F[j] is its mean over
 * face j (the P1 field's face average: the mean of the three vertices);
 * faces of kind 1 (reflecting) carry none.  Weighted by surf =
 * |s_x + s_y + s_z| (kernels.cu:419) these sum to 4 pi times the lm = 0
 * rows of the boundary term summed over all vertices. */
void or_boundary_flux(const or_problem *pr, const double *f, double *F) {
    const double fourpi = 12.5663706143591729538505735331;
    const or_idx slm = pr->slm, aslm = pr->aslm;
    for (size_t j = 0; j < pr->nF; j++) {
        const or_face *tr = &pr->bnd[j];
        if (pr->face_kind && pr->face_kind[j] != 0) {
            F[j] = 0;
            continue;
        }
        double surf = fabs(tr->s[0] + tr->s[1] + tr->s[2]);
        const double *On = (2 * fabs(tr->s[0]) > surf) ? pr->Ox
                          : (2 * fabs(tr->s[1]) > surf) ? pr->Oy : pr->Oz;
        double acc = 0;
        for (or_idx lm = 0; lm < slm; lm++) {
            double fs = 0;
            for (int k = 0; k < 3; k++) fs += f[(size_t)tr->p[k] * aslm + lm];
            acc += On[(size_t)aslm * lm + 0] * fs;
        }
        F[j] = fourpi * acc / 3.0;
    }
}

/* ===================================================================== */
/*  Exponential integrals and the step problem                           */
/* ===================================================================== */

/* E_1 by its power series for x <= 1 (A&S 5.1.11), E_n by the upward
 * recurrence E_{n+1} = (e^-x - x E_n)/n (A&S 5.1.14); for x > 1 the
 * continued fraction of A&S 5.1.22 evaluated by the modified Lentz method. */
double or_expn(int n, double x) {
    if (x < 0 || n < 0) return NAN;
    if (x == 0) return n > 1 ? 1.0 / (n - 1) : INFINITY;
    if (n == 0) return exp(-x) / x;
    if (x <= 1.0) {
        const long double gamma = 0.57721566490153286060651209008240243L;
        long double s = 0, term = 1;
        for (int k = 1; k < 60; k++) {
            term *= -x / k;           /* (-x)^k / k! */
            s += term / k;
        }
        long double e = -gamma - logl(x) - s;
        for (int m = 1; m < n; m++) e = (expl(-x) - x * e) / m;
        return (double)e;
    }
    /* E_n(x) = e^-x / (x + n - 1*n/(x + n + 2 - 2(n+1)/(x + n + 4 - ...))) */
    const long double tiny = 1e-300L;
    long double b = x + n, c = 1.0L / tiny, d = 1.0L / b, h = d;
    for (int i = 1; i < 500; i++) {
        long double an = -(long double)i * (n - 1 + i);
        b += 2;
        d = 1.0L / (an * d + b);
        c = b + an / c;
        long double del = c * d;
        h *= del;
        if (fabsl(del - 1) < 1e-19L) break;
    }
    return (double)(h * expl(-x));
}

/* main.cpp:498-515 restated with E_n: U = 2 pi E2(|x|), W = 2 pi E3(|x|),
 * T_xx = 2 pi E4(|x|); the hot side (x < 0) is the complement of the
 * uniform field 4pi, 4pi/3. */
double or_step_U(double x) {
    double r = 2 * M_PI * or_expn(2, fabs(x));
    return x < 0 ? 4 * M_PI - r : r;
}
double or_step_W(double x) { return 2 * M_PI * or_expn(3, fabs(x)); }
double or_step_T(double x) {
    double r = 2 * M_PI * or_expn(4, fabs(x));
    return x < 0 ? 4 * M_PI / 3.0 - r : r;
}

/* Purely absorbing slab [-L, L], kappa = 1, I_p = 1 on [-L, 0], vacuum at
 * both ends: U(x) = 2 pi (int_0^1 of the two one-sided intensities). */
double or_slab_U(double x, double L) {
    if (x <= 0)
        return 2 * M_PI * (2 - or_expn(2, x + L) - or_expn(2, -x));
    return 2 * M_PI * (or_expn(2, x) - or_expn(2, x + L));
}
