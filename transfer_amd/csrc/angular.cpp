// angular.cpp -- AngularData: the angular coupling matrices of the even-parity
// spherical-harmonic method, computed on the host at setup.
//
// Restates AngularData.cpp:9-68 with the harmonic convention of
// Spherical.cpp:39-82 (real harmonics, no Condon-Shortley phase, <Y^2> = 1,
// cos(m phi) for m >= 0, sin(|m| phi) for m < 0, lm ordered by even l
// ascending then m = -l..l).  The integrals use a Gauss-Legendre x trapezoid
// product rule that is exact for the polynomial degrees involved (the
// reference used Lebedev and Gauss-Legendre tables; any exact rule gives the
// same matrices up to rounding).
#include <cmath>
#include <stdexcept>

#include "internal.h"
#include "stencil.h"

namespace tr {

// Gauss-Legendre nodes on [-1, 1] by Newton's method on P_n, started from
// the asymptotic node estimate cos(pi (i - 1/4) / (n + 1/2)).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = n == 0 ? 1 : (n == 1 ? z : p1);
            double pm = n == 1 ? 1 : p0;
            dp = n * (z * pn - pm) / (z * z - 1);
            double dz = pn / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double pn = n == 1 ? z : p1, pm = n == 1 ? 1 : p0;
            dp = n * (z * pn - pm) / (z * z - 1);
        }
        double wi = 2.0 / ((1 - z * z) * dp * dp);
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = wi;
        w[n - 1 - i] = wi;
    }
}

// Y_lm for all l <= lmax at the unit vector (x, y, z), out[l*l + l + m].
// Associated Legendre functions without the Condon-Shortley phase by the
// standard three-term recurrence in l at fixed m.
void sh_eval_all(int lmax, double x, double y, double z, double *out) {
    double st = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    double phi = std::atan2(y, x);
    std::vector<double> P((lmax + 1) * (lmax + 1), 0.0);  // P[l*(lmax+1)+m]
    double pmm = 1.0;
    for (int m = 0; m <= lmax; m++) {
        if (m > 0) pmm *= (2 * m - 1) * st;
        P[m * (lmax + 1) + m] = pmm;
        if (m + 1 <= lmax) P[(m + 1) * (lmax + 1) + m] = z * (2 * m + 1) * pmm;
        for (int l = m + 2; l <= lmax; l++)
            P[l * (lmax + 1) + m] = ((2 * l - 1) * z * P[(l - 1) * (lmax + 1) + m] -
                                     (l + m - 1) * P[(l - 2) * (lmax + 1) + m]) /
                                    (l - m);
    }
    for (int l = 0; l <= lmax; l++)
        for (int m = -l; m <= l; m++) {
            int am = m < 0 ? -m : m;
            // c_lm^2 = (2l+1)(2 - delta_m0)(l-|m|)!/(l+|m|)!  (Spherical.cpp:39-47)
            double c2 = (2.0 * l + 1.0) * (am == 0 ? 1.0 : 2.0);
            for (int i = l - am + 1; i <= l + am; i++) c2 /= i;
            double ang = m >= 0 ? std::cos(am * phi) : std::sin(am * phi);
            out[l * l + l + m] = std::sqrt(c2) * P[l * (lmax + 1) + am] * ang;
        }
}

void build_angular(int maxk, double albedo, double g, int hg_lmax, Angular &a) {
    if (maxk < 0 || maxk > 7)
        throw std::runtime_error("angorder must be in 0..7 (slm <= ASLM_MAX=128)");
    a.maxk = maxk;
    const int slm = (1 + maxk) * (1 + 2 * maxk);
    a.slm = slm;
    a.l_of.resize(slm);
    a.m_of.resize(slm);
    for (int l = 0, lm = 0; l <= 2 * maxk; l += 2)
        for (int m = -l; m <= l; m++, lm++) {
            a.l_of[lm] = l;
            a.m_of[lm] = m;
        }
    const int lfull = 2 * maxk + 1;  // odd harmonics up to 2K+1 for scattering
    const int nall = (lfull + 1) * (lfull + 1);

    // Full-sphere product rule: GL in mu (exact to degree 2n-1) x M azimuths
    // (exact for |k| < M); integrands have degree <= 4K+2.
    std::vector<double> gx, gw;
    const int nmu = 2 * maxk + 3, M = 4 * maxk + 6;
    gauss_legendre(nmu, gx, gw);
    const int npt = nmu * M;
    std::vector<double> Yev((size_t)npt * slm), Yall((size_t)npt * nall),
        dir((size_t)npt * 3), wt(npt);
    for (int i = 0; i < nmu; i++)
        for (int j = 0; j < M; j++) {
            int q = i * M + j;
            double mu = gx[i], s = std::sqrt(std::fmax(0.0, 1 - mu * mu));
            double ph = 2 * M_PI * (j + 0.5) / M;
            dir[3 * q + 0] = s * std::cos(ph);
            dir[3 * q + 1] = s * std::sin(ph);
            dir[3 * q + 2] = mu;
            wt[q] = gw[i] * 0.5 / M;
            sh_eval_all(lfull, dir[3 * q], dir[3 * q + 1], dir[3 * q + 2],
                        &Yall[(size_t)q * nall]);
            for (int lm = 0; lm < slm; lm++) {
                int l = a.l_of[lm], m = a.m_of[lm];
                Yev[(size_t)q * slm + lm] = Yall[(size_t)q * nall + l * l + l + m];
            }
        }

    // val[i][j][lm1][lm2]: the coupling that multiplies input lm1 into output
    // lm2 (the index order of kernels.cu:276-282).
    std::vector<double> val((size_t)9 * slm * slm, 0.0);
    if (albedo == 0) {
        for (int q = 0; q < npt; q++) {
            const double *Y = &Yev[(size_t)q * slm];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double wq = wt[q] * dir[3 * q + i] * dir[3 * q + j];
                    double *v = &val[(size_t)(i * 3 + j) * slm * slm];
                    for (int a1 = 0; a1 < slm; a1++) {
                        double t = wq * Y[a1];
                        for (int a2 = 0; a2 < slm; a2++) v[a1 * slm + a2] += t * Y[a2];
                    }
                }
        }
    } else {
        // W = sum over odd o of |Y_o> (1 - albedo g_o)^-1 <Y_o|; the coupling is
        // <Y_out Omega_i W Omega_j Y_in>.
        std::vector<int> odd;
        for (int l = 1; l <= lfull; l += 2)
            for (int m = -l; m <= l; m++) odd.push_back(l * l + l + m);
        const int no = (int)odd.size();
        std::vector<double> A((size_t)3 * slm * no, 0.0), wo(no);
        for (int o = 0; o < no; o++) {
            int l = (int)std::lround(std::floor(std::sqrt((double)odd[o])));
            double gl = l <= hg_lmax ? std::pow(g, l) : 0.0;
            wo[o] = 1.0 / (1.0 - albedo * gl);
        }
        for (int q = 0; q < npt; q++)
            for (int i = 0; i < 3; i++)
                for (int e = 0; e < slm; e++) {
                    double t = wt[q] * dir[3 * q + i] * Yev[(size_t)q * slm + e];
                    for (int o = 0; o < no; o++)
                        A[((size_t)i * slm + e) * no + o] += t * Yall[(size_t)q * nall + odd[o]];
                }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int a1 = 0; a1 < slm; a1++)
                    for (int a2 = 0; a2 < slm; a2++) {
                        double s = 0;
                        for (int o = 0; o < no; o++)
                            s += A[((size_t)i * slm + a2) * no + o] * wo[o] *
                                 A[((size_t)j * slm + a1) * no + o];
                        val[((size_t)(i * 3 + j) * slm + a1) * slm + a2] = s;
                    }
    }
    a.massfac.resize(slm);
    for (int lm = 0; lm < slm; lm++) {
        int l = a.l_of[lm];
        double gl = (l == 0) ? 1.0 : (l <= hg_lmax ? std::pow(g, l) : 0.0);
        a.massfac[lm] = 1.0 - albedo * gl;
    }
    // Selector of AngularData.cpp:40: first j with |value| > 1e-10, else j = 2.
    a.omega.assign((size_t)3 * slm * slm, 0.0);
    a.omega_pos.assign((size_t)3 * slm * slm, 0);
    for (int i = 0; i < 3; i++)
        for (int a1 = 0; a1 < slm; a1++)
            for (int a2 = 0; a2 < slm; a2++)
                for (int j = 0; j < 3; j++) {
                    double v = val[((size_t)(i * 3 + j) * slm + a1) * slm + a2];
                    if (std::fabs(v) > 1e-10 || j == 2) {
                        a.omega[(size_t)i * slm * slm + a1 * slm + a2] = v;
                        a.omega_pos[(size_t)i * slm * slm + a1 * slm + a2] = j;
                        break;
                    }
                }

    // <|Omega_n| Y1 Y2> = (1/2pi) int over the hemisphere Omega_n > 0 of
    // Omega_n Y1 Y2 (the integrand is even under Omega -> -Omega); GL on
    // [0, 1] in the axial cosine x M azimuths.
    const int nh = 2 * maxk + 2;
    gauss_legendre(nh, gx, gw);
    a.Ox.assign((size_t)slm * slm, 0.0);
    a.Oy.assign((size_t)slm * slm, 0.0);
    a.Oz.assign((size_t)slm * slm, 0.0);
    std::vector<double> Y(nall), Ye(slm);
    for (int axis = 0; axis < 3; axis++) {
        std::vector<double> &O = axis == 0 ? a.Ox : (axis == 1 ? a.Oy : a.Oz);
        for (int i = 0; i < nh; i++)
            for (int j = 0; j < M; j++) {
                double mu = 0.5 * (gx[i] + 1.0), w = 0.5 * gw[i] / M;
                double s = std::sqrt(std::fmax(0.0, 1 - mu * mu));
                double ph = 2 * M_PI * (j + 0.5) / M;
                double d[3];
                d[axis] = mu;
                d[(axis + 1) % 3] = s * std::cos(ph);
                d[(axis + 2) % 3] = s * std::sin(ph);
                sh_eval_all(2 * maxk, d[0], d[1], d[2], Y.data());
                for (int lm = 0; lm < slm; lm++) {
                    int l = a.l_of[lm], m = a.m_of[lm];
                    Ye[lm] = Y[l * l + l + m];
                }
                for (int a1 = 0; a1 < slm; a1++)
                    for (int a2 = 0; a2 < slm; a2++)
                        O[(size_t)a1 * slm + a2] += w * mu * Ye[a1] * Ye[a2];
            }
    }
}

}  // namespace tr

namespace tr {

// an orthonormal frame (t1, t2, n) around the unit vector of nn
static void normal_frame(const double nn[3], double t1[3], double t2[3], double n[3]) {
    const double len = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    for (int a = 0; a < 3; a++) n[a] = nn[a] / len;
    if (std::fabs(n[0]) < 0.9) {
        t1[0] = 0; t1[1] = -n[2]; t1[2] = n[1];
    } else {
        t1[0] = n[2]; t1[1] = 0; t1[2] = -n[0];
    }
    const double l1 = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    for (int a = 0; a < 3; a++) t1[a] /= l1;
    t2[0] = n[1] * t1[2] - n[2] * t1[1];
    t2[1] = n[2] * t1[0] - n[0] * t1[2];
    t2[2] = n[0] * t1[1] - n[1] * t1[0];
}

void abs_normal_matrix(int maxk, const double nn[3], std::vector<double> &O) {
    const int slm = (1 + maxk) * (1 + 2 * maxk), lmax = 2 * maxk;
    double t1[3], t2[3], n[3];
    normal_frame(nn, t1, t2, n);
    std::vector<double> gx, gw;
    const int nh = 2 * maxk + 2, M = 4 * maxk + 6;
    gauss_legendre(nh, gx, gw);
    std::vector<double> Yall((size_t)(lmax + 1) * (lmax + 1)), Y(slm);
    O.assign((size_t)slm * slm, 0.0);
    for (int i = 0; i < nh; i++) {
        const double mu = 0.5 * (gx[i] + 1.0), wmu = 0.5 * gw[i], s = std::sqrt(std::fmax(0.0, 1 - mu * mu));
        for (int j = 0; j < M; j++) {
            const double ph = 2 * M_PI * (j + 0.5) / M, c = std::cos(ph), sn = std::sin(ph);
            double d[3];
            for (int a = 0; a < 3; a++) d[a] = mu * n[a] + s * (c * t1[a] + sn * t2[a]);
            sh_eval_all(lmax, d[0], d[1], d[2], Yall.data());
            for (int l = 0, lm = 0; l <= lmax; l += 2)
                for (int m = -l; m <= l; m++, lm++) Y[lm] = Yall[l * l + l + m];
            // (1/2pi) int over the hemisphere of mu Y Y' = sum wmu (1/M) mu Y Y'
            const double w = wmu * mu / M;
            for (int a1 = 0; a1 < slm; a1++)
                for (int a2 = 0; a2 < slm; a2++) O[(size_t)a1 * slm + a2] += w * Y[a1] * Y[a2];
        }
    }
}

// Column lm' = 00 of abs_normal_matrix(n): <|Omega.n| Y_lm>.  |Omega.n|
// is zonal about n, so by the Funk-Hecke formula <g(Omega.n) Y_lm> =
// lambda_l Y_lm(n) with lambda_l = (1/2) int_{-1}^{1} g(t) P_l(t) dt, which
// for g = |t| and even l is int_0^1 t P_l(t) dt (Gauss-Legendre on [0, 1],
// exact for its degree l + 1).
void abs_normal_column(int maxk, const double nn[3], double *col) {
    const int lmax = 2 * maxk;
    const double len = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    std::vector<double> Yall((size_t)(lmax + 1) * (lmax + 1));
    sh_eval_all(lmax, nn[0] / len, nn[1] / len, nn[2] / len, Yall.data());
    std::vector<double> gx, gw;
    gauss_legendre(maxk + 2, gx, gw);
    for (int l = 0, lm = 0; l <= lmax; l += 2) {
        double lam = 0;
        for (size_t i = 0; i < gx.size(); i++) {
            const double t = 0.5 * (gx[i] + 1.0);
            double p0 = 1, p1 = t;
            for (int k = 2; k <= l; k++) {
                const double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            lam += 0.5 * gw[i] * t * (l == 0 ? 1.0 : p1);
        }
        for (int m = -l; m <= l; m++, lm++) col[lm] = lam * Yall[l * l + l + m];
    }
}

uint32_t rotation_size(int maxk) {
    uint32_t s = 0;
    for (int l = 0; l <= 2 * maxk; l += 2) s += (uint32_t)((2 * l + 1) * (2 * l + 1));
    return s;
}

void normal_rotation(int maxk, const double nn[3], std::vector<double> &R) {
    const int lmax = 2 * maxk;
    double t1[3], t2[3], n[3];
    normal_frame(nn, t1, t2, n);
    // Gauss-Legendre in mu x trapezoid in phi over the whole sphere, exact
    // for the products of two harmonics of degree <= 2K
    std::vector<double> gx, gw;
    const int nq = 2 * maxk + 1, M = 4 * maxk + 2;
    gauss_legendre(nq, gx, gw);
    std::vector<double> Yl((size_t)(lmax + 1) * (lmax + 1)), Yg((size_t)(lmax + 1) * (lmax + 1));
    R.assign(rotation_size(maxk), 0.0);
    for (int i = 0; i < nq; i++) {
        const double mu = gx[i], s = std::sqrt(std::fmax(0.0, 1 - mu * mu));
        for (int j = 0; j < M; j++) {
            const double ph = 2 * M_PI * (j + 0.5) / M, c = std::cos(ph), sn = std::sin(ph);
            const double w = 0.5 * gw[i] / M;  // the average over the sphere
            const double lx = s * c, ly = s * sn, lz = mu;
            double d[3];
            for (int a = 0; a < 3; a++) d[a] = lx * t1[a] + ly * t2[a] + lz * n[a];
            sh_eval_all(lmax, lx, ly, lz, Yl.data());
            sh_eval_all(lmax, d[0], d[1], d[2], Yg.data());
            size_t off = 0;
            for (int l = 0; l <= lmax; l += 2) {
                const int nb = 2 * l + 1;
                for (int a = 0; a < nb; a++)
                    for (int b = 0; b < nb; b++) R[off + (size_t)a * nb + b] += w * Yg[l * l + a] * Yl[l * l + b];
                off += (size_t)nb * nb;
            }
        }
    }
}

}  // namespace tr
