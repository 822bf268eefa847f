// analytic.cpp -- exponential integrals and the step-problem solutions the
// reference compares against (main.cpp:498-515 with E1.cpp:136-159).
//
// The reference's xE1(x) is meant to be x*E1(x) but returns x*e^x*E1(x) for
// x > 1 (E1.cpp:148-157), so its U2/W2/T2 are wrong beyond one mean free
// path.  Here the same closed forms are written with a correct E_n:
//   U2 = 2 pi E2(|x|)            (4 pi - 2 pi E2(|x|) for x < 0)
//   W2 = 2 pi E3(|x|)
//   T2 = 2 pi E4(|x|)            (4 pi/3 - 2 pi E4(|x|) for x < 0)
// which is what e^-x - x E1 = E2, ((1-x) e^-x + x^2 E1)/2 = E3 and
// ((2 - x + x^2) e^-x - x^3 E1)/6 = E4 give for the reference's expressions.
#include <cmath>
#include <limits>

#include "internal.h"
#include "transfer.h"

namespace {

const double kEuler = 0.577215664901532860606512090082402431;

// E_n(x) for n >= 0, x >= 0: the power series of A&S 5.1.12 for x <= 1 and
// the continued fraction of A&S 5.1.22, evaluated by the modified Lentz
// method, for x > 1.
double expn(int n, double x) {
    if (n < 0 || x < 0 || std::isnan(x)) return std::numeric_limits<double>::quiet_NaN();
    if (n == 0) return x == 0 ? std::numeric_limits<double>::infinity() : std::exp(-x) / x;
    if (x == 0) return n == 1 ? std::numeric_limits<double>::infinity() : 1.0 / (n - 1);
    const double eps = 1e-17, tiny = 1e-300;
    if (x > 1.0) {
        double b = x + n, c = 1.0 / tiny, d = 1.0 / b, h = d;
        for (int i = 1; i < 100000; i++) {
            const double a = -(double)i * (n - 1 + i);
            b += 2.0;
            d = a * d + b;
            if (std::fabs(d) < tiny) d = tiny;
            c = b + a / c;
            if (std::fabs(c) < tiny) c = tiny;
            d = 1.0 / d;
            const double del = c * d;
            h *= del;
            if (std::fabs(del - 1.0) < eps) break;
        }
        return h * std::exp(-x);
    }
    const int nm1 = n - 1;
    double sum = nm1 != 0 ? 1.0 / nm1 : -std::log(x) - kEuler;
    double fact = 1.0;
    for (int i = 1; i < 1000; i++) {
        fact *= -x / i;
        double del;
        if (i != nm1) {
            del = -fact / (i - nm1);
        } else {
            double psi = -kEuler;
            for (int k = 1; k <= nm1; k++) psi += 1.0 / k;
            del = fact * (-std::log(x) + psi);
        }
        sum += del;
        if (std::fabs(del) < std::fabs(sum) * eps) break;
    }
    return sum;
}

}  // namespace

extern "C" {

double tr_expn(int n, double x) { return expn(n, x); }

int tr_step_analytic(double x, double *uwt) {
    if (!uwt || std::isnan(x)) return tr::fail(TR_ERR_ARG, "tr_step_analytic: bad argument");
    const double ax = std::fabs(x), tp = 2.0 * M_PI;
    const double u = tp * expn(2, ax), t = tp * expn(4, ax);
    uwt[0] = x < 0 ? 4.0 * M_PI - u : u;
    uwt[1] = tp * expn(3, ax);
    uwt[2] = x < 0 ? 4.0 * M_PI / 3.0 - t : t;
    return TR_OK;
}

double tr_slab_analytic_U(double x, double L) {
    if (!(L > 0) || x < -L || x > L) return std::numeric_limits<double>::quiet_NaN();
    const double tp = 2.0 * M_PI;
    if (x <= 0) return tp * (2.0 - expn(2, x + L) - expn(2, -x));
    return tp * (expn(2, x) - expn(2, x + L));
}

}  // extern "C"
