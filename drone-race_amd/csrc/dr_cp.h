// Clopper-Pearson bounds for the plan verification of the branching MPPI
// controller (state.h:134-151, MPPIController::computeCertifiedLoss /
// computeEpsilon, mppi_controller.cu:468-514).  The reference evaluates the
// beta quantiles with Boost.Math; this is an independent FP64 evaluation:
// the regularized incomplete beta function by its continued fraction
// (modified Lentz), inverted by safeguarded Newton steps inside a bisection
// bracket.  Host and device compile the same code.
#pragma once

#include <math.h>

#include "dr_rng.h"

namespace dr {
namespace cp {

// log B(a, b)
DR_HD double lbeta(double a, double b)
{
    return lgamma(a) + lgamma(b) - lgamma(a + b);
}

// Continued fraction of I_x(a, b) (valid for x < (a + 1) / (a + b + 2)).
DR_HD double betacf(double a, double b, double x)
{
    const double tiny = 1e-300, eps = 1e-16;
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (fabs(d) < tiny)
        d = tiny;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m <= 1000; ++m)
    {
        const double m2 = 2.0 * m;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < tiny)
            d = tiny;
        c = 1.0 + aa / c;
        if (fabs(c) < tiny)
            c = tiny;
        d = 1.0 / d;
        h *= d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < tiny)
            d = tiny;
        c = 1.0 + aa / c;
        if (fabs(c) < tiny)
            c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) < eps)
            break;
    }
    return h;
}

// Regularized incomplete beta I_x(a, b), a, b > 0.
DR_HD double ibeta(double a, double b, double x)
{
    if (x <= 0.0)
        return 0.0;
    if (x >= 1.0)
        return 1.0;
    const double lf = a * log(x) + b * log1p(-x) - lbeta(a, b);
    if (x < (a + 1.0) / (a + b + 2.0))
        return exp(lf) * betacf(a, b, x) / a;
    return 1.0 - exp(lf) * betacf(b, a, 1.0 - x) / b;
}

// x with I_x(a, b) = p, 0 < p < 1.
DR_HD double ibeta_inv(double a, double b, double p)
{
    double lo = 0.0, hi = 1.0;
    double x = a / (a + b); // the mean as the first guess
    const double lb = lbeta(a, b);
    for (int it = 0; it < 200; ++it)
    {
        const double f = ibeta(a, b, x) - p;
        if (f < 0.0)
            lo = x;
        else
            hi = x;
        const double dens = exp((a - 1.0) * log(x) + (b - 1.0) * log1p(-x) - lb);
        double nx = (dens > 0.0) ? x - f / dens : 0.5 * (lo + hi);
        if (!(nx > lo && nx < hi)) // Newton left the bracket: bisect
            nx = 0.5 * (lo + hi);
        if (fabs(nx - x) <= 1e-15 * x || hi - lo <= 1e-15 * hi)
        {
            x = nx;
            break;
        }
        x = nx;
    }
    return x;
}

// clopperPearsonLowerBound / UpperBound (state.h:134-151): the beta
// quantiles of the exact binomial interval, with the reference's special
// cases k = 0 and k = n.
DR_HD double cp_lower(unsigned k, unsigned n, double alpha)
{
    if (k == 0)
        return 0.0;
    return ibeta_inv((double)k, (double)(n - k + 1), alpha);
}

DR_HD double cp_upper(unsigned k, unsigned n, double alpha)
{
    if (k == n)
        return 1.0;
    return ibeta_inv((double)(k + 1), (double)(n - k), 1.0 - alpha);
}

} // namespace cp
} // namespace dr
