// p4g_gamma.hip -- discrete-gamma category rates (host, FP64).
//
// Yang (1994) equal-probability categories with the mean rate of each
// category, as p4 computes them (Pf/gamma.c:17-50, DiscreteGamma with
// median = 0, called from Pf/p4_tree.c:236-258 and pf.gdasrvCalcRates).
// The reference implements published algorithms, restated here:
//   chi-square percentage points: Best & Roberts (1975), AS 91;
//   incomplete gamma ratio: Bhattacharjee (1970), AS 32;
//   log gamma: Pike & Hill (1966), CACM algorithm 291;
//   normal percentage points: Odeh & Evans (1974), AS 70.
// Their constants and iteration tolerances (0.5e-6 relative in AS 91, 1e-8
// in AS 32) determine how far the rates sit from exact quantiles (about
// 1e-7 relative), so they are the published ones: a lnL parity target of
// 1e-10 against the reference needs the same rates, not better ones.
#include "p4g_internal.h"

#include <cmath>

namespace p4g {
namespace {

double lnGamma(double alpha)
{
    double x = alpha, f = 0.0;
    if (x < 7) {
        f = 1.0;
        double z = x - 1.0;
        while ((z += 1.0) < 7.0) f *= z;
        x = z;
        f = -std::log(f);
    }
    const double z = 1.0 / (x * x);
    return f + (x - 0.5) * std::log(x) - x + 0.918938533204673 +
           (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z +
            0.083333333333333) / x;
}

double pointNormal(double p)
{
    static const double a[5] = {-0.322232431088, -1.0, -0.342242088547, -0.0204231210245,
                                -0.453642210148e-4};
    static const double b[5] = {0.0993484626060, 0.588581570495, 0.531103462366, 0.103537752850,
                                0.0038560700634};
    const double p1 = (p < 0.5 ? p : 1 - p);
    if (p1 < 1e-20) return -9999;
    const double y = std::sqrt(std::log(1 / (p1 * p1)));
    const double z = y + ((((y * a[4] + a[3]) * y + a[2]) * y + a[1]) * y + a[0]) /
                             ((((y * b[4] + b[3]) * y + b[2]) * y + b[1]) * y + b[0]);
    return p < 0.5 ? -z : z;
}

double incompleteGamma(double x, double p, double g)
{
    const double accurate = 1e-8, overflow = 1e30;
    if (x == 0.0) return 0.0;
    if (x < 0 || p <= 0) return -1.0;
    const double factor = std::exp(p * std::log(x) - x - g);
    if (!(x > 1 && x >= p)) {
        double gin = 1.0, term = 1.0, rn = p;
        do {
            rn++;
            term *= x / rn;
            gin += term;
        } while (term > accurate);
        return gin * (factor / p);
    }
    double a = 1.0 - p, b = a + x + 1.0, term = 0.0;
    double pn[6] = {1.0, x, x + 1, x * b, 0.0, 0.0};
    double gin = pn[2] / pn[3];
    for (;;) {
        a++;
        b += 2.0;
        term++;
        const double an = a * term;
        pn[4] = b * pn[2] - an * pn[0];
        pn[5] = b * pn[3] - an * pn[1];
        if (pn[5] != 0) {
            const double rn = pn[4] / pn[5];
            const double dif = std::fabs(gin - rn);
            if (!(dif > accurate) && dif <= accurate * rn) return 1.0 - factor * gin;
            gin = rn;
        }
        for (int i = 0; i < 4; i++) pn[i] = pn[i + 2];
        if (!(std::fabs(pn[4]) < overflow))
            for (int i = 0; i < 4; i++) pn[i] /= overflow;
    }
}

double pointChi2(double p, double v)
{
    const double e = 0.5e-6, aa = 0.6931471805;
    if (p < 0.000002 || p > 0.999998 || v <= 0.0) return -1.0;
    const double g = lnGamma(v / 2.0);
    const double xx = v / 2.0;
    const double c = xx - 1.0;
    double ch, q, p1, p2, t, a, b;
    if (!(v >= -1.24 * std::log(p))) {
        ch = std::pow((p * xx * std::exp(g + xx * aa)), 1.0 / xx);
        if (ch - e < 0) return ch;
    } else if (!(v > 0.32)) {
        ch = 0.4;
        a = std::log(1.0 - p);
        do {
            q = ch;
            p1 = 1.0 + ch * (4.67 + ch);
            p2 = ch * (6.73 + ch * (6.66 + ch));
            t = -0.5 + (4.67 + 2.0 * ch) / p1 - (6.73 + ch * (13.32 + 3.0 * ch)) / p2;
            ch -= (1.0 - std::exp(a + g + 0.5 * ch + c * aa) * p2 / p1) / t;
        } while (!(std::fabs(q / ch - 1.0) - 0.01 <= 0.0));
    } else {
        const double x = pointNormal(p);
        p1 = 0.222222 / v;
        ch = v * std::pow((x * std::sqrt(p1) + 1.0 - p1), 3.0);
        if (ch > 2.2 * v + 6.0) ch = -2.0 * (std::log(1.0 - p) - c * std::log(0.5 * ch) + g);
    }
    do {
        q = ch;
        p1 = 0.5 * ch;
        if ((t = incompleteGamma(p1, xx, g)) < 0.0) return -1.0;
        p2 = p - t;
        t = p2 * std::exp(xx * aa + g + p1 - c * std::log(ch));
        b = t / ch;
        a = 0.5 * t - b * c;
        const double s1 = (210.0 + a * (140.0 + a * (105.0 + a * (84.0 + a * (70.0 + 60.0 * a))))) / 420.0;
        const double s2 = (420.0 + a * (735.0 + a * (966.0 + a * (1141.0 + 1278.0 * a)))) / 2520.0;
        const double s3 = (210.0 + a * (462.0 + a * (707.0 + 932.0 * a))) / 2520.0;
        const double s4 =
            (252.0 + a * (672.0 + 1182.0 * a) + c * (294.0 + a * (889.0 + 1740.0 * a))) / 5040.0;
        const double s5 = (84.0 + 264.0 * a + c * (175.0 + 606.0 * a)) / 2520.0;
        const double s6 = (120.0 + c * (346.0 + 127.0 * c)) / 5040.0;
        ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
    } while (std::fabs(q / ch - 1.0) > e);
    return ch;
}

}  // namespace

int discreteGamma(double *freqK, double *rK, double alfa, double beta, int K, int median)
{
    if (K < 1) return 1;
    if (K == 1) {
        freqK[0] = 1.0;
        rK[0] = 1.0;
        return 0;
    }
    const double gap05 = 1.0 / (2.0 * K), factor = alfa / beta * K;
    if (median) {
        double t = 0;
        for (int i = 0; i < K; i++) rK[i] = pointChi2((i * 2.0 + 1) * gap05, 2.0 * alfa) / (2.0 * beta);
        for (int i = 0; i < K; i++) t += rK[i];
        for (int i = 0; i < K; i++) rK[i] *= factor / t;
    } else {
        const double lnga1 = lnGamma(alfa + 1);
        for (int i = 0; i < K - 1; i++) freqK[i] = pointChi2((i + 1.0) / K, 2.0 * alfa) / (2.0 * beta);
        for (int i = 0; i < K - 1; i++) freqK[i] = incompleteGamma(freqK[i] * beta, alfa + 1, lnga1);
        rK[0] = freqK[0] * factor;
        rK[K - 1] = (1 - freqK[K - 2]) * factor;
        for (int i = 1; i < K - 1; i++) rK[i] = (freqK[i] - freqK[i - 1]) * factor;
    }
    for (int i = 0; i < K; i++) freqK[i] = 1.0 / K;
    return 0;
}

}  // namespace p4g
