// fresnel.hpp -- FP64 Fresnel integrals on the device (template taps and the
// candidate optimiser's response function).
#pragma once

#include <hip/hip_runtime.h>

namespace fdas {

// ------------------------------------------------------------------ Fresnel
// C(x) = int_0^x cos(pi t^2/2) dt, S(x) = int_0^x sin(pi t^2/2) dt in FP64.
// |x| < 1.5: Maclaurin series  x * sum_k s_k a^k / (k! (2k+1)), a = pi x^2 / 2
//            (even k -> C, odd k -> S, s_k = + + - - ...).
// |x| >= 1.5: C + iS = (1+i)/2 (1 - erfc(w)), w = sqrt(pi)/2 (1-i) x, with
//            erfc(w) = (1-i) x e^{i pi x^2/2} h and h the Laplace continued
//            fraction 1/(b1 + a2/(b2 + ...)), b_k = 1 - i pi x^2 + 4(k-1),
//            a_k = -(2k-3)(2k-2), evaluated by the modified Lentz method.
// |x| > 36974: 0.5 (the saturation of the reference's fresnl, bpd_fresnl.c:170).
__device__ inline void fresnel_d(double xin, double* s_out, double* c_out)
{
    const double kPi = 3.14159265358979323846;
    double x = fabs(xin), c, s;
    if (x < 1.5) {
        const double a = 0.5 * kPi * x * x;
        double t = 1.0, sc = 0.0, ss = 0.0;
        for (int k = 0; k < 60; k++) {
            if (k > 0) t *= a / k;
            const double term = t / (2 * k + 1);
            const double sg = ((k & 3) < 2) ? 1.0 : -1.0;
            if (k & 1) ss += sg * term;
            else sc += sg * term;
            if (t < 1e-18) break;
        }
        c = x * sc;
        s = x * ss;
    } else if (x > 36974.0) {
        c = 0.5;
        s = 0.5;
    } else {
        const double px2 = kPi * x * x;
        double br = 1.0, bi = -px2;
        double cr = 1e300, ci = 0.0;
        double den = br * br + bi * bi;
        double dr = br / den, di = -bi / den;
        double hr = dr, hi = di;
        for (int k = 2; k < 2000; k++) {
            const double n = 2 * k - 3;
            const double an = -n * (n + 1.0);
            br += 4.0;
            const double tr = an * dr + br, ti = an * di + bi;
            den = tr * tr + ti * ti;
            dr = tr / den;
            di = -ti / den;
            den = cr * cr + ci * ci;
            cr = br + an * cr / den;
            ci = bi - an * ci / den;
            const double er = cr * dr - ci * di, ei = cr * di + ci * dr;
            const double nr = hr * er - hi * ei, ni = hr * ei + hi * er;
            hr = nr;
            hi = ni;
            if (fabs(er - 1.0) + fabs(ei) < 1e-17) break;
        }
        const double gr = (hr + hi) * x, gi = (hi - hr) * x;
        double sp, cp;
        sincos(0.5 * px2, &sp, &cp);
        const double pr = cp * gr - sp * gi, pim = cp * gi + sp * gr;
        const double ur = 1.0 - pr, ui = -pim;
        c = 0.5 * (ur - ui);
        s = 0.5 * (ur + ui);
    }
    if (xin < 0.0) {
        c = -c;
        s = -s;
    }
    *c_out = c;
    *s_out = s;
}

}  // namespace fdas
