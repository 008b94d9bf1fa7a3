// template_tap.hpp -- one tap of a drift-z template, FP64 on the device (shared by
// the FFT'd template bank and the direct-correlation bank).
#pragma once

#include "fdas_common.hpp"
#include "fresnel.hpp"

namespace fdas {

// Tap i (of numkern) of the drift-z response with roffset 0 and 2 points per
// bin: r_i = numkern/4 - z/2 - i/2 and the response is
// int_0^1 exp(2 pi i (r_i u + z u^2 / 2)) du
// (source/bpd_subharmonics.cu:146-207; z = 0 uses the interpolation kernel of
// :78-135, whose centre tap is exactly 1).
__device__ inline cplx template_tap(int z, int numkern, int i)
{
    const double kPi = 3.14159265358979323846;
    const double kSqrt2 = 1.41421356237309504880;
    if (z == 0) {
        if (i == numkern / 2) return mk(1.0f, 0.0f);
        const double rr = kPi * (numkern / 4.0) - kPi * 0.5 * i;
        double sr, cr;
        sincos(rr, &sr, &cr);
        const double sinc = sr / rr;
        return mk((float)(cr * sinc), (float)(sr * sinc));
    }
    const double zd_abs = fabs((double)z);
    const int signz = (z < 0) ? -1 : 1;
    const double zd = signz * kSqrt2 / sqrt(zd_abs);
    const double cons = zd / 2.0;
    const double pibyz = kPi / (double)z;
    const double r = (0.0 - 0.5 * z) + numkern / 4.0 - 0.5 * i;
    const double yy = r * zd;
    const double zz = yy + z * zd;
    const double xx = pibyz * r * r;
    double sx, cx;
    sincos(xx, &sx, &cx);
    double fsy, fcy, fsz, fcz;
    fresnel_d(yy, &fsy, &fcy);
    fresnel_d(zz, &fsz, &fcz);
    const double tr = signz * (fcz - fcy);
    const double ti = fsy - fsz;
    return mk((float)((tr * cx - ti * sx) * cons), (float)(-(tr * sx + ti * cx) * cons));
}

__device__ __forceinline__ int d_z_response_halfwidth(double z)
{
    z = fabs(z);
    int m = (int)(z * (0.00089 * z + 0.3131) + NUMFINTBINS);
    m = (m < NUMFINTBINS) ? NUMFINTBINS : m;
    if (z > 100 && m > 0.6 * z) m = (int)(0.6 * z);
    return m;
}

}  // namespace fdas
