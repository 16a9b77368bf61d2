// deterministic elementary functions of the HIP path
#pragma once
#include <math.h>
// Deterministic sin/cos shared by the HIP path and the oracle: Cody-Waite
// reduction by pi/2 (pi/2 split into three parts of 33, 33 and 53 bits) and
// Taylor polynomials on [-pi/4, pi/4] in Horner form.  Evaluated with the
// same operations on host and device (no contraction), both sides produce the
// same bits; the error against libm is below 2 ulp for |x| < 1e3.
#define DET_PIO2_1 1.5707963267341256
#define DET_PIO2_2 6.077100506303966e-11
#define DET_PIO2_3 2.0222662487959506e-21
#define DET_2OPI 0.6366197723675814
#define DET_FN __host__ __device__ __forceinline__

DET_FN double det_sincos(double x, int cosine)
{
    double k = rint(x * DET_2OPI);
    double r = ((x - k * DET_PIO2_1) - k * DET_PIO2_2) - k * DET_PIO2_3;
    int q = (int)(((long long)k + cosine) & 3);
    double r2 = r * r;
    double s = r + r * r2 * (-0.16666666666666666 + r2 * (0.008333333333333333 + r2 * (-0.0001984126984126984 +
               r2 * (2.7557319223985893e-06 + r2 * (-2.505210838544172e-08 + r2 * (1.6059043836821613e-10 +
               r2 * (-7.647163731819816e-13 + r2 * 2.8114572543455206e-15)))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (0.041666666666666664 + r2 * (-0.001388888888888889 +
               r2 * (2.48015873015873e-05 + r2 * (-2.755731922398589e-07 + r2 * (2.08767569878681e-09 +
               r2 * (-1.1470745597729725e-11 + r2 * (4.779477332387385e-14 + r2 * -1.5619206968586225e-16))))))));
    if (q == 0) return s;
    if (q == 1) return c;
    if (q == 2) return -s;
    return -c;
}
DET_FN double det_sin(double x) { return det_sincos(x, 0); }
DET_FN double det_cos(double x) { return det_sincos(x, 1); }

// Deterministic natural logarithm, same operations on host and device:
// x = 2^e f with f in [sqrt(1/2), sqrt(2)), log f = 2 atanh(s) with
// s = (f-1)/(f+1), |s| < 0.172, from the Taylor series of atanh through
// s^23 (the next term is below 1e-20), plus e ln2 with ln2 split so that
// e * DET_LN2_HI is exact.  Error against libm below 2 ulp.
#define DET_LN2_HI 0.6931471803691238
#define DET_LN2_LO 1.9082149292705877e-10
DET_FN double det_log(double x)
{
    if (!(x > 0)) return x == 0 ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int e;
    double f = frexp(x, &e);
    if (f < 0.7071067811865476) {
        f *= 2;
        e -= 1;
    }
    double s = (f - 1) / (f + 1);
    double s2 = s * s;
    double q = 1.0 / 3 + s2 * (1.0 / 5 + s2 * (1.0 / 7 + s2 * (1.0 / 9 + s2 * (1.0 / 11 + s2 * (1.0 / 13 +
               s2 * (1.0 / 15 + s2 * (1.0 / 17 + s2 * (1.0 / 19 + s2 * (1.0 / 21 + s2 * (1.0 / 23))))))))));
    double lf = 2 * s + 2 * s * (s2 * q);
    return e * DET_LN2_HI + (e * DET_LN2_LO + lf);
}

// Deterministic exponential, same operations on host and device: x = k ln2
// + r with k = rint(x / ln2) and ln2 split as above (|r| <= 0.35), exp r
// from its Taylor series through r^14 (the next term is below 1e-18),
// scaled by 2^k.  Error against libm below 2 ulp.
DET_FN double det_exp(double x)
{
    if (x != x) return x;
    if (x > 709.78) return INFINITY;
    if (x < -745.2) return 0.0;
    double k = rint(x * 1.4426950408889634);
    double r = (x - k * DET_LN2_HI) - k * DET_LN2_LO;
    double p = 1.0 + r * (1.0 + r * (1.0 / 2 + r * (1.0 / 6 + r * (1.0 / 24 + r * (1.0 / 120 + r * (1.0 / 720 +
               r * (1.0 / 5040 + r * (1.0 / 40320 + r * (1.0 / 362880 + r * (1.0 / 3628800 + r * (1.0 / 39916800 +
               r * (1.0 / 479001600 + r * (1.0 / 6227020800.0 + r * (1.0 / 87178291200.0))))))))))))));
    return ldexp(p, (int)k);
}
