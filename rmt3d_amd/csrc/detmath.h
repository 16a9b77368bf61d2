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
