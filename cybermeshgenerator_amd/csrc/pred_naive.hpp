// pred_naive.hpp — the reference's naive FP64 orient3d and insphere determinants, in its
// operation order (device code; every TU that includes it is built without FMA
// contraction, as the reference's x86-64 build evaluates them).
#pragma once

#include <hip/hip_runtime.h>

namespace cmg::hip {

__device__ __forceinline__ double o3_naive(const double* v, double* perm) {
    // predicates.cpp:2201-2222 with pa = v[0..2], pb = v[3..5], pc = v[6..8], pd = v[9..11]
    double adx = v[0] - v[9], ady = v[1] - v[10], adz = v[2] - v[11];
    double bdx = v[3] - v[9], bdy = v[4] - v[10], bdz = v[5] - v[11];
    double cdx = v[6] - v[9], cdy = v[7] - v[10], cdz = v[8] - v[11];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    if (perm)
        *perm = (__builtin_fabs(bdxcdy) + __builtin_fabs(cdxbdy)) * __builtin_fabs(adz) +
                (__builtin_fabs(cdxady) + __builtin_fabs(adxcdy)) * __builtin_fabs(bdz) +
                (__builtin_fabs(adxbdy) + __builtin_fabs(bdxady)) * __builtin_fabs(cdz);
    return det;
}

__device__ __forceinline__ double is_naive(const double* v, double* perm) {
    // predicates.cpp:4063-4106 with pa..pe = v[0..14]
    double aex = v[0] - v[12], bex = v[3] - v[12], cex = v[6] - v[12], dex = v[9] - v[12];
    double aey = v[1] - v[13], bey = v[4] - v[13], cey = v[7] - v[13], dey = v[10] - v[13];
    double aez = v[2] - v[14], bez = v[5] - v[14], cez = v[8] - v[14], dez = v[11] - v[14];
    double aexbey = aex * bey, bexaey = bex * aey;
    double ab = aexbey - bexaey;
    double bexcey = bex * cey, cexbey = cex * bey;
    double bc = bexcey - cexbey;
    double cexdey = cex * dey, dexcey = dex * cey;
    double cd = cexdey - dexcey;
    double dexaey = dex * aey, aexdey = aex * dey;
    double da = dexaey - aexdey;
    double aexcey = aex * cey, cexaey = cex * aey;
    double ac = aexcey - cexaey;
    double bexdey = bex * dey, dexbey = dex * bey;
    double bd = bexdey - dexbey;
    double abc = aez * bc - bez * ac + cez * ab;
    double bcd = bez * cd - cez * bd + dez * bc;
    double cda = cez * da + dez * ac + aez * cd;
    double dab = dez * ab + aez * bd + bez * da;
    double alift = aex * aex + aey * aey + aez * aez;
    double blift = bex * bex + bey * bey + bez * bez;
    double clift = cex * cex + cey * cey + cez * cez;
    double dlift = dex * dex + dey * dey + dez * dez;
    double det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
    if (perm) {
        // predicates.cpp:4122-4156
        double aezp = __builtin_fabs(aez), bezp = __builtin_fabs(bez),
               cezp = __builtin_fabs(cez), dezp = __builtin_fabs(dez);
        double aexbeyp = __builtin_fabs(aexbey), bexaeyp = __builtin_fabs(bexaey);
        double bexceyp = __builtin_fabs(bexcey), cexbeyp = __builtin_fabs(cexbey);
        double cexdeyp = __builtin_fabs(cexdey), dexceyp = __builtin_fabs(dexcey);
        double dexaeyp = __builtin_fabs(dexaey), aexdeyp = __builtin_fabs(aexdey);
        double aexceyp = __builtin_fabs(aexcey), cexaeyp = __builtin_fabs(cexaey);
        double bexdeyp = __builtin_fabs(bexdey), dexbeyp = __builtin_fabs(dexbey);
        *perm = ((cexdeyp + dexceyp) * bezp + (dexbeyp + bexdeyp) * cezp +
                 (bexceyp + cexbeyp) * dezp) * alift +
                ((dexaeyp + aexdeyp) * cezp + (aexceyp + cexaeyp) * dezp +
                 (cexdeyp + dexceyp) * aezp) * blift +
                ((aexbeyp + bexaeyp) * dezp + (bexdeyp + dexbeyp) * aezp +
                 (dexaeyp + aexdeyp) * bezp) * clift +
                ((bexceyp + cexbeyp) * aezp + (cexaeyp + aexceyp) * bezp +
                 (aexbeyp + bexaeyp) * cezp) * dlift;
    }
    return det;
}

}  // namespace cmg::hip
