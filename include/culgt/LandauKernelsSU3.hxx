// culgt/LandauKernelsSU3.hxx -- drop-in for the reference's
// gaugefixing/LandauKernelsSU3.hxx:55-105 (static kernel wrappers of Landau
// gauge fixing).  Same names, arguments and meaning; a, b are the reference's
// launch dimensions (V/2/NSB blocks of 8*NSB threads for sweeps, V/NSB x NSB
// for the quality) and are checked; nnt is accepted and not read (neighbours
// are computed on the fly).  Failures throw std::runtime_error.
#ifndef CULGT_LANDAUKERNELSSU3_HXX_
#define CULGT_LANDAUKERNELSSU3_HXX_

#include "GlobalConstants.h"

class LandauKernelsSU3
{
public:
    static const int gauge = GF_LANDAU;
    static const int Ndim = 4;
    static const int Nc = 3;

    static void initCacheConfig() {}

    static void generateGaugeQualityPerSite(int a, int b, Real *U, double *dGff, double *dA)
    {
        culgt_dropin::quality_per_site(HOST_CONSTANTS::SIZE, GF_LANDAU, a, b, U, dGff, dA);
    }
    static double getGaugeQualityPrefactorA() { return 1. / (double)Nc; }
    static double getGaugeQualityPrefactorGff() { return 1. / (double)(Nc * Ndim); }

    // never launched by the reference apps; reprojects every link
    static void restoreThirdLine(int, int, Real *U, lat_index_t *) { culgt_dropin::project(HOST_CONSTANTS::SIZE, U); }

    static void randomTrafo(int a, int b, Real *U, lat_index_t *, bool parity, int rngSeed, int rngCounter)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE, GF_LANDAU, GF_RANDOM, a, b, U, U, parity, 0.f, rngSeed, rngCounter);
    }
    static void orStep(int a, int b, Real *U, lat_index_t *, bool parity, float orParameter)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE, GF_LANDAU, GF_OR, a, b, U, U, parity, orParameter, 0, 0);
    }
    static void microStep(int a, int b, Real *U, lat_index_t *, bool parity)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE, GF_LANDAU, GF_MICRO, a, b, U, U, parity, 0.f, 0, 0);
    }
    static void saStep(int a, int b, Real *U, lat_index_t *, bool parity, float temperature, int rngSeed, int rngCounter)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE, GF_LANDAU, GF_SA, a, b, U, U, parity, temperature, rngSeed,
                            rngCounter);
    }
};

#endif
