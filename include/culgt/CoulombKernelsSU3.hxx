// culgt/CoulombKernelsSU3.hxx -- drop-in for the reference's
// gaugefixing/CoulombKernelsSU3.hxx:53-101: Coulomb gauge fixing of one
// timeslice.  UtUp holds slice t and UtDw slice t-1, both in the GpuPattern
// layout over {1, Nx, Ny, Nz}; the spatial links of slice t and the time
// links of both slices are transformed (CoulombKernelsSU3.hxx:208-250).
#ifndef CULGT_COULOMBKERNELSSU3_HXX_
#define CULGT_COULOMBKERNELSSU3_HXX_

#include "GlobalConstants.h"

class CoulombKernelsSU3
{
public:
    static const int gauge = GF_COULOMB;
    static const int Ndim = 4;
    static const int Nc = 3;

    static void initCacheConfig() {}

    static void generateGaugeQualityPerSite(int a, int b, Real *U, double *dGff, double *dA)
    {
        culgt_dropin::quality_per_site(HOST_CONSTANTS::SIZE_TIMESLICE, GF_COULOMB, a, b, U, dGff, dA);
    }
    static double getGaugeQualityPrefactorA() { return 1. / (double)Nc; }
    static double getGaugeQualityPrefactorGff() { return 1. / (double)(Nc * (Ndim - 1)); }

    static void restoreThirdLine(int, int, Real *U, lat_index_t *)
    {
        culgt_dropin::project(HOST_CONSTANTS::SIZE_TIMESLICE, U);
    }
    static void randomTrafo(int a, int b, Real *UtUp, Real *UtDw, lat_index_t *, bool parity, int rngSeed, int rngCounter)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE_TIMESLICE, GF_COULOMB, GF_RANDOM, a, b, UtUp, UtDw, parity, 0.f, rngSeed,
                            rngCounter);
    }
    static void orStep(int a, int b, Real *UtUp, Real *UtDw, lat_index_t *, bool parity, float orParameter)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE_TIMESLICE, GF_COULOMB, GF_OR, a, b, UtUp, UtDw, parity, orParameter, 0,
                            0);
    }
    static void microStep(int a, int b, Real *UtUp, Real *UtDw, lat_index_t *, bool parity)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE_TIMESLICE, GF_COULOMB, GF_MICRO, a, b, UtUp, UtDw, parity, 0.f, 0, 0);
    }
    static void saStep(int a, int b, Real *UtUp, Real *UtDw, lat_index_t *, bool parity, float temperature, int rngSeed,
                       int rngCounter)
    {
        culgt_dropin::sweep(HOST_CONSTANTS::SIZE_TIMESLICE, GF_COULOMB, GF_SA, a, b, UtUp, UtDw, parity, temperature,
                            rngSeed, rngCounter);
    }
};

#endif
