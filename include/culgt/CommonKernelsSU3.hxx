// culgt/CommonKernelsSU3.hxx -- drop-in for the reference's
// gaugefixing/CommonKernelsSU3.hxx:40-62: projectSU3 (reunitarize every link
// and write its third row) and setHot (random links).  ptrToDeviceSize is
// HOST_CONSTANTS::getPtrToDeviceSize() or getPtrToDeviceSizeTimeslice(), as
// in the reference apps; a*b must be the number of sites of that lattice.
#ifndef CULGT_COMMONKERNELSSU3_HXX_
#define CULGT_COMMONKERNELSSU3_HXX_

#include "GlobalConstants.h"

class CommonKernelsSU3
{
public:
    static void initCacheConfig() {}

    static void projectSU3(int a, int b, Real *U, lat_coord_t *ptrToDeviceSize)
    {
        const lat_coord_t *size = lattice_of(ptrToDeviceSize);
        if ((int64_t)a * b != culgt_dropin::volume(size)) throw std::invalid_argument("su3gf: launch does not cover V sites");
        culgt_dropin::project(size, U);
    }

    static void setHot(int a, int b, Real *U, lat_coord_t *ptrToDeviceSize, int rngSeed, int rngCounter)
    {
        const lat_coord_t *size = lattice_of(ptrToDeviceSize);
        if ((int64_t)a * b != culgt_dropin::volume(size)) throw std::invalid_argument("su3gf: launch does not cover V sites");
        const gf_lattice l = culgt_dropin::lattice(size);
        culgt_dropin::check(gf_set_hot(&l, U, rngSeed, (uint32_t)rngCounter, nullptr));
    }

private:
    static const lat_coord_t *lattice_of(const lat_coord_t *p)
    {
        if (p == HOST_CONSTANTS::SIZE_TIMESLICE) return HOST_CONSTANTS::SIZE_TIMESLICE;
        if (p == HOST_CONSTANTS::SIZE) return HOST_CONSTANTS::SIZE;
        throw std::invalid_argument("su3gf: size pointer is not HOST_CONSTANTS::getPtrToDeviceSize[Timeslice]()");
    }
};

#endif
