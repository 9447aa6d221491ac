// culgt/GlobalConstants.h -- drop-in for the reference's gaugefixing/GlobalConstants.h
// (GlobalConstants.h:35-91) and lattice/datatype/{lattice_typedefs,datatypes}.h.
//
// The lattice extents stay compile-time constants set by the same macros as
// in the reference: -D_X_=<Nx> -D_T_=<Nt> [-D_Y_ -D_Z_] [-D_NSB_] and
// -DDOUBLEPRECISION selects Real = double.  Everything forwards to the C ABI
// of libsu3gf.so (include/su3gf.h); link with -lsu3gf.
#ifndef CULGT_GLOBALCONSTANTS_H_
#define CULGT_GLOBALCONSTANTS_H_

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../su3gf.h"

// lattice/datatype/lattice_typedefs.h
typedef short lat_dim_t;
typedef short lat_group_dim_t;
typedef short lat_coord_t;
typedef int lat_index_t;
typedef int lat_array_index_t;
enum GaugeType { LANDAU, COULOMB, MAG, U1xU1 };
enum StoppingCrit { MAX, AVERAGE };
enum ParityType { NO_SPLIT, FULL_SPLIT, TIMESLICE_SPLIT };

// lattice/datatype/datatypes.h
#ifdef DOUBLEPRECISION
typedef double Real;
#else
typedef float Real;
#endif

#ifndef _X_
#error "Define _X_ (the lattice size in x-direction)"
#endif
#ifndef _T_
#error "Define _T_ (the lattice size in t-direction)"
#endif
const lat_coord_t Nx = _X_;
#ifdef _Y_
const lat_coord_t Ny = _Y_;
#else
const lat_coord_t Ny = _X_;
#endif
#ifdef _Z_
const lat_coord_t Nz = _Z_;
#else
const lat_coord_t Nz = _X_;
#endif
const lat_coord_t Nt = _T_;
#ifdef _NSB_
const int NSB = _NSB_;
#else
const int NSB = 32;
#endif

namespace HOST_CONSTANTS {
const lat_coord_t SIZE[4] = {Nt, Nx, Ny, Nz};
const lat_coord_t SIZE_TIMESLICE[4] = {1, Nx, Ny, Nz};
// The reference returns device copies of SIZE / SIZE_TIMESLICE that the apps
// only hand back to CommonKernelsSU3; here they are tokens naming the lattice.
inline lat_coord_t *getPtrToDeviceSize() { return const_cast<lat_coord_t *>(SIZE); }
inline lat_coord_t *getPtrToDeviceSizeTimeslice() { return const_cast<lat_coord_t *>(SIZE_TIMESLICE); }
inline void init() {}
}  // namespace HOST_CONSTANTS

namespace culgt_dropin {

inline void check(gf_status s)
{
    if (s != GF_OK) throw std::runtime_error(std::string("su3gf: ") + gf_last_error());
}

inline gf_lattice lattice(const lat_coord_t *size)
{
    gf_lattice l;
    for (int i = 0; i < 4; i++) l.dims[i] = size[i];
    l.dtype = sizeof(Real) == sizeof(double) ? GF_F64 : GF_F32;
    l.layout = GF_LAYOUT_PATTERN18;
    return l;
}

inline int64_t volume(const lat_coord_t *size)
{
    return (int64_t)size[0] * size[1] * size[2] * size[3];
}

// The reference launches V/2/NSB blocks of 8*NSB threads for a sweep
// (LandauGaugeFixingSU3_4D.cu:124-125); NSB enters the RNG key.
inline int sweep_nsb(int a, int b, const lat_coord_t *size)
{
    if (b <= 0 || b % 8) throw std::invalid_argument("su3gf: threads per block must be a positive multiple of 8");
    const int nsb = b / 8;
    if ((int64_t)a * nsb != volume(size) / 2) throw std::invalid_argument("su3gf: launch does not cover V/2 sites");
    return nsb;
}

// The reference's single-precision build with DPUPDATES=true
// (gaugefixing/apps/Makefile:46-48) computes the OR, microcanonical and
// heatbath updates in double (algorithms/OrUpdate.hxx:54-64,
// MicroUpdate.hxx:45-54, SaUpdate.hxx:54-98): the same macros select the
// library's double-precision update rules.
inline int update_rule(int algo)
{
    if (sizeof(Real) == sizeof(double)) return algo;
#ifdef USE_DP_ORUPDATE
    if (algo == GF_OR) return algo | GF_UPDATE_DP;
#endif
#ifdef USE_DP_MICROUPDATE
    if (algo == GF_MICRO) return algo | GF_UPDATE_DP;
#endif
#ifdef USE_DP_SAUPDATE
    if (algo == GF_SA) return algo | GF_UPDATE_DP;
#endif
    return algo;
}

inline void sweep(const lat_coord_t *size, int gauge, int algo, int a, int b, Real *U, Real *Udw, bool parity,
                  float param, int seed, unsigned counter)
{
    const gf_lattice l = lattice(size);
    check(gf_sweep(&l, U, Udw, parity ? 1 : 0, gauge, update_rule(algo), param, seed, counter,
                   sweep_nsb(a, b, size), nullptr));
}

inline void quality_per_site(const lat_coord_t *size, int gauge, int a, int b, Real *U, double *dGff, double *dA)
{
    if ((int64_t)a * b != volume(size)) throw std::invalid_argument("su3gf: launch does not cover V sites");
    const gf_lattice l = lattice(size);
    check(gf_quality_per_site(&l, gauge, U, dGff, dA, nullptr));
}

inline void project(const lat_coord_t *size, Real *U)
{
    const gf_lattice l = lattice(size);
    check(gf_project(&l, U, nullptr));
}

}  // namespace culgt_dropin

#endif  // CULGT_GLOBALCONSTANTS_H_
