// culgt/GaugeFixingStats.hxx -- drop-in for the reference's
// gaugefixing/GaugeFixingStats.hxx:55-232: per-site quality of GType
// (Landau, Coulomb, MAG or U1xU1 kernels), a reduction and the prefactors;
// getCurrentGff / getCurrentA after generateGaugeQuality.  The reduction is a
// fixed-order tree on the device instead of the reference's single-thread loop.
#ifndef CULGT_GAUGEFIXINGSTATS_HXX_
#define CULGT_GAUGEFIXINGSTATS_HXX_

#include "GlobalConstants.h"

template <int Ndim, int Nc, class GType, StoppingCrit ma> class GaugeFixingStats
{
public:
    GaugeFixingStats(Real *U, const lat_coord_t *size) : U(U), reqPrec(0) { init(size); }
    GaugeFixingStats(Real *U, const lat_coord_t *size, double prec) : U(U), reqPrec(prec) { init(size); }
    ~GaugeFixingStats() { gf_stats_destroy(st); }
    GaugeFixingStats(const GaugeFixingStats &) = delete;
    GaugeFixingStats &operator=(const GaugeFixingStats &) = delete;

    void setGaugePrecision(double prec) { reqPrec = prec; }
    double getGaugePrecision() { return reqPrec; }
    double getCurrentGff() { return currentGff; }
    double getCurrentA() { return currentA; }
    void setPointer(Real *U) { this->U = U; }

    void generateGaugeQuality()
    {
        culgt_dropin::check(gf_stats_quality(st, &lat, U, &currentGff, &currentA, nullptr));
    }

private:
    void init(const lat_coord_t *size)
    {
        lat = culgt_dropin::lattice(size);
        culgt_dropin::check(gf_stats_create(lat.dims, GType::gauge, ma == MAX ? GF_CRIT_MAX : GF_CRIT_AVERAGE, &st));
    }
    Real *U;
    gf_lattice lat;
    gf_stats *st = nullptr;
    double currentGff = 0, currentA = 0, reqPrec;
};

#endif
