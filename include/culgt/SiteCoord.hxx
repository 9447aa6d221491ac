// culgt/SiteCoord.hxx -- host-side drop-in for the reference's
// lattice/SiteCoord.hxx:40-355: a lattice site held as coordinates
// (t, x, y, z) -- public `site[]` and `size[]`, as the reference's file code
// reads and writes them -- with its index in one of three orderings:
//
//   NO_SPLIT        n = ((t Nx + x) Ny + y) Nz + z
//   FULL_SPLIT      n/2, plus V/2 for odd t+x+y+z (the GpuPattern order)
//   TIMESLICE_SPLIT t V3 + (n3/2, plus V3/2 for odd x+y+z), n3 the index in the slice
//
// getLatticeIndexTimeslice() is the parity-split index within the site's
// timeslice for both split orderings (SiteCoord.hxx:289-330).
#ifndef CULGT_SITECOORD_HXX_
#define CULGT_SITECOORD_HXX_

#include "GlobalConstants.h"

template <lat_dim_t Nd, ParityType par> class SiteCoord
{
public:
    static const lat_dim_t Ndim = Nd;
    lat_coord_t size[Nd];
    lat_coord_t site[Nd];

    explicit SiteCoord(const lat_coord_t size_[Nd])
    {
        for (int i = 0; i < Nd; i++) {
            size[i] = size_[i];
            site[i] = 0;
        }
    }

    lat_coord_t &operator[](lat_dim_t i) { return site[i]; }
    lat_coord_t getLatticeSizeDirection(lat_dim_t i) const { return size[i]; }

    lat_index_t getLatticeSize() const
    {
        lat_index_t v = 1;
        for (int i = 0; i < Nd; i++) v *= size[i];
        return v;
    }
    lat_index_t getLatticeSizeTimeslice() const { return getLatticeSize() / size[0]; }

    lat_index_t getLatticeIndex() const
    {
        if (par == NO_SPLIT) return lex(0);
        if (par == FULL_SPLIT) return lex(0) / 2 + (parity(0) ? getLatticeSize() / 2 : 0);
        return site[0] * getLatticeSizeTimeslice() + slice_index();
    }

    lat_index_t getLatticeIndexTimeslice() const
    {
        if (par == NO_SPLIT) return lex(1);
        return slice_index();
    }

    void setLatticeIndex(lat_index_t i)
    {
        if (par == NO_SPLIT) {
            from_lex(i, 0);
        } else if (par == FULL_SPLIT) {
            from_split(i, getLatticeSize(), 0);
        } else {
            const lat_index_t V3 = getLatticeSizeTimeslice();
            from_split(i % V3, V3, 1);
            site[0] = (lat_coord_t)(i / V3);
        }
    }

    void setLatticeIndexFromNonParitySplitOrder(lat_index_t n) { from_lex(n, 0); }
    void setLatticeIndexFromParitySplitOrder(lat_index_t i) { from_split(i, getLatticeSize(), 0); }

    // the neighbour in direction mu, periodic
    void setNeighbour(lat_dim_t mu, bool up)
    {
        if (up) site[mu] = (lat_coord_t)((site[mu] + 1 == size[mu]) ? 0 : site[mu] + 1);
        else site[mu] = (lat_coord_t)((site[mu] == 0) ? size[mu] - 1 : site[mu] - 1);
    }

private:
    lat_index_t lex(int from) const
    {
        lat_index_t n = 0;
        for (int i = from; i < Nd; i++) n = n * size[i] + site[i];
        return n;
    }
    int parity(int from) const
    {
        int s = 0;
        for (int i = from; i < Nd; i++) s += site[i];
        return s & 1;
    }
    lat_index_t slice_index() const
    {
        return lex(1) / 2 + (parity(1) ? getLatticeSizeTimeslice() / 2 : 0);
    }
    void from_lex(lat_index_t n, int from)
    {
        for (int i = Nd - 1; i >= from; i--) {
            site[i] = (lat_coord_t)(n % size[i]);
            n /= size[i];
        }
    }
    // coordinates from a parity-split index over dimensions from..Nd-1: the
    // lexicographic index 2m or 2m+1, whichever has the wanted parity
    void from_split(lat_index_t i, lat_index_t vol, int from)
    {
        const int odd = i >= vol / 2;
        from_lex(2 * (odd ? i - vol / 2 : i), from);
        if (parity(from) != odd) site[Nd - 1]++;
    }
};

#endif
