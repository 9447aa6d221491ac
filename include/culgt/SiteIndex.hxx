// culgt/SiteIndex.hxx -- host-side drop-in for the reference's
// lattice/SiteIndex.hxx:45-352: the site index of a lattice in one of three
// orderings and the neighbour table the reference's apps build with it.
//
//   NO_SPLIT        n = ((t Nx + x) Ny + y) Nz + z
//   FULL_SPLIT      n/2, plus V/2 for odd t+x+y+z (the GpuPattern order)
//   TIMESLICE_SPLIT t V3 + (n3/2, plus V3/2 for odd x+y+z), n3 the index in the slice
//
// nn[(2 mu + up) V + i] is the index of the backward (up = 0) or forward
// (up = 1) neighbour of site i in direction mu, periodic.  The gfx950 kernels
// compute neighbours on the fly and never read such a table; apps that build
// one for their own use keep working.  Members the reference leaves
// unimplemented for the split orders (asserting) are implemented here.
#ifndef CULGT_SITEINDEX_HXX_
#define CULGT_SITEINDEX_HXX_

#include <stdexcept>

#include "GlobalConstants.h"

template <lat_dim_t Nd, ParityType par> class SiteIndex
{
public:
    explicit SiteIndex(const lat_coord_t size_[Nd])
    {
        V = 1;
        for (lat_dim_t i = 0; i < Nd; i++) {
            size[i] = size_[i];
            V *= size_[i];
        }
    }

    lat_index_t *nn = nullptr;  // neighbour table used by setNeighbour

    lat_index_t getLatticeIndex() const { return index; }
    void setLatticeIndex(lat_index_t i) { index = i; }
    lat_index_t getLatticeSize() const { return V; }
    lat_index_t getLatticeSizeTimeslice() const { return V / size[0]; }
    lat_coord_t getLatticeSizeDirection(lat_dim_t i) const { return size[i]; }

    // coordinate i of the current site
    lat_coord_t operator[](lat_dim_t i) const
    {
        lat_coord_t c[Nd];
        coords(index, c);
        return c[i];
    }

    // index of site (t, ...) from the index within timeslice t in the same
    // ordering restricted to the slice (NO_SPLIT: lexicographic; the split
    // orders: parity split within the slice)
    void setLatticeIndexTimeslice(lat_index_t i3, lat_coord_t t)
    {
        if (par == FULL_SPLIT) {
            lat_coord_t c[Nd];
            slice_coords(i3, c);
            c[0] = t;
            index = getLatticeIndex(c);
        } else {
            index = t * getLatticeSizeTimeslice() + i3;
        }
    }

    lat_index_t getLatticeIndexTimeslice() const
    {
        if (par == FULL_SPLIT) {
            lat_coord_t c[Nd];
            coords(index, c);
            return slice_index(c);
        }
        return index % getLatticeSizeTimeslice();
    }

    void setLatticeIndexFromParitySplitOrder(lat_index_t i)
    {
        lat_coord_t c[Nd];
        full_split_coords(i, c);
        index = getLatticeIndex(c);
    }

    void setLatticeIndexFromNonParitySplitOrder(lat_index_t n)
    {
        lat_coord_t c[Nd];
        lex_coords(n, c);
        index = getLatticeIndex(c);
    }

    void setNeighbour(lat_dim_t mu, bool up)
    {
        if (!nn) throw std::logic_error("SiteIndex::setNeighbour: no neighbour table (calculateNeighbourTable first)");
        index = nn[(2 * mu + (up ? 1 : 0)) * V + index];
    }

    void calculateNeighbourTable(lat_index_t *table)
    {
        for (lat_index_t n = 0; n < V; n++) {
            lat_coord_t c[Nd], d[Nd];
            lex_coords(n, c);
            const lat_index_t here = getLatticeIndex(c);
            for (lat_dim_t mu = 0; mu < Nd; mu++) {
                for (lat_dim_t k = 0; k < Nd; k++) d[k] = c[k];
                d[mu] = (c[mu] == 0) ? size[mu] - 1 : c[mu] - 1;
                table[(2 * mu) * V + here] = getLatticeIndex(d);
                d[mu] = (c[mu] + 1 == size[mu]) ? 0 : c[mu] + 1;
                table[(2 * mu + 1) * V + here] = getLatticeIndex(d);
            }
        }
        nn = table;
    }

    // index of the site with coordinates c in this ordering
    lat_index_t getLatticeIndex(const lat_coord_t c[Nd]) const
    {
        if (par == NO_SPLIT) return lex_index(c, 0);
        if (par == FULL_SPLIT) {
            const lat_index_t n = lex_index(c, 0);
            return n / 2 + (parity(c, 0) ? V / 2 : 0);
        }
        return c[0] * getLatticeSizeTimeslice() + slice_index(c);
    }

private:
    lat_coord_t size[Nd];
    lat_index_t V;
    lat_index_t index = 0;

    lat_index_t lex_index(const lat_coord_t c[Nd], lat_dim_t from) const
    {
        lat_index_t n = 0;
        for (lat_dim_t i = from; i < Nd; i++) n = n * size[i] + c[i];
        return n;
    }
    static int parity(const lat_coord_t c[Nd], lat_dim_t from)
    {
        int s = 0;
        for (lat_dim_t i = from; i < Nd; i++) s += c[i];
        return s & 1;
    }
    // parity-split index within the timeslice of c
    lat_index_t slice_index(const lat_coord_t c[Nd]) const
    {
        const lat_index_t V3 = getLatticeSizeTimeslice();
        return lex_index(c, 1) / 2 + (parity(c, 1) ? V3 / 2 : 0);
    }
    void lex_coords(lat_index_t n, lat_coord_t c[Nd]) const
    {
        for (lat_dim_t i = Nd - 1; i >= 0; i--) {
            c[i] = (lat_coord_t)(n % size[i]);
            n /= size[i];
        }
    }
    // coordinates from a parity-split index over the dimensions from..Nd-1
    // (2m or 2m+1 in lexicographic order, whichever has the wanted parity)
    void split_coords(lat_index_t i, lat_index_t vol, lat_dim_t from, lat_coord_t c[Nd]) const
    {
        const int odd = i >= vol / 2;
        lat_index_t n = 2 * (odd ? i - vol / 2 : i);
        for (lat_dim_t k = Nd - 1; k >= from; k--) {
            c[k] = (lat_coord_t)(n % size[k]);
            n /= size[k];
        }
        if (parity(c, from) != odd) c[Nd - 1]++;
    }
    void full_split_coords(lat_index_t i, lat_coord_t c[Nd]) const
    {
        // the parity of a 4D site includes t, so decode over all dimensions
        split_coords(i, V, 0, c);
    }
    void slice_coords(lat_index_t i3, lat_coord_t c[Nd]) const
    {
        split_coords(i3, getLatticeSizeTimeslice(), 1, c);
    }
    void coords(lat_index_t i, lat_coord_t c[Nd]) const
    {
        if (par == NO_SPLIT) {
            lex_coords(i, c);
        } else if (par == FULL_SPLIT) {
            full_split_coords(i, c);
        } else {
            const lat_index_t V3 = getLatticeSizeTimeslice();
            slice_coords(i % V3, c);
            c[0] = (lat_coord_t)(i / V3);
        }
    }
};

#endif
