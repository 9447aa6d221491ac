// culgt/GpuPattern.hxx -- host-side drop-in for the reference's
// lattice/access_pattern/{GpuPattern,GpuPatternTimeslice,StandardPattern}.hxx:
// where entry (i, j), real part c = 0 or imaginary part c = 1, of the link in
// direction mu at a site lies in an array of all 18 reals per link.
//
//   GpuPattern            idx + V (c + 2 (j + 3 (i + 3 mu))), idx the site's
//                         index in its own ordering (GpuPattern.hxx:53-55)
//   GpuPatternTimeslice   72 V3 t + i3 + V3 (c + 2 (j + 3 (i + 3 mu))), i3 the
//                         parity-split index within timeslice t
//                         (GpuPatternTimeslice.hxx:57-61)
//   StandardPattern       c + 2 (j + 3 (i + 3 (mu + 4 n))), n the site's index
//                         (StandardPattern.hxx:53-70); the "unique" order of the
//                         configuration files
//
// getIndexByUnique maps a StandardPattern index of the NO_SPLIT lexicographic
// order to the pattern's index, as LinkFile does when it loads and saves.
#ifndef CULGT_GPUPATTERN_HXX_
#define CULGT_GPUPATTERN_HXX_

#include "SiteCoord.hxx"

namespace culgt_dropin {
struct UniqueEntry {
    int c, j, i, mu;
    lat_array_index_t n;  // NO_SPLIT site index
};
inline UniqueEntry split_unique(lat_array_index_t u, int Ndim, int Nc)
{
    UniqueEntry e;
    e.c = (int)(u % 2);
    u /= 2;
    e.j = (int)(u % Nc);
    u /= Nc;
    e.i = (int)(u % Nc);
    u /= Nc;
    e.mu = (int)(u % Ndim);
    e.n = u / Ndim;
    return e;
}
}  // namespace culgt_dropin

template <class Site, lat_dim_t T_Ndim, lat_group_dim_t T_Nc> class GpuPattern
{
public:
    static const lat_group_dim_t Nc = T_Nc;
    static const lat_index_t Ndim = T_Ndim;
    static const int memory_pattern = 0;  // GF_PATTERN_GPU

    static lat_index_t getSiteIndex(Site s) { return s.getLatticeIndex(); }
    static lat_array_index_t getLinkIndex(Site s, lat_dim_t mu)
    {
        return s.getLatticeIndex() + mu * T_Nc * T_Nc * 2 * s.getLatticeSize();
    }
    static lat_array_index_t getIndex(Site s, lat_dim_t mu, lat_group_dim_t i, lat_group_dim_t j, bool c)
    {
        return s.getLatticeIndex() + s.getLatticeSize() * ((int)c + 2 * (j + T_Nc * (i + T_Nc * mu)));
    }
    static lat_array_index_t getIndexByUnique(lat_array_index_t u, lat_coord_t size[T_Ndim])
    {
        const culgt_dropin::UniqueEntry e = culgt_dropin::split_unique(u, T_Ndim, T_Nc);
        Site s(size);
        s.setLatticeIndexFromNonParitySplitOrder(e.n);
        return s.getLatticeIndex() + s.getLatticeSize() * (e.c + 2 * (e.j + T_Nc * (e.i + T_Nc * e.mu)));
    }
};

template <class Site, lat_dim_t T_Ndim, lat_group_dim_t T_Nc> class GpuPatternTimeslice
{
public:
    static const lat_group_dim_t Nc = T_Nc;
    static const lat_index_t Ndim = T_Ndim;
    static const int memory_pattern = 1;  // GF_PATTERN_GPU_TIMESLICE

    static lat_index_t getSiteIndex(Site s)
    {
        return s.getLatticeIndexTimeslice() + s[0] * s.getLatticeSizeTimeslice() * T_Ndim * T_Nc * T_Nc * 2;
    }
    static lat_array_index_t getLinkIndex(Site s, lat_dim_t mu)
    {
        const lat_array_index_t V3 = s.getLatticeSizeTimeslice();
        return s.getLatticeIndexTimeslice() + mu * T_Nc * T_Nc * 2 * V3 + s[0] * V3 * T_Ndim * T_Nc * T_Nc * 2;
    }
    static lat_array_index_t getIndex(Site s, lat_dim_t mu, lat_group_dim_t i, lat_group_dim_t j, bool c)
    {
        const lat_array_index_t V3 = s.getLatticeSizeTimeslice();
        return s.getLatticeIndexTimeslice() + V3 * ((int)c + 2 * (j + T_Nc * (i + T_Nc * mu))) +
               s[0] * V3 * T_Ndim * T_Nc * T_Nc * 2;
    }
    static lat_array_index_t getIndexByUnique(lat_array_index_t u, lat_coord_t size[T_Ndim])
    {
        const culgt_dropin::UniqueEntry e = culgt_dropin::split_unique(u, T_Ndim, T_Nc);
        Site s(size);
        s.setLatticeIndexFromNonParitySplitOrder(e.n);
        const lat_array_index_t V3 = s.getLatticeSizeTimeslice();
        return s.getLatticeIndexTimeslice() + V3 * (e.c + 2 * (e.j + T_Nc * (e.i + T_Nc * e.mu))) +
               s[0] * V3 * T_Ndim * T_Nc * T_Nc * 2;
    }
};

template <class Site, lat_dim_t T_Ndim, lat_group_dim_t T_Nc> class StandardPattern
{
public:
    static const lat_group_dim_t Nc = T_Nc;
    static const lat_index_t Ndim = T_Ndim;
    static const bool is_standard = true;

    static lat_index_t getSiteIndex(Site s) { return s.getLatticeIndex() * T_Ndim * T_Nc * T_Nc * 2; }
    static lat_array_index_t getLinkIndex(Site s, lat_dim_t mu)
    {
        return (mu + T_Ndim * s.getLatticeIndex()) * T_Nc * T_Nc * 2;
    }
    static lat_array_index_t getIndex(Site s, lat_dim_t mu, lat_group_dim_t i, lat_group_dim_t j, bool c)
    {
        return (int)c + 2 * (j + T_Nc * (i + T_Nc * (mu + T_Ndim * s.getLatticeIndex())));
    }
    static lat_array_index_t getUniqueIndex(Site s, lat_dim_t mu, lat_group_dim_t i, lat_group_dim_t j, bool c)
    {
        return getIndex(s, mu, i, j, c);
    }
    static lat_array_index_t getUniqueIndex(lat_array_index_t index, lat_coord_t *) { return index; }
    static lat_array_index_t getIndexByUnique(lat_array_index_t u, lat_coord_t *) { return u; }
};

#endif
