// Prints what include/culgt/SiteIndex.hxx computes for one lattice, for the
// CPU test tests/test_culgt_headers.py: the neighbour table of each ordering
// and, for every site, the index round trips through the coordinates.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "culgt/SiteIndex.hxx"

template <ParityType P> void dump(const lat_coord_t size[4], const char *name)
{
    SiteIndex<4, P> s(size);
    const lat_index_t V = s.getLatticeSize();
    std::vector<lat_index_t> nn(8 * (size_t)V);
    s.calculateNeighbourTable(nn.data());
    printf("%s nn", name);
    for (lat_index_t k = 0; k < 8 * V; k++) printf(" %d", nn[k]);
    printf("\n%s coords", name);
    for (lat_index_t i = 0; i < V; i++) {
        s.setLatticeIndex(i);
        lat_coord_t c[4] = {s[0], s[1], s[2], s[3]};
        printf(" %d %d %d %d %d", c[0], c[1], c[2], c[3], s.getLatticeIndex(c));
    }
    printf("\n%s slice", name);
    for (lat_index_t i = 0; i < V; i++) {
        s.setLatticeIndex(i);
        const lat_index_t i3 = s.getLatticeIndexTimeslice();
        const lat_coord_t t = s[0];
        s.setLatticeIndexTimeslice(i3, t);
        printf(" %d %d", i3, s.getLatticeIndex());
    }
    printf("\n%s step", name);
    for (lat_index_t i = 0; i < V; i++) {
        s.setLatticeIndex(i);
        s.setNeighbour(3, true);
        s.setNeighbour(0, false);
        printf(" %d", s.getLatticeIndex());
    }
    printf("\n");
}

int main(int argc, char **argv)
{
    if (argc != 5) return 2;
    const lat_coord_t size[4] = {(lat_coord_t)atoi(argv[1]), (lat_coord_t)atoi(argv[2]), (lat_coord_t)atoi(argv[3]),
                                 (lat_coord_t)atoi(argv[4])};
    dump<NO_SPLIT>(size, "NO_SPLIT");
    dump<FULL_SPLIT>(size, "FULL_SPLIT");
    dump<TIMESLICE_SPLIT>(size, "TIMESLICE_SPLIT");
    SiteIndex<4, FULL_SPLIT> s(size);
    for (lat_index_t n = 0; n < s.getLatticeSize(); n++) {
        s.setLatticeIndexFromNonParitySplitOrder(n);
        const lat_index_t i = s.getLatticeIndex();
        s.setLatticeIndexFromParitySplitOrder(i);
        if (s.getLatticeIndex() != i) return 3;
    }
    return 0;
}
