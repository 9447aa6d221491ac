// ref_harness.cpp -- exposes the reference's own host-compilable templates
// through a C ABI so that the oracle can be checked against them.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target "ref") from the
// reference sources where they lie (-I <reference>/src); the output goes to
// oracle/_ref/ and is loaded only by the test-suite.
//
// What is compiled from the reference (paths relative to its src/):
//   lattice/SiteIndex.hxx, lattice/SiteCoord.hxx, lattice/access_pattern/GpuPattern.hxx,
//   lattice/Link.hxx, lattice/SU3.hxx, lattice/Matrix.hxx, lattice/Quaternion.hxx,
//   lattice/Complex.hxx, external/Random123/philox.h, external/Random123/u01.h.
// Everything under gaugefixing/ includes <cuda.h> (GlobalConstants.h:33) and is
// therefore not buildable here.  The few lines this harness needs from there --
// the FP64 accumulation ladder (GaugeFixingSubgroupStep.hxx:126-208), OrUpdate
// (OrUpdate.hxx:52-78), MicroUpdate (MicroUpdate.hxx:43-66), the apply() bodies
// (LandauKernelsSU3.hxx:216-256) and the Landau quality body
// (LandauKernelsSU3.hxx:114-191) -- are composed here from the compiled
// reference classes, so every load, store, reconstruction, subgroup
// extraction, subgroup multiplication and matrix operation below is the
// reference's own code.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "lattice/SU3.hxx"
#include "lattice/SiteIndex.hxx"
#include "lattice/SiteCoord.hxx"
#include "lattice/access_pattern/GpuPattern.hxx"
#include "lattice/Link.hxx"
#include "lattice/Matrix.hxx"
#include "lattice/Quaternion.hxx"
#include "external/Random123/philox.h"
#include "external/Random123/u01.h"

typedef SiteIndex<4, FULL_SPLIT> RSite;
typedef SiteCoord<4, FULL_SPLIT> RCoord;
typedef GpuPattern<RSite, 4, 3> RPattern;
typedef GpuPattern<RCoord, 4, 3> RPatternC;
typedef Link<RPattern, RSite, 4, 3> RLink;
typedef Link<RPatternC, RCoord, 4, 3> RLinkC;
typedef Matrix<Complex<Real>, 3> RMat;

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

static void to_size(const int dims[4], lat_coord_t size[4])
{
    for (int i = 0; i < 4; i++) size[i] = (lat_coord_t)dims[i];
}

REF_EXPORT int ref_real_size() { return (int)sizeof(Real); }

// ---------------------------------------------------------------- geometry
REF_EXPORT void ref_neighbour_table(const int dims[4], int32_t *nn)
{
    lat_coord_t size[4];
    to_size(dims, size);
    RSite s(size);
    s.calculateNeighbourTable(nn);
}

REF_EXPORT int32_t ref_site_index(const int dims[4], const int coord[4])
{
    lat_coord_t size[4], c[4];
    to_size(dims, size);
    to_size(coord, c);
    RSite s(size);
    return s.getLatticeIndex(c);
}

REF_EXPORT void ref_site_coords(const int dims[4], int32_t idx, int coord[4])
{
    lat_coord_t size[4];
    to_size(dims, size);
    RCoord s(size);
    s.setLatticeIndex(idx);
    for (int i = 0; i < 4; i++) coord[i] = s[i];
}

REF_EXPORT int32_t ref_pattern_index(const int dims[4], int32_t idx, int mu, int i, int j, int c)
{
    lat_coord_t size[4];
    to_size(dims, size);
    RSite s(size);
    s.setLatticeIndex(idx);
    return RPattern::getIndex(s, mu, i, j, c != 0);
}

// --------------------------------------------------------------------- RNG
REF_EXPORT void ref_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    philox4x32_ctr_t c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    philox4x32_key_t k = {{key[0], key[1]}};
    philox4x32_ctr_t r = philox4x32(c, k);
    for (int i = 0; i < 4; i++) out[i] = r.v[i];
}

REF_EXPORT double ref_u01_64(uint64_t i) { return u01_open_open_64_53(i); }
REF_EXPORT float ref_u01_32(uint32_t i) { return u01_open_open_32_24(i); }

// ----------------------------------------------------------- SU(3) helpers
static void from18(SU3<RMat> &u, const Real *a)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) u.set(i, j, Complex<Real>(a[2 * (3 * i + j)], a[2 * (3 * i + j) + 1]));
}
static void to18(SU3<RMat> &u, Real *a)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            a[2 * (3 * i + j)] = u.get(i, j).x;
            a[2 * (3 * i + j) + 1] = u.get(i, j).y;
        }
}

REF_EXPORT void ref_reconstruct_third_line(Real m[18])
{
    RMat mm; SU3<RMat> u(mm);
    from18(u, m); u.reconstructThirdLine(); to18(u, m);
}
REF_EXPORT void ref_project_su3(Real m[18])
{
    RMat mm; SU3<RMat> u(mm);
    from18(u, m); u.projectSU3(); to18(u, m);
}
REF_EXPORT void ref_subgroup_quaternion(const Real m[18], int i, int j, Real q[4])
{
    RMat mm; SU3<RMat> u(mm);
    from18(u, m);
    Quaternion<Real> r = u.getSubgroupQuaternion(i, j);
    for (int k = 0; k < 4; k++) q[k] = r[k];
}
REF_EXPORT void ref_left_subgroup_mult(Real m[18], int i, int j, const Real q[4])
{
    RMat mm; SU3<RMat> u(mm);
    from18(u, m);
    Quaternion<Real> r;
    for (int k = 0; k < 4; k++) r[k] = q[k];
    u.leftSubgroupMult(i, j, &r);
    to18(u, m);
}
REF_EXPORT void ref_right_subgroup_mult(Real m[18], int i, int j, const Real q[4])
{
    RMat mm; SU3<RMat> u(mm);
    from18(u, m);
    Quaternion<Real> r;
    for (int k = 0; k < 4; k++) r[k] = q[k];
    u.rightSubgroupMult(i, j, &r);
    to18(u, m);
}
REF_EXPORT void ref_det(const Real m[18], Real out[2])
{
    RMat mm; SU3<RMat> u(mm);
    from18(u, m);
    Complex<Real> d = u.det();
    out[0] = d.x; out[1] = d.y;
}

// -------------------------------------------------- composed Landau sweep
// OrUpdate::calculateUpdate (OrUpdate.hxx:52-78), MicroUpdate (MicroUpdate.hxx:43-66)
static void update(int algo, Real a[4], float param)
{
    if (algo == 0) {
        Real ai_sq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
        Real a0_sq = a[0] * a[0];
        Real b = (param * a0_sq + ai_sq) / (a0_sq + ai_sq);
        Real c = 1.0 / std::sqrt(a0_sq + b * b * ai_sq);
        a[0] *= c; a[1] *= b * c; a[2] *= b * c; a[3] *= b * c;
    } else {
        Real ai_sq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
        Real a0_sq = a[0] * a[0];
        Real b = 2. * a[0] / (a0_sq + ai_sq);
        a[0] = (a0_sq - ai_sq) / (a0_sq + ai_sq);
        a[1] *= b; a[2] *= b; a[3] *= b;
    }
}

// One parity of LKSU3::apply (LandauKernelsSU3.hxx:216-256) for algo 0 (OR)
// or 1 (micro).  The 8 threads of a site are run one after the other for each
// ladder phase, which is the order the DP build's __syncthreads enforce.
REF_EXPORT int ref_sweep_landau(int algo, const int dims[4], Real *U, int parity, float param)
{
    if (algo != 0 && algo != 1) return -1;
    lat_coord_t size[4];
    to_size(dims, size);
    RSite proto(size);
    const int V = proto.getLatticeSize();
    int32_t *nn = new int32_t[8 * (size_t)V];
    proto.calculateNeighbourTable(nn);
    for (int p = 0; p < V / 2; p++) {
        RMat loc[8];
        SU3<RMat> locU[8] = {SU3<RMat>(loc[0]), SU3<RMat>(loc[1]), SU3<RMat>(loc[2]), SU3<RMat>(loc[3]),
                             SU3<RMat>(loc[4]), SU3<RMat>(loc[5]), SU3<RMat>(loc[6]), SU3<RMat>(loc[7])};
        RSite sites[8] = {proto, proto, proto, proto, proto, proto, proto, proto};
        for (int t = 0; t < 8; t++) {
            const int updown = t / 4, mu = t % 4;
            sites[t].nn = nn;
            sites[t].setLatticeIndex(p + (parity ? V / 2 : 0));
            if (updown == 1) sites[t].setNeighbour(mu, false);
            RLink link(U, sites[t], mu);
            SU3<RLink> glob(link);
            locU[t].assignWithoutThirdLine(glob);
            locU[t].reconstructThirdLine();
        }
        const int sub[3][2] = {{0, 2}, {1, 2}, {0, 1}};
        for (int s = 0; s < 3; s++) {
            const int i = sub[s][0], j = sub[s][1];
            Real shA[4] = {0, 0, 0, 0};
            for (int t = 0; t < 8; t++) {
                Quaternion<Real> q = locU[t].getSubgroupQuaternion(i, j);
                if (t < 4) { shA[0] += q[0]; shA[1] -= q[1]; shA[2] -= q[2]; shA[3] -= q[3]; }
                else { shA[0] += q[0]; shA[1] += q[1]; shA[2] += q[2]; shA[3] += q[3]; }
            }
            update(algo, shA, param);
            for (int t = 0; t < 8; t++) {
                Quaternion<Real> q;
                for (int k = 0; k < 4; k++) q[k] = shA[k];
                if (t < 4) locU[t].leftSubgroupMult(i, j, &q);
                else { q.hermitian(); locU[t].rightSubgroupMult(i, j, &q); }
            }
        }
        for (int t = 0; t < 8; t++) {
            RLink link(U, sites[t], t % 4);
            SU3<RLink> glob(link);
            glob.assignWithoutThirdLine(locU[t]);
        }
    }
    delete[] nn;
    return 0;
}

// CommonKernelsSU3::projectSU3 kernel body (CommonKernelsSU3.hxx:71-90),
// run on the global storage through Link exactly as the kernel does.
REF_EXPORT void ref_project_all(const int dims[4], Real *U)
{
    lat_coord_t size[4];
    to_size(dims, size);
    RSite s(size);
    const int V = s.getLatticeSize();
    for (int site = 0; site < V; site++) {
        s.setLatticeIndex(site);
        for (int mu = 0; mu < 4; mu++) {
            RLink linkUp(U, s, mu);
            SU3<RLink> globUp(linkUp);
            globUp.projectSU3();
        }
    }
}

// Landau quality per site (LandauKernelsSU3.hxx:114-191) composed from the
// compiled SU3 / Matrix operations, and the serial sum
// (GaugeFixingStats.hxx:655-668) with the Landau prefactors.
REF_EXPORT void ref_quality_landau(const int dims[4], Real *U, double *gff, double *theta)
{
    lat_coord_t size[4];
    to_size(dims, size);
    RCoord s(size);
    const int V = s.getLatticeSize();
    double sg = 0, sa = 0;
    for (int site = 0; site < V; site++) {
        RMat locMatSum;
        SU3<RMat> Sum(locMatSum);
        Sum.zero();
        for (int mu = 0; mu < 4; mu++) {
            s.setLatticeIndex(site);
            RMat locMat;
            SU3<RMat> temp(locMat);
            RLinkC linkUp(U, s, mu);
            SU3<RLinkC> globUp(linkUp);
            temp.assignWithoutThirdLine(globUp);
            temp.reconstructThirdLine();
            Sum += temp;
            s.setNeighbour(mu, false);
            RLinkC linkDw(U, s, mu);
            SU3<RLinkC> globDw(linkDw);
            temp.assignWithoutThirdLine(globDw);
            temp.reconstructThirdLine();
            Sum -= temp;
        }
        Sum -= Sum.trace() / Real(3.);
        RMat locMatSumHerm;
        SU3<RMat> SumHerm(locMatSumHerm);
        SumHerm = Sum;
        SumHerm.hermitian();
        Sum -= SumHerm;
        double prec = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) prec += Sum.get(i, j).abs_squared();
        s.setLatticeIndex(site);
        double result = 0;
        RMat locTemp;
        SU3<RMat> temp(locTemp);
        for (int mu = 0; mu < 4; mu++) {
            RLinkC linkUp(U, s, mu);
            SU3<RLinkC> globUp(linkUp);
            temp.assignWithoutThirdLine(globUp);
            temp.reconstructThirdLine();
            result += temp.trace().x;
        }
        sg += result;
        sa += prec;
    }
    *gff = sg * (1. / 12.) / (double)V;
    *theta = sa * (1. / 3.) / (double)V;
}
