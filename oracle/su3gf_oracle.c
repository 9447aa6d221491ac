/*
 * su3gf_oracle.c -- CPU restatement of cuLGT's SU(3) gauge-fixing sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Only the test-suite, the smoke check and the CPU-baseline
 * leg of the benchmark may load it; the product library never calls into it.
 *
 * It restates, operation by operation, the FP64 ("DOUBLEPRECISION") code path
 * of the reference (paths relative to the reference's src/ directory):
 *   - geometry        lattice/SiteIndex.hxx:256-313, lattice/SiteCoord.hxx:104-196,
 *                     lattice/access_pattern/GpuPattern.hxx:53-55
 *   - RNG             external/Random123/philox.h:206-326, u01.h:111,129,
 *                     lattice/rng/PhiloxWrapper.hxx:93-140
 *   - SU(3) algebra   lattice/SU3.hxx:193-241 (subgroup quaternion / matrix),
 *                     :536-583 (projectSU3), :639-676 (reconstructThirdLine),
 *                     :682-714 (subgroup multiplications), :762-771 (order);
 *                     lattice/Quaternion.hxx:100-132,174-180; lattice/Complex.hxx
 *   - update rules    gaugefixing/algorithms/{Or,Micro,Sa,Sr,Random}Update.hxx
 *   - subgroup step   gaugefixing/GaugeFixingSubgroupStep.hxx:74-705 (the
 *                     deterministic DP ladder: up mu=0..3, then down mu=0..3)
 *   - apply kernels   gaugefixing/{Landau,Coulomb,MAG,U1xU1}KernelsSU3.hxx apply()
 *   - quality         gaugefixing/{Landau,Coulomb,MAG}KernelsSU3.hxx
 *                     generateGaugeQualityPerSite, GaugeFixingStats.hxx:186-232,655-675
 *   - projection      gaugefixing/CommonKernelsSU3.hxx:71-90
 *   - driver loop     gaugefixing/apps/LandauGaugeFixingSU3_4D.cu:244-262
 *
 * The arithmetic is written in the same order as the reference and this file
 * is compiled with -ffp-contract=off, so every multiply and add rounds as in
 * the reference source text.  CUDA's rsqrt is restated as 1/sqrt; cospi and
 * sinpi are restated with an exact argument reduction (glibc has neither).
 * Built three times: the default is Real = double; -DORC_F32 gives Real =
 * float; -DORC_F32 -DORC_DP_UPDATES the float build whose OR, microcanonical
 * and heatbath updates are computed in double (DPUPDATES=true,
 * gaugefixing/apps/Makefile:46-48).
 *
 * Pins (tests/test_oracle.py): golden vectors written from the reference's
 * own headers, Random123 known answers, closed forms and invariants of every
 * gauge's sweep and quality, and sampling-distribution tests of the random
 * update rules.  Parity unpinned: the last ulp of log, cospi and sinpi in the
 * heatbath and the random transformation (the reference takes them from
 * CUDA's math library, which is not part of its sources).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_F32
typedef float Real;
#define ORC_SQRT(x) sqrtf(x)
#else
typedef double Real;
#define ORC_SQRT(x) sqrt(x)
#endif

#define ORC_EXPORT __attribute__((visibility("default")))

enum { G_LANDAU = 0, G_COULOMB = 1, G_MAG = 2, G_U1XU1 = 3 };
enum { A_OR = 0, A_MICRO = 1, A_SA = 2, A_SR = 3, A_RANDOM = 4 };

/* ------------------------------------------------------------------ */
/* Complex arithmetic (lattice/Complex.hxx:104-243)                    */
/* ------------------------------------------------------------------ */
typedef struct { Real x, y; } cplx;

static inline cplx c_make(Real x, Real y) { cplx c; c.x = x; c.y = y; return c; }
static inline cplx c_conj(cplx a) { return c_make(a.x, -a.y); }
/* operator*=(Complex): x = re*a.x - im*a.y; y = re*a.y + im*a.x (Complex.hxx:158-165) */
static inline cplx c_mul(cplx a, cplx b)
{
    cplx c;
    c.x = a.x * b.x - a.y * b.y;
    c.y = a.x * b.y + a.y * b.x;
    return c;
}
static inline cplx c_add(cplx a, cplx b) { return c_make(a.x + b.x, a.y + b.y); }
static inline cplx c_sub(cplx a, cplx b) { return c_make(a.x - b.x, a.y - b.y); }
static inline cplx c_scale(cplx a, Real s) { return c_make(a.x * s, a.y * s); }
static inline cplx c_div_real(cplx a, Real s) { return c_make(a.x / s, a.y / s); }
static inline Real c_abs2(cplx a) { return a.x * a.x + a.y * a.y; }

typedef struct { cplx m[3][3]; } mat3;

/* ------------------------------------------------------------------ */
/* Geometry (lattice/SiteIndex.hxx, SiteCoord.hxx, GpuPattern.hxx)     */
/* dims = {Nt, Nx, Ny, Nz}; mu = 0 is time; z runs fastest.            */
/* ------------------------------------------------------------------ */
ORC_EXPORT int64_t orc_volume(const int dims[4])
{
    return (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
}

/* SiteIndex::getLatticeIndex, FULL_SPLIT (SiteIndex.hxx:289-313) */
ORC_EXPORT int64_t orc_site_index(const int dims[4], const int coord[4])
{
    int64_t parity = 0, index = 0;
    for (int i = 0; i < 4; i++) parity += coord[i];
    for (int i = 0; i < 4; i++) { index *= dims[i]; index += coord[i]; }
    if (parity % 2 == 0) return index / 2;
    return index / 2 + orc_volume(dims) / 2;
}

/* SiteCoord::setLatticeIndex, FULL_SPLIT (SiteCoord.hxx:168-196) */
ORC_EXPORT void orc_site_coords(const int dims[4], int64_t latticeIndex, int coord[4])
{
    const int64_t V = orc_volume(dims);
    int curParity;
    if (latticeIndex >= V / 2) { latticeIndex -= V / 2; latticeIndex *= 2; curParity = 1; }
    else { latticeIndex *= 2; curParity = 0; }
    int64_t parity = 0;
    for (int i = 3; i >= 0; i--) {
        coord[i] = (int)(latticeIndex % dims[i]);
        parity += coord[i];
        latticeIndex /= dims[i];
    }
    if (parity % 2 != curParity) coord[3]++;
}

/* SiteIndex::calculateNeighbourTable (SiteIndex.hxx:256-284):
 * nn[(2*mu+0)*V + idx] = backward neighbour, nn[(2*mu+1)*V + idx] = forward. */
ORC_EXPORT void orc_neighbour_table(const int dims[4], int32_t *nn)
{
    const int64_t V = orc_volume(dims);
    for (int64_t n = 0; n < V; n++) {
        int site[4];
        int64_t li = n;
        for (int j = 3; j >= 0; j--) { site[j] = (int)(li % dims[j]); li /= dims[j]; }
        const int64_t here = orc_site_index(dims, site);
        for (int j = 0; j < 4; j++) {
            int cp[4];
            memcpy(cp, site, sizeof cp);
            cp[j]--; if (cp[j] < 0) cp[j] += dims[j];
            nn[(2 * j) * V + here] = (int32_t)orc_site_index(dims, cp);
            memcpy(cp, site, sizeof cp);
            cp[j]++; if (cp[j] >= dims[j]) cp[j] -= dims[j];
            nn[(2 * j + 1) * V + here] = (int32_t)orc_site_index(dims, cp);
        }
    }
}

/* GpuPattern::getIndex (GpuPattern.hxx:53-55): idx + V*(c + 2*(j + 3*(i + 3*mu))) */
ORC_EXPORT int64_t orc_pattern_index(int64_t V, int64_t idx, int mu, int i, int j, int c)
{
    return idx + V * (int64_t)(c + 2 * (j + 3 * (i + 3 * mu)));
}

static inline int64_t backward(const int dims[4], int64_t idx, int mu)
{
    int co[4];
    orc_site_coords(dims, idx, co);
    co[mu]--; if (co[mu] < 0) co[mu] += dims[mu];
    return orc_site_index(dims, co);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 and u01 (external/Random123/philox.h:206-326, u01.h)   */
/* ------------------------------------------------------------------ */
static inline uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t *hi)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

ORC_EXPORT void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo32(0xD2511F53u, c0, &hi0);
        uint32_t lo1 = mulhilo32(0xCD9E8D57u, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* u01_open_open_64_53 (u01.h:129): (1+(i>>11)) * 0x1.fffffffffffffp-54 */
ORC_EXPORT double orc_u01_64(uint64_t i)
{
    return (double)(1 + (i >> 11)) * 0x1.fffffffffffffp-54;
}
/* u01_open_open_32_24 (u01.h:111): (1+(i>>8)) * 0x1.fffffep-25f */
ORC_EXPORT float orc_u01_32(uint32_t i)
{
    return (float)(1 + (i >> 8)) * 0x1.fffffep-25f;
}

/* PhiloxWrapper (PhiloxWrapper.hxx:93-140): key = (tid, seed);
 * ctr = (n, launch counter, 0x12345678, 0xabcdef09); the first draw uses n = 1.
 * DP hands out the u64 words [1] then [0]; SP the u32 words 3,2,1,0. */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t res[4];
    int local;
} orc_rng;

static void rng_init(orc_rng *r, int tid, int seed, uint32_t counter)
{
    r->key[0] = (uint32_t)tid;
    r->key[1] = (uint32_t)seed;
    r->ctr[0] = 0;
    r->ctr[1] = counter;
    r->ctr[2] = 0x12345678u;
    r->ctr[3] = 0xabcdef09u;
    orc_philox4x32_10(r->ctr, r->key, r->res);
    r->local = 0;
}

static Real rng_rand(orc_rng *r)
{
#ifdef ORC_F32
    if (r->local == 0) {
        r->local = 3;
        r->ctr[0]++;
        orc_philox4x32_10(r->ctr, r->key, r->res);
        return orc_u01_32(r->res[3]);
    }
    r->local--;
    return orc_u01_32(r->res[r->local]);
#else
    if (r->local == 0) {
        r->local = 1;
        r->ctr[0]++;
        orc_philox4x32_10(r->ctr, r->key, r->res);
        return orc_u01_64((uint64_t)r->res[2] | ((uint64_t)r->res[3] << 32));
    }
    r->local--;
    return orc_u01_64((uint64_t)r->res[0] | ((uint64_t)r->res[1] << 32));
#endif
}

ORC_EXPORT void orc_philox_stream(int tid, int seed, uint32_t counter, int n, Real *out)
{
    orc_rng r;
    rng_init(&r, tid, seed, counter);
    for (int k = 0; k < n; k++) out[k] = rng_rand(&r);
}

/* The reference keys every per-site RNG with the global thread id of the
 * (mu = 0, updown = 0) thread of its 8-thread group: blockIdx*8*NSB + id with
 * site-in-parity = blockIdx*NSB + id (LandauKernelsSU3.hxx:225-230,272). */
ORC_EXPORT int orc_rng_tid(int64_t site_in_parity, int nsb)
{
    return (int)((site_in_parity / nsb) * 8 * nsb + site_in_parity % nsb);
}

/* cospi / sinpi with exact argument reduction (CUDA math functions called at
 * SaUpdate.hxx:122,132-133 and RandomUpdate.hxx:54-57). */
static double red_pi(double x)
{
    return x - 2.0 * nearbyint(0.5 * x); /* r in [-1, 1], exact */
}
static double orc_cospi(double x)
{
    double r = fabs(red_pi(x));
    if (r <= 0.25) return cos(M_PI * r);
    if (r <= 0.75) return sin(M_PI * (0.5 - r));
    return -cos(M_PI * (1.0 - r));
}
static double orc_sinpi(double x)
{
    double r = red_pi(x);
    double s = r < 0 ? -1.0 : 1.0;
    double a = fabs(r);
    double v;
    if (a <= 0.25) v = sin(M_PI * a);
    else if (a <= 0.75) v = cos(M_PI * (0.5 - a));
    else v = sin(M_PI * (1.0 - a));
    return s * v;
}

/* ------------------------------------------------------------------ */
/* SU(3) helpers (lattice/SU3.hxx)                                      */
/* ------------------------------------------------------------------ */

/* reconstructThirdLine (SU3.hxx:655-664) */
static void reconstruct(mat3 *u)
{
    cplx a = c_sub(c_mul(c_conj(u->m[0][1]), c_conj(u->m[1][2])), c_mul(c_conj(u->m[0][2]), c_conj(u->m[1][1])));
    cplx b = c_sub(c_mul(c_conj(u->m[0][2]), c_conj(u->m[1][0])), c_mul(c_conj(u->m[0][0]), c_conj(u->m[1][2])));
    cplx c = c_sub(c_mul(c_conj(u->m[0][0]), c_conj(u->m[1][1])), c_mul(c_conj(u->m[0][1]), c_conj(u->m[1][0])));
    Real norm = c_abs2(a) + c_abs2(b) + c_abs2(c);
    norm = (Real)(1. / ORC_SQRT(norm));
    u->m[2][0] = c_scale(a, norm);
    u->m[2][1] = c_scale(b, norm);
    u->m[2][2] = c_scale(c, norm);
}

/* projectSU3 (SU3.hxx:536-583) */
static void project(mat3 *u)
{
    Real abs_u = 0, abs_v = 0;
    cplx sp = c_make(0, 0);
    for (int i = 0; i < 3; i++) abs_u += c_abs2(u->m[0][i]);
    abs_u = ORC_SQRT(abs_u);
    for (int i = 0; i < 3; i++) u->m[0][i] = c_div_real(u->m[0][i], abs_u);
    for (int i = 0; i < 3; i++) sp = c_add(sp, c_mul(u->m[1][i], c_conj(u->m[0][i])));
    for (int i = 0; i < 3; i++) u->m[1][i] = c_sub(u->m[1][i], c_mul(u->m[0][i], sp));
    for (int i = 0; i < 3; i++) abs_v += c_abs2(u->m[1][i]);
    abs_v = ORC_SQRT(abs_v);
    for (int i = 0; i < 3; i++) u->m[1][i] = c_div_real(u->m[1][i], abs_v);
    reconstruct(u);
}

/* getSubgroupQuaternion (SU3.hxx:193-217) */
static void subgroup_quat(const mat3 *u, int i, int j, Real q[4])
{
    cplx t;
    t = u->m[i][i]; q[0] = t.x; q[3] = t.y;
    t = u->m[j][j]; q[0] += t.x; q[3] -= t.y;
    t = u->m[i][j]; q[2] = t.x; q[1] = t.y;
    t = u->m[j][i]; q[2] -= t.x; q[1] += t.y;
}

/* Quaternion::get (Quaternion.hxx:100-132) */
static inline cplx qget(const Real q[4], int i, int j)
{
    if (i == 0) return j == 0 ? c_make(q[0], q[3]) : c_make(q[2], q[1]);
    return j == 0 ? c_make(-q[2], q[1]) : c_make(q[0], -q[3]);
}

/* leftSubgroupMult (SU3.hxx:701-714) */
static void left_mult(mat3 *u, int i, int j, const Real q[4])
{
    for (int k = 0; k < 3; k++) {
        cplx IK = c_mul(qget(q, 0, 0), u->m[i][k]);
        IK = c_add(IK, c_mul(qget(q, 0, 1), u->m[j][k]));
        cplx JK = c_mul(qget(q, 1, 0), u->m[i][k]);
        JK = c_add(JK, c_mul(qget(q, 1, 1), u->m[j][k]));
        u->m[i][k] = IK;
        u->m[j][k] = JK;
    }
}

/* rightSubgroupMult (SU3.hxx:682-696) */
static void right_mult(mat3 *u, int i, int j, const Real q[4])
{
    for (int k = 0; k < 3; k++) {
        cplx KI = c_mul(qget(q, 0, 0), u->m[k][i]);
        KI = c_add(KI, c_mul(qget(q, 1, 0), u->m[k][j]));
        cplx KJ = c_mul(qget(q, 0, 1), u->m[k][i]);
        KJ = c_add(KJ, c_mul(qget(q, 1, 1), u->m[k][j]));
        u->m[k][i] = KI;
        u->m[k][j] = KJ;
    }
}

/* 18-real row-major (re, im) <-> mat3, for the helper exports */
static void m_from18(mat3 *u, const Real *a)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) u->m[i][j] = c_make(a[2 * (3 * i + j)], a[2 * (3 * i + j) + 1]);
}
static void m_to18(const mat3 *u, Real *a)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { a[2 * (3 * i + j)] = u->m[i][j].x; a[2 * (3 * i + j) + 1] = u->m[i][j].y; }
}

ORC_EXPORT void orc_reconstruct_third_line(Real m[18]) { mat3 u; m_from18(&u, m); reconstruct(&u); m_to18(&u, m); }
ORC_EXPORT void orc_project_su3(Real m[18]) { mat3 u; m_from18(&u, m); project(&u); m_to18(&u, m); }
ORC_EXPORT void orc_subgroup_quaternion(const Real m[18], int i, int j, Real q[4]) { mat3 u; m_from18(&u, m); subgroup_quat(&u, i, j, q); }
ORC_EXPORT void orc_left_subgroup_mult(Real m[18], int i, int j, const Real q[4]) { mat3 u; m_from18(&u, m); left_mult(&u, i, j, q); m_to18(&u, m); }
ORC_EXPORT void orc_right_subgroup_mult(Real m[18], int i, int j, const Real q[4]) { mat3 u; m_from18(&u, m); right_mult(&u, i, j, q); m_to18(&u, m); }

/* ------------------------------------------------------------------ */
/* Link load / store in the GpuPattern layout                           */
/* ------------------------------------------------------------------ */
/* assignWithoutThirdLine + reconstructThirdLine (LandauKernelsSU3.hxx:246-247) */
static void load_link(const Real *U, int64_t V, int64_t idx, int mu, mat3 *u)
{
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 3; j++)
            u->m[i][j] = c_make(U[orc_pattern_index(V, idx, mu, i, j, 0)], U[orc_pattern_index(V, idx, mu, i, j, 1)]);
    reconstruct(u);
}
/* globU.assignWithoutThirdLine(locU) (LandauKernelsSU3.hxx:255) */
static void store_link(Real *U, int64_t V, int64_t idx, int mu, const mat3 *u)
{
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 3; j++) {
            U[orc_pattern_index(V, idx, mu, i, j, 0)] = u->m[i][j].x;
            U[orc_pattern_index(V, idx, mu, i, j, 1)] = u->m[i][j].y;
        }
}

/* ------------------------------------------------------------------ */
/* Update rules: calculateUpdate(shA) (gaugefixing/algorithms/)          */
/* ------------------------------------------------------------------ */
typedef struct {
    int algo;
    float param;   /* OR/SR parameter or SA temperature, a float as in the reference */
    orc_rng *rng;
} orc_algo;

#if defined(ORC_F32) && defined(ORC_DP_UPDATES)
/* The single-precision build with DPUPDATES=true (gaugefixing/apps/Makefile:46-48):
 * the OR, microcanonical and heatbath updates in double on float links.  The
 * squares of the subgroup sums are float expressions assigned to doubles (shA
 * is Real = float); every result is rounded to float as it is stored. */
typedef double SaReal;
#define SA_SQRT(x) sqrt(x)

/* OrUpdate.hxx:54-64 (USE_DP_ORUPDATE) */
static void upd_or(Real a[4], float orParameter)
{
    double ai_sq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    double a0_sq = a[0] * a[0];
    double b = ((double)orParameter * a0_sq + ai_sq) / (a0_sq + ai_sq);
    double c = 1.0 / sqrt(a0_sq + b * b * ai_sq);
    a[0] *= c;
    a[1] *= b * c;
    a[2] *= b * c;
    a[3] *= b * c;
}

/* MicroUpdate.hxx:45-54 (USE_DP_MICROUPDATE) */
static void upd_micro(Real a[4])
{
    double ai_sq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    double a0_sq = a[0] * a[0];
    double b = 2. * a[0] / (a0_sq + ai_sq);
    a[0] = (Real)((a0_sq - ai_sq) / (a0_sq + ai_sq));
    a[1] *= b;
    a[2] *= b;
    a[3] *= b;
}
#else
typedef Real SaReal;
#define SA_SQRT(x) ORC_SQRT(x)

/* OrUpdate.hxx:52-78 */
static void upd_or(Real a[4], float orParameter)
{
    Real ai_sq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    Real a0_sq = a[0] * a[0];
    Real b = (orParameter * a0_sq + ai_sq) / (a0_sq + ai_sq);
    Real c = (Real)(1.0 / ORC_SQRT(a0_sq + b * b * ai_sq));
    a[0] *= c;
    a[1] *= b * c;
    a[2] *= b * c;
    a[3] *= b * c;
}

/* MicroUpdate.hxx:43-66 */
static void upd_micro(Real a[4])
{
    Real ai_sq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    Real a0_sq = a[0] * a[0];
    Real b = (Real)2. * a[0] / (a0_sq + ai_sq);
    a[0] = (a0_sq - ai_sq) / (a0_sq + ai_sq);
    a[1] *= b;
    a[2] *= b;
    a[3] *= b;
}
#endif

/* SrUpdate.hxx:54-96 */
static void upd_sr(Real a[4], float srParameter, orc_rng *rng)
{
    Real rnd = rng_rand(rng);
    Real a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    Real keep = (Real)(rnd >= srParameter), flip = (Real)(rnd < srParameter);
    a[0] = keep * a0 + flip * (a0 * a0 - a1 * a1 - a2 * a2 - a3 * a3);
    a[1] = (Real)(keep * a1 + flip * (2.0 * a0 * a1));
    a[2] = (Real)(keep * a2 + flip * (2.0 * a0 * a2));
    a[3] = (Real)(keep * a3 + flip * (2.0 * a0 * a3));
    Real c = (Real)(1.0 / ORC_SQRT(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]));
    a[0] *= c;
    a[1] *= c;
    a[2] *= c;
    a[3] *= c;
}

/* RandomUpdate.hxx:47-59 */
static void upd_random(Real a[4], orc_rng *rng)
{
    Real alpha, phi, cos_theta, sin_theta, sin_alpha;
    alpha = rng_rand(rng);
    phi = (Real)(2.0 * rng_rand(rng));
    cos_theta = (Real)(2.0 * rng_rand(rng) - 1.0);
    sin_theta = (Real)sqrt(1.0 - cos_theta * cos_theta);
    sin_alpha = (Real)orc_sinpi(alpha);
    a[0] = (Real)orc_cospi(alpha);
    a[1] = sin_alpha * sin_theta * (Real)orc_cospi(phi);
    a[2] = sin_alpha * sin_theta * (Real)orc_sinpi(phi);
    a[3] = sin_alpha * cos_theta;
}

/* SaUpdate.hxx:52-147 (Kennedy-Pendleton heatbath; temperature = 1/beta);
 * SaReal is double in the DPUPDATES build (SaUpdate.hxx:54-98), where the
 * generator still returns floats */
static void upd_sa(Real a[4], float temperature, orc_rng *rng)
{
    SaReal e0, e1, e2, e3, dk, p0;
    SaReal r1, r2, r3, r4;
    SaReal a0, a1, a2, a3;
    SaReal delta, phi, sin_alpha, sin_theta, cos_theta;
    e0 = a[0];
    e1 = -a[1];
    e2 = -a[2];
    e3 = -a[3];
    dk = (SaReal)(1.0 / SA_SQRT(e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3));
    p0 = dk * temperature;
    do {
        while ((r1 = rng_rand(rng)) < 0.0001) {}
        r1 = (SaReal)(-log(r1)) * p0;
        while ((r2 = rng_rand(rng)) < 0.0001) {}
        r2 = (SaReal)(-log(r2)) * p0;
        r3 = (SaReal)orc_cospi(2.0 * rng_rand(rng));
        r3 = r3 * r3;
        delta = r2 + r1 * r3;
        r4 = rng_rand(rng);
    } while (r4 * r4 > (1.0 - 0.5 * delta));
    a0 = (SaReal)(1.0 - delta);
    cos_theta = (SaReal)(2.0 * rng_rand(rng) - 1.0);
    sin_theta = (SaReal)sqrt(1.0 - cos_theta * cos_theta);
    sin_alpha = (SaReal)sqrt(1 - a0 * a0);
    phi = (SaReal)(2.0 * rng_rand(rng));
    a1 = sin_alpha * sin_theta * (SaReal)orc_cospi(phi);
    a2 = sin_alpha * sin_theta * (SaReal)orc_sinpi(phi);
    a3 = sin_alpha * cos_theta;
    e0 *= dk;
    e1 *= dk;
    e2 *= dk;
    e3 *= dk;
    a[0] = a0 * e0 + a3 * e3 + a2 * e2 + e1 * a1;
    a[3] = e0 * a3 - e3 * a0 + a1 * e2 - a2 * e1;
    a[2] = a3 * e1 - a0 * e2 + a2 * e0 - a1 * e3;
    a[1] = a2 * e3 + a1 * e0 - a3 * e2 - e1 * a0;
}

static void calculate_update(const orc_algo *al, Real a[4])
{
    switch (al->algo) {
    case A_OR: upd_or(a, al->param); break;
    case A_MICRO: upd_micro(a); break;
    case A_SA: upd_sa(a, al->param, al->rng); break;
    case A_SR: upd_sr(a, al->param, al->rng); break;
    case A_RANDOM: upd_random(a, al->rng); break;
    default: break;
    }
}

/* One update rule on a subgroup quaternion, with the generator of thread
 * `tid` (key (tid, seed), counter `counter`): what calculateUpdate returns
 * for that thread (gaugefixing/algorithms/{Or,Micro,Sa,Sr,Random}Update.hxx). */
ORC_EXPORT void orc_update(int algo, Real a[4], float param, int tid, int seed, uint32_t counter)
{
    orc_rng rng;
    rng_init(&rng, tid, seed, counter);
    orc_algo al = {algo, param, &rng};
    calculate_update(&al, a);
}

/* ------------------------------------------------------------------ */
/* One site update: GaugeFixingSubgroupStep::subgroup for the three      */
/* subgroups (0,2),(1,2),(0,1) (SU3.hxx:768-770).                        */
/* L[0..3] = up links U_mu(x), L[4..7] = down links U_mu(x-mu).          */
/* ------------------------------------------------------------------ */
static void site_subgroup(int gauge, const orc_algo *al, mat3 L[8], int i, int j)
{
    Real shA[4] = {0, 0, 0, 0};
    Real q[4];
    if (gauge == G_LANDAU || gauge == G_COULOMB) {
        /* DP ladder (GaugeFixingSubgroupStep.hxx:126-208) */
        for (int mu = 0; mu < 4; mu++) {
            if (gauge == G_COULOMB && mu == 0) continue;
            subgroup_quat(&L[mu], i, j, q);
            shA[0] += q[0]; shA[1] -= q[1]; shA[2] -= q[2]; shA[3] -= q[3];
        }
        for (int mu = 0; mu < 4; mu++) {
            if (gauge == G_COULOMB && mu == 0) continue;
            subgroup_quat(&L[4 + mu], i, j, q);
            shA[0] += q[0]; shA[1] += q[1]; shA[2] += q[2]; shA[3] += q[3];
        }
        calculate_update(al, shA);               /* :211-219 */
        q[0] = shA[0]; q[1] = shA[1]; q[2] = shA[2]; q[3] = shA[3];
    } else if (gauge == G_U1XU1) {
        /* GaugeFixingSubgroupStep.hxx:240-336 */
        for (int mu = 0; mu < 4; mu++) {
            subgroup_quat(&L[mu], i, j, q);
            shA[0] += q[0]; shA[3] -= q[3];
        }
        for (int mu = 0; mu < 4; mu++) {
            subgroup_quat(&L[4 + mu], i, j, q);
            shA[0] += q[0]; shA[3] += q[3];
        }
        shA[1] = 0.0; shA[2] = 0.0;
        calculate_update(al, shA);
        q[0] = shA[0]; q[1] = 0.0; q[2] = 0.0; q[3] = shA[3];
        Real norm = (Real)(1.0 / ORC_SQRT(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]));
        q[0] *= norm; q[1] *= norm; q[2] *= norm; q[3] *= norm;
    } else {
        /* MAG (GaugeFixingSubgroupStep.hxx:354-575) */
        for (int k = 0; k < 8; k++) {
            const mat3 *u = &L[k];
            cplx a00 = u->m[i][i], a01 = u->m[i][j], a10 = u->m[j][i], a11 = u->m[j][j];
            shA[0] += c_abs2(a00) - c_abs2(a01) - c_abs2(a10) + c_abs2(a11);
            cplx c;
            if (k < 4) c = c_sub(c_mul(a00, c_conj(a10)), c_mul(a01, c_conj(a11)));
            else c = c_sub(c_mul(c_conj(a00), a01), c_mul(c_conj(a10), a11));
            shA[1] += c.y;
            shA[2] += c.x;
        }
        shA[1] *= 2.;
        shA[2] *= 2.;
        shA[3] = 0;
        shA[0] = shA[0] + ORC_SQRT(shA[0] * shA[0] + shA[1] * shA[1] + shA[2] * shA[2]);
        calculate_update(al, shA);
        q[0] = shA[0]; q[1] = shA[1]; q[2] = shA[2]; q[3] = 0;
        Real norm = (Real)(1.0 / ORC_SQRT(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]));
        q[0] *= norm; q[1] *= norm; q[2] *= norm; q[3] *= norm;
    }
    /* apply: up links from the left, down links with q^dagger from the right
     * (GaugeFixingSubgroupStep.hxx:228-236) */
    Real qh[4] = {q[0], -q[1], -q[2], -q[3]};
    for (int mu = 0; mu < 4; mu++) {
        left_mult(&L[mu], i, j, q);
        right_mult(&L[4 + mu], i, j, qh);
    }
}

static void site_update(int gauge, const orc_algo *al, mat3 L[8])
{
    site_subgroup(gauge, al, L, 0, 2);
    site_subgroup(gauge, al, L, 1, 2);
    site_subgroup(gauge, al, L, 0, 1);
}

/* ------------------------------------------------------------------ */
/* Half sweeps: apply<Algorithm>(U, nn, parity, algo)                   */
/* ------------------------------------------------------------------ */

/* Landau / MAG / U1xU1 on a 4D lattice: LandauKernelsSU3.hxx:216-256,
 * MAGKernelsSU3.hxx:248-293, U1xU1KernelsSU3.hxx:216-256.
 * Returns 0 on success, -1 on bad arguments. */
ORC_EXPORT int orc_sweep(int gauge, int algo, const int dims[4], Real *U, int parity,
                         float param, int seed, uint32_t counter, int nsb)
{
    if (gauge == G_COULOMB || gauge < 0 || gauge > 3 || algo < 0 || algo > 4) return -1;
    const int64_t V = orc_volume(dims);
    const int64_t Vh = V / 2;
    if (nsb <= 0) nsb = 32;
    int64_t p;
#pragma omp parallel for schedule(static)
    for (p = 0; p < Vh; p++) {
        const int64_t site = p + (parity ? Vh : 0);
        orc_rng rng;
        rng_init(&rng, orc_rng_tid(p, nsb), seed, counter);
        orc_algo al = {algo, param, &rng};
        mat3 L[8];
        int64_t dn[4];
        for (int mu = 0; mu < 4; mu++) {
            dn[mu] = backward(dims, site, mu);
            load_link(U, V, site, mu, &L[mu]);
            load_link(U, V, dn[mu], mu, &L[4 + mu]);
        }
        site_update(gauge, &al, L);
        for (int mu = 0; mu < 4; mu++) {
            store_link(U, V, site, mu, &L[mu]);
            store_link(U, V, dn[mu], mu, &L[4 + mu]);
        }
    }
    return 0;
}

/* Coulomb on one timeslice: CoulombKernelsSU3.hxx:208-250.  UtUp holds slice t,
 * UtDw slice t-1, both in the GpuPattern layout over {1,Nx,Ny,Nz}.
 * dims3 = {1, Nx, Ny, Nz}. */
ORC_EXPORT int orc_sweep_coulomb(int algo, const int dims3[4], Real *UtUp, Real *UtDw, int parity,
                                 float param, int seed, uint32_t counter, int nsb)
{
    if (dims3[0] != 1 || algo < 0 || algo > 4) return -1;
    const int64_t V = orc_volume(dims3);
    const int64_t Vh = V / 2;
    if (nsb <= 0) nsb = 32;
    int64_t p;
#pragma omp parallel for schedule(static)
    for (p = 0; p < Vh; p++) {
        const int64_t site = p + (parity ? Vh : 0);
        orc_rng rng;
        rng_init(&rng, orc_rng_tid(p, nsb), seed, counter);
        orc_algo al = {algo, param, &rng};
        mat3 L[8];
        int64_t dn[4];
        for (int mu = 0; mu < 4; mu++) {
            dn[mu] = (mu == 0) ? site : backward(dims3, site, mu);
            load_link(UtUp, V, site, mu, &L[mu]);
            load_link(mu == 0 ? UtDw : UtUp, V, dn[mu], mu, &L[4 + mu]);
        }
        site_update(G_COULOMB, &al, L);
        for (int mu = 0; mu < 4; mu++) {
            store_link(UtUp, V, site, mu, &L[mu]);
            store_link(mu == 0 ? UtDw : UtUp, V, dn[mu], mu, &L[4 + mu]);
        }
    }
    return 0;
}

/* CommonKernelsSU3::projectSU3 (CommonKernelsSU3.hxx:71-90): every link of
 * every site is reunitarized and its third row written. */
ORC_EXPORT void orc_project_all(const int dims[4], Real *U)
{
    const int64_t V = orc_volume(dims);
    int64_t s;
#pragma omp parallel for schedule(static)
    for (s = 0; s < V; s++) {
        for (int mu = 0; mu < 4; mu++) {
            mat3 u;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    u.m[i][j] = c_make(U[orc_pattern_index(V, s, mu, i, j, 0)], U[orc_pattern_index(V, s, mu, i, j, 1)]);
            project(&u);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    U[orc_pattern_index(V, s, mu, i, j, 0)] = u.m[i][j].x;
                    U[orc_pattern_index(V, s, mu, i, j, 1)] = u.m[i][j].y;
                }
        }
    }
}

/* CommonKernelsSU3::setHot (CommonKernelsSU3.hxx:92-130): per site, an RNG
 * keyed by (site, seed); per link the product of three random SU(2) subgroup
 * elements (Quaternion::projectSU2, Quaternion.hxx:220-228) multiplied from
 * the right onto the identity; all 18 reals are written. */
ORC_EXPORT void orc_set_hot(const int dims[4], Real *U, int seed, uint32_t counter)
{
    const int64_t V = orc_volume(dims);
    int64_t s;
#pragma omp parallel for schedule(static)
    for (s = 0; s < V; s++) {
        orc_rng rng;
        rng_init(&rng, (int)s, seed, counter);
        for (int mu = 0; mu < 4; mu++) {
            mat3 u;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) u.m[i][j] = c_make((Real)(i == j), 0);
            for (int i = 0; i < 2; i++)
                for (int j = i + 1; j < 3; j++) {
                    Real q[4];
                    for (int k = 0; k < 4; k++) q[k] = (Real)(rng_rand(&rng) * 2.0 - 1.0);
                    Real det = (Real)(1.0 / ORC_SQRT(q[0] * q[0] + q[3] * q[3] + q[2] * q[2] + q[1] * q[1]));
                    for (int k = 0; k < 4; k++) q[k] *= det;
                    right_mult(&u, i, j, q);
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    U[orc_pattern_index(V, s, mu, i, j, 0)] = u.m[i][j].x;
                    U[orc_pattern_index(V, s, mu, i, j, 1)] = u.m[i][j].y;
                }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Gauge quality                                                         */
/* ------------------------------------------------------------------ */

/* Landau/U1xU1 (LandauKernelsSU3.hxx:114-191) and Coulomb (mu0 = 1,
 * CoulombKernelsSU3.hxx:110-183): theta_x and gff_x for one site. */
static void quality_site_landau(const int dims[4], const Real *U, int64_t site, int mu0, double *dA, double *dGff)
{
    const int64_t V = orc_volume(dims);
    mat3 sum, tmp;
    memset(&sum, 0, sizeof sum);
    for (int mu = mu0; mu < 4; mu++) {
        load_link(U, V, site, mu, &tmp);
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) sum.m[i][j] = c_add(sum.m[i][j], tmp.m[i][j]);
        load_link(U, V, backward(dims, site, mu), mu, &tmp);
        for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) sum.m[i][j] = c_sub(sum.m[i][j], tmp.m[i][j]);
    }
    /* Sum -= Sum.trace()/Real(3.)  (Matrix::trace, Matrix::operator-=(T)) */
    cplx tr = c_make(0, 0);
    for (int i = 0; i < 3; i++) tr = c_add(tr, sum.m[i][i]);
    tr = c_div_real(tr, (Real)3.);
    for (int i = 0; i < 3; i++) sum.m[i][i] = c_sub(sum.m[i][i], tr);
    mat3 herm;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) herm.m[j][i] = c_conj(sum.m[i][j]);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) sum.m[i][j] = c_sub(sum.m[i][j], herm.m[i][j]);
    double prec = 0;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) prec += c_abs2(sum.m[i][j]);
    *dA = prec;
    double result = 0;
    for (int mu = mu0; mu < 4; mu++) {
        load_link(U, V, site, mu, &tmp);
        cplx t = c_make(0, 0);
        for (int i = 0; i < 3; i++) t = c_add(t, tmp.m[i][i]);
        result += t.x;
    }
    *dGff = result;
}

/* MAG (MAGKernelsSU3.hxx:116-223) */
static void quality_site_mag(const int dims[4], const Real *U, int64_t site, double *dA, double *dGff)
{
    const int64_t V = orc_volume(dims);
    double gff = 0.0;
    double Xx[3] = {0, 0, 0}, Xy[3] = {0, 0, 0};
    for (int mu = 0; mu < 4; mu++) {
        mat3 up, dw;
        load_link(U, V, site, mu, &up);
        load_link(U, V, backward(dims, site, mu), mu, &dw);
        for (int j = 0; j < 3; j++) gff += c_abs2(up.m[j][j]);
        for (int i = 0; i < 2; i++)
            for (int j = i + 1; j < 3; j++) {
                /* getSubgroupMatrix(i,j) of the up link */
                cplx a00 = up.m[i][i], a01 = up.m[i][j], a10 = up.m[j][i], a11 = up.m[j][j];
                cplx c = c_sub(c_mul(a00, c_conj(a10)), c_mul(a01, c_conj(a11)));
                Xx[i + j - 1] += (double)c.x;
                Xy[i + j - 1] += (double)c.y;
                /* getSubgroupMatrixHermitian(i,j) of the down link (SU3.hxx:231-241) */
                cplx b00 = c_conj(dw.m[i][i]), b01 = c_conj(dw.m[j][i]), b10 = c_conj(dw.m[i][j]), b11 = c_conj(dw.m[j][j]);
                c = c_sub(c_mul(b00, c_conj(b10)), c_mul(b01, c_conj(b11)));
                Xx[i + j - 1] += (double)c.x;
                Xy[i + j - 1] += (double)c.y;
            }
    }
    *dGff = gff;
    *dA = sqrt(Xx[0] * Xx[0] + Xy[0] * Xy[0]) + sqrt(Xx[1] * Xx[1] + Xy[1] * Xy[1]) + sqrt(Xx[2] * Xx[2] + Xy[2] * Xy[2]);
}

/* Per-site values, as the generateGaugeQualityPerSite kernels write them. */
ORC_EXPORT int orc_quality_per_site(int gauge, const int dims[4], const Real *U, double *dGff, double *dA)
{
    const int64_t V = orc_volume(dims);
    if (gauge == G_COULOMB && dims[0] != 1) return -1;
    int64_t s;
#pragma omp parallel for schedule(static)
    for (s = 0; s < V; s++) {
        if (gauge == G_MAG) quality_site_mag(dims, U, s, &dA[s], &dGff[s]);
        else quality_site_landau(dims, U, s, gauge == G_COULOMB ? 1 : 0, &dA[s], &dGff[s]);
    }
    return 0;
}

/* GaugeFixingStats::generateGaugeQuality with StoppingCrit AVERAGE:
 * serial sum in index order (reduceGaugeQuality, GaugeFixingStats.hxx:655-668),
 * then the prefactors (LandauKernelsSU3.hxx:75-82, CoulombKernelsSU3.hxx:71-78,
 * MAGKernelsSU3.hxx:76-83) and the division by V (GaugeFixingStats.hxx:203-206). */
ORC_EXPORT int orc_quality(int gauge, const int dims[4], const Real *U, double *gff, double *theta)
{
    const int64_t V = orc_volume(dims);
    double *g = (double *)malloc(sizeof(double) * V);
    double *a = (double *)malloc(sizeof(double) * V);
    if (!g || !a) { free(g); free(a); return -1; }
    int rc = orc_quality_per_site(gauge, dims, U, g, a);
    double sg = 0, sa = 0;
    for (int64_t s = 0; s < V; s++) { sg += g[s]; sa = sa + a[s]; }
    free(g);
    free(a);
    if (rc) return rc;
    const double Nc = 3.0, Ndim = 4.0;
    double pg = (gauge == G_COULOMB) ? 1. / (Nc * (Ndim - 1)) : 1. / (Nc * Ndim);
    *gff = sg * pg / (double)V;
    *theta = sa * (1. / Nc) / (double)V;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Driver loop (LandauGaugeFixingSU3_4D.cu:244-271): OR sweeps with       */
/* reprojection when i % reproject == 0 and a check when                */
/* i % checkprecision == 0 that stops once theta < precision (a float). */
/* Returns the number of completed sweeps; *gff / *theta hold the last   */
/* check's values.  A final projectSU3 writes the third rows.            */
/* ------------------------------------------------------------------ */
ORC_EXPORT int orc_run_or(int gauge, const int dims[4], Real *U, float omega, float precision,
                          int maxiter, int checkprecision, int reproject, double *gff, double *theta)
{
    double g = 0, t = 0;
    int i;
    int done = 0;
    for (i = 0; i < maxiter; i++) {
        orc_sweep(gauge, A_OR, dims, U, 0, omega, 0, 0, 32);
        orc_sweep(gauge, A_OR, dims, U, 1, omega, 0, 0, 32);
        done = i + 1;
        if (reproject > 0 && i % reproject == 0) orc_project_all(dims, U);
        if (checkprecision > 0 && i % checkprecision == 0) {
            orc_quality(gauge, dims, U, &g, &t);
            if (t < precision) break;
        }
    }
    /* reconstruct the third rows (LandauGaugeFixingSU3_4D.cu:271) */
    orc_project_all(dims, U);
    *gff = g;
    *theta = t;
    return done;
}

ORC_EXPORT int orc_real_size(void) { return (int)sizeof(Real); }

ORC_EXPORT void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
