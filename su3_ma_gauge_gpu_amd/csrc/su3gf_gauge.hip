// su3gf_gauge.hip -- the single-precision sweep of the mixed-precision driver,
// written on the gauge transformation instead of the links.
//
// Every update of the checkerboard sweep is a gauge transformation: after any
// number of sweeps the links are U_mu(x) = g(x) U0_mu(x) g(x+mu)^dagger for the
// starting configuration U0 and the accumulated g.  This kernel keeps U0
// read-only and updates g alone.  For a site x, with y = x+mu and z = x-mu,
//
//   P = sum_mu g(y) U0_mu(x)^dagger,   Q = sum_mu g(z) U0_mu(z),
//   M = sum_mu U_mu(x)^dagger + U_mu(z) = (P + Q) g(x)^dagger.
//
// The reference accumulates the subgroup quaternion over the eight links as
// sum conj(q(U_up)) + q(U_down) (GaugeFixingSubgroupStep.hxx:126-208); since
// q(A^dagger) = conj(q(A)) and q is linear this is q(M), and the update of the
// up links from the left and the down links from the right by the SU(2)
// element turns M into M G^dagger (SU3.hxx:682-714).  So the three subgroup
// steps -- (0,2), (1,2), (0,1), as in SU3.hxx:768-770 -- run on the one 3x3
// matrix M, and g(x) is multiplied from the left by the same elements.  For
// Landau, Coulomb (mu = 0 left out of M) and U1xU1 (q1 = q2 = 0, then
// normalised) the result is the reference's update in exact arithmetic; the
// rounding differs, as it does for any single-precision sweep.
//
// Arithmetic: the stored g and U0 are FP32; the kernel computes either in
// FP32 or in FP64 (template parameter A).  On this hardware FP64 FMAs run at
// the rate of unpacked FP32 ones, and FP64 arithmetic removes the rounding
// noise of the 3x3 products, which otherwise stops convergence near
// theta ~ 2e-12; what remains is the FP32 storage of U0 and g (theta ~ 4e-13).
// Row 2 of U0 and of g is rebuilt as conj(row0 x row1) without normalising:
// the rows are unit to FP32 precision already.
//
// Bytes per site-update: the links of U0 are read once per parity kernel
// (4 V links for V/2 sites: 8 x 48 B), g of the other parity is read once
// (48 B) and g of the updated site read and written (96 B): 528 B, against
// 864 B for updating FP32 links together with g.
//
// Fused check: with THETA = 1 (even kernel, before the update) or 2 (odd
// kernel, after it) the kernel also sums, over its sites, theta_x from
// Delta = sum U_up - sum U_down = g P^dagger - Q g^dagger and the functional
// Re tr of every link (each link has exactly one endpoint of either parity).
// These are the quantities of LandauKernelsSU3.hxx:114-191.
#include <algorithm>
#include <atomic>
#include <type_traits>

#include "su3gf_sweep.h"

namespace su3gf {
namespace {

constexpr int kGBlock = 256;  // 64 sites x 4 directions

// Waves per block of the gauge-transformation sweep: one per direction that
// enters S, so three for Coulomb, whose S leaves out the time direction.
template <int G> constexpr int gsweep_waves() { return G == GF_COULOMB ? 3 : 4; }

// The complex products below are written as explicit fused multiply-adds:
// the compiler keeps a * b + c unfused when c is a sum formed elsewhere,
// which costs a third more instructions in these issue-bound loops.
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// s += a b
template <typename A> __device__ __forceinline__ void cmac(Cx<A> &s, const Cx<A> &a, const Cx<A> &b)
{
    s.x = fma_(a.x, b.x, s.x);
    s.x = fma_(-a.y, b.y, s.x);
    s.y = fma_(a.x, b.y, s.y);
    s.y = fma_(a.y, b.x, s.y);
}

// s += a conj(b)
template <typename A> __device__ __forceinline__ void cmacc(Cx<A> &s, const Cx<A> &a, const Cx<A> &b)
{
    s.x = fma_(a.x, b.x, s.x);
    s.x = fma_(a.y, b.y, s.x);
    s.y = fma_(a.y, b.x, s.y);
    s.y = fma_(-a.x, b.y, s.y);
}

template <typename A> __device__ __forceinline__ Cx<A> czero() { return cx((A)0, (A)0); }

// conj(a b - c d)
template <typename A> __device__ __forceinline__ Cx<A> cross_term(Cx<A> a, Cx<A> b, Cx<A> c, Cx<A> d)
{
    A x = a.x * b.x;
    x = fma_(-a.y, b.y, x);
    x = fma_(-c.x, d.x, x);
    x = fma_(c.y, d.y, x);
    A y = c.x * d.y;
    y = fma_(c.y, d.x, y);
    y = fma_(-a.x, b.y, y);
    y = fma_(-a.y, b.x, y);
    return cx(x, y);
}

// row 2 = conj(row0 x row1) without normalising: the rows are unit to the
// precision they are stored in
template <typename A> __device__ __forceinline__ void third_row(M3<A> &u)
{
    u.e[2][0] = cross_term(u.e[0][1], u.e[1][2], u.e[0][2], u.e[1][1]);
    u.e[2][1] = cross_term(u.e[0][2], u.e[1][0], u.e[0][0], u.e[1][2]);
    u.e[2][2] = cross_term(u.e[0][0], u.e[1][1], u.e[0][1], u.e[1][0]);
}

// Buffer resource over one component-major field of 12 FP32 components per
// site (rows 0,1): component c of site i is at byte c*4V + 4i.  Loads through
// it are a per-lane 32-bit byte offset plus a uniform per-component offset,
// with no 64-bit address arithmetic per load; the hardware checks every
// access against the field's size.
struct Field {
    __amdgpu_buffer_rsrc_t r;
    uint32_t V4;  // bytes per component plane
};
__device__ __forceinline__ Field make_field(const float *base, int64_t V, int words = 12)
{
    Field f;
    // word 3: the untyped-access format the gfx9-family buffer instructions expect
    f.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, (int)(4 * words * V), 0x00020000);
    f.V4 = (uint32_t)(4 * V);
    return f;
}

// A field in tiles of 64 sites, `words` 32-bit components per site: component
// c of site i at word (i & ~63) * words + c * 64 + (i & 63) (the layout of
// gauge_ref for g, words = 12, and of the 24-bit links, words = 9).  V4 is
// then the component stride of a tile, 256 B, and a site's byte offset 4 i
// becomes tile_off(4 i, words).
__device__ __forceinline__ Field make_tiled_field(const float *base, int64_t V, int words)
{
    Field f;
    const int64_t tiles = (V + 63) / 64;
    f.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, (int)(4 * words * 64 * tiles),
                                            0x00020000);
    f.V4 = 256u;
    return f;
}
__device__ __forceinline__ uint32_t tile_off(uint32_t boff, uint32_t words)
{
    return (boff & ~255u) * words + (boff & 255u);
}

// Cache policy of a buffer access (the instruction's sc1 bit): coherent at
// device scope, for g that a block on another XCD writes or reads within the
// same launch (the pipelined sweep); 0 is the default policy.
constexpr int kSc1 = 16;

// rows 0,1 of site byte offset boff
template <typename A, int CP = 0> __device__ __forceinline__ void load_rows01(const Field &f, uint32_t boff, M3<A> &u)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            const float re =
                __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)boff, (int)(comp * f.V4), CP));
            const float im = __builtin_bit_cast(
                float, __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)boff, (int)((comp + 1) * f.V4), CP));
            u.e[i][j] = cx((A)re, (A)im);
        }
}

// Links in 24-bit fixed point: each of the 12 reals of rows 0,1 is stored as
// round(v * kQ24), a 24-bit two's-complement integer (|v| <= 1, so the
// quantisation error is at most 2^-24, the size of an FP32 rounding error of
// an entry near 1), and four of them share three 32-bit words:
//   w0 = a | b << 24,   w1 = b >> 8 | c << 16,   w2 = c >> 16 | d << 8.
// Nine words per link instead of twelve: 36 B instead of 48, in tiles of 64
// sites per direction (make_tiled_field, words = 9).  The
// loads scale by 1/kQ24 at once: row 2 is rebuilt from rows 0,1 and is
// quadratic in them, so unscaled integers would not form an SU(3) matrix.
constexpr double kQ24 = 8388607.0;  // 2^23 - 1: +-1 are exact
constexpr float kInvQ24 = (float)(1.0 / kQ24);

template <typename A> __device__ __forceinline__ void load_rows01_q24(const Field &f, uint32_t boff, M3<A> &u)
{
    A v[12];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t w0 = __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)boff, (int)((3 * k) * f.V4), 0);
        const uint32_t w1 = __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)boff, (int)((3 * k + 1) * f.V4), 0);
        const uint32_t w2 = __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)boff, (int)((3 * k + 2) * f.V4), 0);
        // the bit-field builtin's result type is unsigned: convert it as the signed value it holds
        const int a = (int)__builtin_amdgcn_sbfe((int)w0, 0, 24);
        const int b = (int)__builtin_amdgcn_sbfe((int)__builtin_amdgcn_alignbit(w1, w0, 24), 0, 24);
        const int c = (int)__builtin_amdgcn_sbfe((int)__builtin_amdgcn_alignbit(w2, w1, 16), 0, 24);
        const int d = (int)w2 >> 8;
        v[4 * k + 0] = (A)a * (A)kInvQ24;
        v[4 * k + 1] = (A)b * (A)kInvQ24;
        v[4 * k + 2] = (A)c * (A)kInvQ24;
        v[4 * k + 3] = (A)d * (A)kInvQ24;
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) u.e[i][j] = cx(v[2 * (j + 3 * i)], v[2 * (j + 3 * i) + 1]);
}

// FP64 storage (LF = 2, the FP64 phase of the mixed driver): g and each
// direction of the links as 12 doubles per site in tiles of 64 sites
// (kTiled12's layout): component c of site i at double
// (i & ~63) * 12 + 64 c + (i & 63).  Byte offsets are 8 i, the component
// stride of a tile 512 B.
__device__ __forceinline__ Field make_tiled_field64(const double *base, int64_t V)
{
    Field f;
    f.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0, (int)(8 * tiled12_reals(V)),
                                            0x00020000);
    f.V4 = 512u;
    return f;
}
__device__ __forceinline__ uint32_t tile_off64(uint32_t boff) { return (boff & ~511u) * 12 + (boff & 511u); }

template <typename A> __device__ __forceinline__ void load_rows01_64(const Field &f, uint32_t boff, M3<A> &u)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            const double re = __builtin_bit_cast(
                double, __builtin_amdgcn_raw_buffer_load_b64(f.r, (int)boff, (int)(comp * f.V4), 0));
            const double im = __builtin_bit_cast(
                double, __builtin_amdgcn_raw_buffer_load_b64(f.r, (int)boff, (int)((comp + 1) * f.V4), 0));
            u.e[i][j] = cx((A)re, (A)im);
        }
}

// byte offset of site i in a field of the storage of link form LF
template <int LF> __device__ __forceinline__ uint32_t site_boff(int64_t i) { return (uint32_t)i * (LF == 2 ? 8u : 4u); }

// U0 rows 0,1 of the site at byte offset boff (site_boff): FP32 planes (LF =
// 0), 24-bit fixed point in tiles (LF = 1) or FP64 in tiles (LF = 2)
template <typename A, int LF> __device__ __forceinline__ void load_links(const Field &f, uint32_t boff, M3<A> &u)
{
    if (LF == 2) load_rows01_64(f, tile_off64(boff), u);
    else if (LF == 1) load_rows01_q24(f, tile_off(boff, 9), u);
    else load_rows01(f, boff, u);
}

// the field of U0_mu (one direction)
template <int LF> __device__ __forceinline__ Field links_field(const float *U0, int64_t V, int mu)
{
    if (LF == 2) return make_tiled_field64((const double *)U0 + tiled12_reals(V) * mu, V);
    if (LF == 1) return make_tiled_field(U0 + 9 * 64 * ((V + 63) / 64) * mu, V, 9);
    return make_field(U0 + 12 * V * mu, V);
}

// g: FP32 tiles, or FP64 tiles with FP64 links (LF = 2)
template <int LF, typename Args> __device__ __forceinline__ Field g_field(const Args &a)
{
    if (LF == 2) return make_tiled_field64((const double *)a.g, a.geo.V);
    return make_tiled_field(a.g, a.geo.V, 12);
}
template <typename A, int LF, int CP = 0> __device__ __forceinline__ void load_g(const Field &f, uint32_t boff, M3<A> &u)
{
    static_assert(LF != 2 || CP == 0, "FP64 g is only swept by the two kernels");
    if (LF == 2) load_rows01_64(f, tile_off64(boff), u);
    else load_rows01<A, CP>(f, tile_off(boff, 12), u);
}

typedef unsigned int B64 __attribute__((__vector_size__(8)));  // the 64-bit buffer store's operand

template <typename A> __device__ __forceinline__ void store_rows01_64(const Field &f, uint32_t boff, const M3<A> &u)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B64, (double)u.e[i][j].x), f.r, (int)boff,
                                                  (int)(comp * f.V4), 0);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(B64, (double)u.e[i][j].y), f.r, (int)boff,
                                                  (int)((comp + 1) * f.V4), 0);
        }
}

template <typename A, int CP = 0> __device__ __forceinline__ void store_rows01(const Field &f, uint32_t boff, const M3<A> &u)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, (float)u.e[i][j].x), f.r, (int)boff,
                                                  (int)(comp * f.V4), CP);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, (float)u.e[i][j].y), f.r, (int)boff,
                                                  (int)((comp + 1) * f.V4), CP);
        }
}

template <typename A, int LF, int CP = 0> __device__ __forceinline__ void store_g(const Field &f, uint32_t boff, const M3<A> &u)
{
    static_assert(LF != 2 || CP == 0, "FP64 g is only swept by the two kernels");
    if (LF == 2) store_rows01_64(f, tile_off64(boff), u);
    else store_rows01<A, CP>(f, tile_off(boff, 12), u);
}

// Z = rows 0,1 of X Y^dagger (Y complete), then row 2 of the SU(3) product
template <typename A> __device__ __forceinline__ void prod_adag(M3<A> &Z, const M3<A> &X, const M3<A> &Y)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Cx<A> s = czero<A>();
#pragma unroll
            for (int j = 0; j < 3; j++) cmacc(s, X.e[i][j], Y.e[k][j]);
            Z.e[i][k] = s;
        }
    third_row(Z);
}

// Z = rows 0,1 of X Y (Y complete), then row 2 of the SU(3) product
template <typename A> __device__ __forceinline__ void prod_ab(M3<A> &Z, const M3<A> &X, const M3<A> &Y)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Cx<A> s = czero<A>();
#pragma unroll
            for (int j = 0; j < 3; j++) cmac(s, X.e[i][j], Y.e[j][k]);
            Z.e[i][k] = s;
        }
    third_row(Z);
}

// acc (+)= X Y^dagger, all nine entries
template <typename A> __device__ __forceinline__ void mac_adag(M3<A> &acc, const M3<A> &X, const M3<A> &Y)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Cx<A> s = acc.e[i][k];
#pragma unroll
            for (int j = 0; j < 3; j++) cmacc(s, X.e[i][j], Y.e[k][j]);
            acc.e[i][k] = s;
        }
}

template <typename A> __device__ __forceinline__ void zero(M3<A> &m)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) m.e[i][j] = czero<A>();
}

// rows I,J of u <- G(q) rows I,J, G(q) = [[q0+iq3, q2+iq1], [-q2+iq1, q0-iq3]]
template <typename A, int I, int J> __device__ __forceinline__ void left_su2(M3<A> &u, const A q[4])
{
    const Cx<A> g00 = cx(q[0], q[3]), g01 = cx(q[2], q[1]), g10 = cx(-q[2], q[1]), g11 = cx(q[0], -q[3]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Cx<A> a = czero<A>(), b = czero<A>();
        cmac(a, g00, u.e[I][k]);
        cmac(a, g01, u.e[J][k]);
        cmac(b, g10, u.e[I][k]);
        cmac(b, g11, u.e[J][k]);
        u.e[I][k] = a;
        u.e[J][k] = b;
    }
}

// entry (r, I) or (r, J) of M G(q)^dagger: new col I = M[:,I] h00 + M[:,J] h10,
// new col J = M[:,I] h01 + M[:,J] h11 with G(q)^dagger = [[h00, h01], [h10, h11]]
template <typename A> struct Hd {
    Cx<A> h00, h01, h10, h11;
};
template <typename A> __device__ __forceinline__ Hd<A> su2_dag(const A q[4])
{
    Hd<A> h;
    h.h00 = cx(q[0], -q[3]);
    h.h01 = cx(-q[2], -q[1]);
    h.h10 = cx(q[2], -q[1]);
    h.h11 = cx(q[0], q[3]);
    return h;
}
template <typename A> __device__ __forceinline__ Cx<A> col_i(const Hd<A> &h, Cx<A> mi, Cx<A> mj)
{
    Cx<A> s = czero<A>();
    cmac(s, mi, h.h00);
    cmac(s, mj, h.h10);
    return s;
}
template <typename A> __device__ __forceinline__ Cx<A> col_j(const Hd<A> &h, Cx<A> mi, Cx<A> mj)
{
    Cx<A> s = czero<A>();
    cmac(s, mi, h.h01);
    cmac(s, mj, h.h11);
    return s;
}

// Gram-Schmidt on rows 0,1 (SU3.hxx:536-583 without the third row)
template <typename A> __device__ __forceinline__ void project_rows01(M3<A> &u)
{
    A n0 = u.e[0][0].x * u.e[0][0].x;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (k) n0 = fma_(u.e[0][k].x, u.e[0][k].x, n0);
        n0 = fma_(u.e[0][k].y, u.e[0][k].y, n0);
    }
    n0 = rsq(n0);
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[0][k] = u.e[0][k] * n0;
    Cx<A> sp = czero<A>();
#pragma unroll
    for (int k = 0; k < 3; k++) cmacc(sp, u.e[1][k], u.e[0][k]);
    A n1 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Cx<A> r = u.e[1][k];
        r.x = fma_(-u.e[0][k].x, sp.x, r.x);
        r.x = fma_(u.e[0][k].y, sp.y, r.x);
        r.y = fma_(-u.e[0][k].x, sp.y, r.y);
        r.y = fma_(-u.e[0][k].y, sp.x, r.y);
        u.e[1][k] = r;
        n1 = fma_(r.x, r.x, n1);
        n1 = fma_(r.y, r.y, n1);
    }
    n1 = rsq(n1);
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[1][k] = u.e[1][k] * n1;
}

// theta_x and the Re tr of the site's eight links from M = S g(x)^dagger,
// S = sum_mu g(x+mu) U0_mu(x)^dagger + g(x-mu) U0_mu(x-mu).  With A and B the
// sums of the site's up links g(x) U0_mu(x) g(x+mu)^dagger and down links
// g(x-mu) U0_mu(x-mu) g(x)^dagger, M = A^dagger + B, so
// M - M^dagger = -(Delta - Delta^dagger) for Delta = A - B: theta_x, which
// depends on Delta only through its traceless anti-Hermitian part
// (LandauKernelsSU3.hxx:114-191), follows from M (the sign drops out of the
// squares), and Re tr M = Re tr A + Re tr B (Landau, Coulomb, U1xU1).
template <typename A> __device__ __forceinline__ void theta_from_m(const M3<A> &M, double &th, double &gf)
{
    const A imtr3 = (M.e[0][0].y + M.e[1][1].y + M.e[2][2].y) / (A)3;
    double s[kThetaVals];
    s[0] = (double)((A)2 * (M.e[0][0].y - imtr3));
    s[1] = (double)((A)2 * (M.e[1][1].y - imtr3));
    s[2] = (double)((A)2 * (M.e[2][2].y - imtr3));
    s[3] = (double)(M.e[0][1].x - M.e[1][0].x);
    s[4] = (double)(M.e[0][1].y + M.e[1][0].y);
    s[5] = (double)(M.e[0][2].x - M.e[2][0].x);
    s[6] = (double)(M.e[0][2].y + M.e[2][0].y);
    s[7] = (double)(M.e[1][2].x - M.e[2][1].x);
    s[8] = (double)(M.e[1][2].y + M.e[2][1].y);
    s[9] = (double)((M.e[0][0].x + M.e[1][1].x) + M.e[2][2].x);
    theta_site<GF_LANDAU>(s, th, gf);
}

// Block order.  The eight XCDs take blocks round-robin (block b runs on XCD
// b % 8), each with its own L2.  remap = 1 deals the blocks out so that each
// XCD works through one contiguous eighth of the sites, which keeps the x, y
// and z neighbours' g in that XCD's L2.  remap = 2 (Nt divisible by 8 and a
// timeslice of one parity a whole number of 64-site blocks) also walks each
// XCD's eighth -- Nt/8 timeslices -- with t innermost, so that the t
// neighbours' g is read from L2 too: in plain order they are a timeslice
// (V3/2 sites, megabytes of traffic) apart.  With t innermost, a block's x
// neighbours are (blocks per timeslice row of x) x (timeslices walked
// together) blocks away in the XCD's sequence; walking the XCD's timeslices in
// passes of tpass shortens that distance (more x neighbours in L2) at the
// cost of the t neighbours across a pass boundary.
__device__ __forceinline__ int64_t logical_block(const GaugeSweepArgs &a)
{
    const uint32_t b = blockIdx.x, n = gridDim.x;  // below 2^31
    if (a.remap == 0) return b;
    const uint32_t xcd = b & 7, k = b >> 3;
    if (a.remap == 1) return (int64_t)xcd * (n >> 3) + k;
    const uint32_t P = (uint32_t)a.tpass;
    const uint32_t pass = fdiv(k, a.pass_div);
    const uint32_t rem = k - pass * P * (uint32_t)a.slice_blocks;
    const uint32_t q = fdiv(rem, a.tpass_div);
    const int64_t t = (int64_t)xcd * a.slices_per_xcd + pass * P + (rem - q * P);  // timeslice
    return t * a.slice_blocks + q;
}

// Wavefront launch: the parity and logical block of this block, or false for
// a padding block (uniform over the block)
__device__ __forceinline__ bool wave_block(const GaugeSweepArgs &a, int64_t &blk, int &parity)
{
    int64_t b = blockIdx.x, n;
    if (b < a.wf_even_grid) {
        parity = 0;
        n = a.wf_even_n;
    } else {
        parity = 1;
        b -= a.wf_even_grid;
        n = a.wf_odd_n + a.wf_odd1_n;
    }
    const int64_t per = (n + 7) / 8;
    const int64_t k = (b & 7) * per + (b >> 3);
    if (k >= n) return false;
    if (parity == 0) blk = a.wf_even0 + k;
    else blk = k < a.wf_odd_n ? a.wf_odd0 + k : a.wf_odd1 + (k - a.wf_odd_n);
    return true;
}

// The update rule on the subgroup quaternion, with the U1xU1 restriction to
// (q0, q3) and normalisation (GaugeFixingSubgroupStep.hxx:320-338).
template <typename A, int G, int ALGO>
__device__ __forceinline__ void gauge_update(A q[4], float param, Philox &rng)
{
    if (G == GF_U1XU1) {
        q[1] = 0;
        q[2] = 0;
    }
    calc_update<A, ALGO, false>(q, param, rng);
    if (G == GF_U1XU1) {
        q[1] = 0;
        q[2] = 0;
        const A n = rsq(q[0] * q[0] + q[3] * q[3]);
        q[0] *= n;
        q[3] *= n;
    }
}

// The direction-mu terms of site x (coordinates c, parity par):
// X = U0_mu(x) g(y)^dagger and Z = g(z) U0_mu(z), y = x+mu, z = x-mu, so that
// S_mu = X^dagger + Z.  coh (uniform): the neighbours' g through loads
// coherent at device scope.
template <typename A, int LF, typename Args>
__device__ __forceinline__ void direction_terms(const Args &a, const Field &fg, int mu, const int c[4],
                                                int par, uint32_t site_off, M3<A> &X, M3<A> &Z, bool coh = false)
{
    M3<A> T, W;
    const uint32_t y_off = site_boff<LF>(forward(a.geo, c, par, mu));
    const uint32_t z_off = site_boff<LF>(backward(a.geo, c, par, mu));
    const Field fu = links_field<LF>(a.U0, a.geo.V, mu);
    load_links<A, LF>(fu, site_off, W);
    if (LF != 2 && coh) load_g<A, LF, LF == 2 ? 0 : kSc1>(fg, y_off, T);
    else load_g<A, LF>(fg, y_off, T);
    third_row(T);
    prod_adag(X, W, T);  // U0_mu(x) g(y)^dagger
    if (LF != 2 && coh) load_g<A, LF, LF == 2 ? 0 : kSc1>(fg, z_off, T);
    else load_g<A, LF>(fg, z_off, T);
    load_links<A, LF>(fu, z_off, W);
    third_row(W);
    prod_ab(Z, T, W);  // g(z) U0_mu(z)
}

// Given S = sum_mu g(y) U0_mu(x)^dagger + g(z) U0_mu(z) (from sum_s, which
// reads it again for the theta of the updated state) and g(x) rows 0,1 in
// Gx: the three subgroup updates of g(x), its store, and the fused-check
// partial sums of the 64-site block blk.
template <typename A, int G, int ALGO, int THETA, int LF, typename SumS, typename Args>
__device__ __forceinline__ void site_tail(const Args &a, const Field &fg, uint32_t site_off, bool active,
                                          int64_t blk, int lane, const SumS &sum_s, M3<A> Gx, Philox &rng,
                                          bool sc1_store = false)
{
    static_assert(THETA == 0 || ALGO != GF_RANDOM, "the random transformation forms no S to check");
    third_row(Gx);

    double th = 0, gf = 0;
    // M = S Gx^dagger; the three subgroup steps read only the entries of M
    // they need and update only the entries a later step reads
    M3<A> M;
    zero(M);
    if (ALGO != GF_RANDOM) {
        M3<A> S;
        sum_s(S);
        mac_adag(M, S, Gx);
    }
    if ((THETA == 1 || THETA == 3) && active) theta_from_m(M, th, gf);  // the state before the update
    if constexpr (THETA == 3) {  // read-only pass: theta only
        th = wave64_sum(th);
        gf = wave64_sum(gf);
        if (lane == 0) {
            a.partial[2 * blk] = th;
            a.partial[2 * blk + 1] = gf;
        }
        return;
    }
    {   // subgroup (0,2): reads M00 M02 M20 M22; M00 M10 M02 M12 M22 are read later
        A q[4];
        quat<A, 0, 2>(M, q);
        gauge_update<A, G, ALGO>(q, a.param, rng);
        const Hd<A> h = su2_dag(q);
        const Cx<A> m00 = col_i(h, M.e[0][0], M.e[0][2]), m10 = col_i(h, M.e[1][0], M.e[1][2]);
        const Cx<A> m02 = col_j(h, M.e[0][0], M.e[0][2]), m12 = col_j(h, M.e[1][0], M.e[1][2]);
        const Cx<A> m22 = col_j(h, M.e[2][0], M.e[2][2]);
        M.e[0][0] = m00; M.e[1][0] = m10; M.e[0][2] = m02; M.e[1][2] = m12; M.e[2][2] = m22;
        left_su2<A, 0, 2>(Gx, q);
    }
    {   // subgroup (1,2): reads M11 M12 M21 M22; M01 M11 are read later
        A q[4];
        quat<A, 1, 2>(M, q);
        gauge_update<A, G, ALGO>(q, a.param, rng);
        const Hd<A> h = su2_dag(q);
        const Cx<A> m01 = col_i(h, M.e[0][1], M.e[0][2]), m11 = col_i(h, M.e[1][1], M.e[1][2]);
        M.e[0][1] = m01; M.e[1][1] = m11;
        left_su2<A, 1, 2>(Gx, q);
    }
    {   // subgroup (0,1): reads M00 M01 M10 M11
        A q[4];
        quat<A, 0, 1>(M, q);
        gauge_update<A, G, ALGO>(q, a.param, rng);
        left_su2<A, 0, 1>(Gx, q);
    }

    project_rows01(Gx);  // keep g unitary at the precision it is stored in
    if (active) {
        if (LF != 2 && sc1_store) store_g<A, LF, LF == 2 ? 0 : kSc1>(fg, site_off, Gx);
        else store_g<A, LF>(fg, site_off, Gx);
    }

    if (THETA == 2 && active) {  // the state after the update, with g(x) as stored
        third_row(Gx);
        M3<A> S;
        sum_s(S);
        zero(M);
        mac_adag(M, S, Gx);
        theta_from_m(M, th, gf);
    }
    if (THETA != 0) {
        th = wave64_sum(th);
        gf = wave64_sum(gf);
        if (lane == 0) {  // indexed by the block's sites, so a timeslice's partials are contiguous
            a.partial[2 * blk] = th;
            a.partial[2 * blk + 1] = gf;
        }
    }
}

// Philox generator of a site (SA, random transformation): key word 0 = global
// id of the reference's (mu = 0, up) thread of this site in its V/2/NSB x
// 8*NSB launch (PhiloxWrapper.hxx:93-110)
template <typename Args> __device__ __forceinline__ void site_rng(const Args &a, int64_t p, int parity, Philox &rng)
{
    const uint32_t tid = (uint32_t)((p / a.nsb) * 8 * a.nsb + p % a.nsb);
    philox_init(rng, tid, a.seed, a.wf && parity ? a.counter_odd : a.counter);
}

// Work mapping 0: a block covers 64 sites with one wave per direction of S
// (mu = 0..3; Coulomb mu = 1..3), each forming the direction-mu terms of its
// 64 sites (four 12-real loads per lane, all coalesced).  The terms go through
// LDS; wave 0 adds them in the order of mu and runs the tail.  gsweep_block
// sweeps the sites of parity `parity` in the 64-site block blk of the lattice
// of a; k_gsweep and k_gsweep_multi choose the block.
template <typename A, int G, int ALGO, int THETA, int LF, typename Args>
__device__ __forceinline__ void gsweep_block(const Args &a, int64_t blk, int parity)
{
    constexpr int kW = gsweep_waves<G>();
    __shared__ A lds[18 * kW * 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mu = wave + (4 - kW);  // one direction per wave
    const int lane = threadIdx.x & 63;
    bool active;
    const int64_t p = block_site(a.geo, blk, lane, active);
    const int64_t site = (active ? p : 0) + (parity ? a.geo.Vh : 0);
    int c[4], par;
    site_coords(a.geo, site, c, par);
    const uint32_t site_off = site_boff<LF>(site);
    const Field fg = g_field<LF>(a);

    M3<A> Gx;
    if (wave == 0) load_g<A, LF>(fg, site_off, Gx);

    Philox rng;
    if (ALGO == GF_SA || ALGO == GF_RANDOM) site_rng(a, p, parity, rng);

    if (ALGO == GF_RANDOM) {
        // the random transformation does not depend on the links
        if (wave != 0) return;
    } else {
        M3<A> X, Z;
        direction_terms<A, LF>(a, fg, mu, c, par, site_off, X, Z);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * (3 * i + j);
                lds[((k + 0) * kW + wave) * 64 + lane] = X.e[j][i].x + Z.e[i][j].x;
                lds[((k + 1) * kW + wave) * 64 + lane] = Z.e[i][j].y - X.e[j][i].y;
            }
        __syncthreads();
        if (wave != 0) return;
    }

    // S = sum_mu g(y) U0_mu(x)^dagger + g(z) U0_mu(z), summed in the order of mu
    const auto sum_s = [&](M3<A> &S) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * (3 * i + j);
                A v[2];
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const A *b = lds + (k + r) * kW * 64 + lane;
                    v[r] = b[0] + b[64];
#pragma unroll
                    for (int w = 2; w < kW; w++) v[r] += b[64 * w];
                }
                S.e[i][j] = cx(v[0], v[1]);
            }
    };
    site_tail<A, G, ALGO, THETA, LF>(a, fg, site_off, active, blk, lane, sum_s, Gx, rng);
}

template <typename A, int G, int ALGO, int THETA, int LF>
__global__ __launch_bounds__(kGBlock) void k_gsweep(GaugeSweepArgs a)
{
    static_assert(LF != 2 || sizeof(A) == 8, "FP64 storage takes FP64 arithmetic");
    if (a.state && a.state->done) return;  // converged: uniform early exit
    if (THETA == 3 && !a.state->near) return;
    int parity = a.parity;
    int64_t blk;
    if (a.wf) {
        if (!wave_block(a, blk, parity)) return;
    } else {
        blk = logical_block(a);
    }
    gsweep_block<A, G, ALGO, THETA, LF>(a, blk, parity);
}

// Several lattices in one launch (the annealing of a batch): their blocks of
// one parity are concatenated in lattice order, and the XCDs take contiguous
// eighths of that sequence (block b runs on XCD b % 8), as remap = 1 does for
// one lattice.  Lattice k's sites are those of the single-lattice sweep with
// its own seed and RNG counter item[k].counter + counter_add + parity: the
// same updates and random numbers.  The lattice's fields as gsweep_block reads
// them, its geometry by reference into the kernel argument (indexed by
// direction, which a local copy would have to keep in scratch memory).
struct MultiView {
    const float *U0;
    float *g;
    const Geom &geo;
    float param;
    uint32_t seed, counter;
    int64_t nsb;
    int wf;
    uint32_t counter_odd;
    double *partial;
};

template <typename A, int G, int ALGO, int LF>
__global__ __launch_bounds__(kGBlock) void k_gsweep_multi(MultiSweepArgs m)
{
    const uint32_t b = blockIdx.x;
    const int32_t L = (int32_t)((b & 7u) * (uint32_t)m.per_xcd + (b >> 3));
    if (L >= m.first[m.n]) return;  // padding of the last XCD's range
    int k = 0;
    while (k + 1 < m.n && m.first[k + 1] <= L) k++;
    const MultiItem &it = m.item[k];
    const MultiView a{it.U0, it.g, it.geo, m.param, it.seed, it.counter + m.counter_add + (uint32_t)m.parity,
                      32, 0, 0, nullptr};
    gsweep_block<A, G, ALGO, 0, LF>(a, L - m.first[k], m.parity);
}

// ------------------------------------------------------ pipelined sweep
// One launch per sweep (su3gf_pipe.h): block b serves queue b % 8 (the XCD it
// runs on) with the item of ticket b / 8 -- the even or the odd sites of one
// 64-site block.  An odd item waits for the flags of the even blocks around it
// (slices t-1..t+1, blocks q-R..q+R), then runs as the two-kernel sweep's odd
// kernel would; an even item runs as the even kernel and then sets its flag to
// the launch's number.  Every site thus sees the neighbour values of the
// two-kernel sweep, and the results are bit-identical.  An item waits only for
// items of lower block index, which the hardware has dispatched before it.
//
// Memory.  An even item writes its g with stores coherent at device scope
// (sc1), waits for them and then sets its flag, also at device scope.  An odd
// item loads its links and its own g (not written in the launch) before it
// waits, and the neighbours' g after seeing the flags: through the XCD's L2
// (where the even items of its own queue, on the same XCD, wrote it) with
// plain loads, and with device-scope loads where the writer may have run on
// another XCD -- the time-direction neighbour in the next queue's or the
// previous queue's timeslice, and any block whose flag says it ran elsewhere
// (never seen on this hardware, where block b runs on XCD b % 8, but handled).
// A CU's L1 holds no stale even g: the only lines of it the CU can have read
// before the flags are those of its own even items, whose stores updated it.
// No device-scope fences: a device-scope acquire (buffer_inv sc1) after the
// wait made the sweep 3x slower (it drops the L2 lines the odd items reuse).

// the XCD this wave runs on (hardware register XCC_ID, bits 3:0)
__device__ __forceinline__ uint32_t xcc_id()
{
    return __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 7u;
}

// iterations (about a microsecond each) one dependency wait spins before it
// gives up and sets the error word
constexpr uint32_t kPipeSpin = 1u << 20;

__device__ __forceinline__ uint32_t *pipe_flags(PipeSync *ps) { return (uint32_t *)(ps + 1); }

// A link's stored words (12 FP32 values, or nine words of 24-bit values),
// loaded ahead of the arithmetic; decode_link gives what load_links gives.
struct LinkWords {
    uint32_t w[12];
};
template <int LF> __device__ __forceinline__ void load_link_words(const Field &f, uint32_t boff, LinkWords &l)
{
    static_assert(LF != 2, "FP32 or 24-bit links");
    if (LF == 1) {
        const uint32_t o = tile_off(boff, 9);
#pragma unroll
        for (int k = 0; k < 9; k++) l.w[k] = __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)o, (int)(k * f.V4), 0);
    } else {
#pragma unroll
        for (int k = 0; k < 12; k++) l.w[k] = __builtin_amdgcn_raw_buffer_load_b32(f.r, (int)boff, (int)(k * f.V4), 0);
    }
}
template <typename A, int LF> __device__ __forceinline__ void decode_link(const LinkWords &l, M3<A> &u)
{
    A v[12];
    if (LF == 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) {  // as load_rows01_q24
            const uint32_t w0 = l.w[3 * k], w1 = l.w[3 * k + 1], w2 = l.w[3 * k + 2];
            const int a = (int)__builtin_amdgcn_sbfe((int)w0, 0, 24);
            const int b = (int)__builtin_amdgcn_sbfe((int)__builtin_amdgcn_alignbit(w1, w0, 24), 0, 24);
            const int c = (int)__builtin_amdgcn_sbfe((int)__builtin_amdgcn_alignbit(w2, w1, 16), 0, 24);
            const int d = (int)w2 >> 8;
            v[4 * k + 0] = (A)a * (A)kInvQ24;
            v[4 * k + 1] = (A)b * (A)kInvQ24;
            v[4 * k + 2] = (A)c * (A)kInvQ24;
            v[4 * k + 3] = (A)d * (A)kInvQ24;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 12; k++) v[k] = (A)__builtin_bit_cast(float, l.w[k]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) u.e[i][j] = cx(v[2 * (j + 3 * i)], v[2 * (j + 3 * i) + 1]);
}

template <typename A, int G, int ALGO, int LF>
__global__ __launch_bounds__(kGBlock) void k_gsweep_pipe(GaugeSweepArgs a)
{
    static_assert(G != GF_MAG, "S a sum of links");
    static_assert(LF != 2, "FP32 or 24-bit links");
    // one wave per direction of S: Coulomb's leaves out the time direction,
    // so its odd sites' neighbours all lie in their own timeslice
    constexpr int kW = gsweep_waves<G>();
    constexpr int kSlices = G == GF_COULOMB ? 1 : 3;  // timeslices an odd block waits on
    __shared__ A lds[18 * kW * 64];
    if (a.state && a.state->done) return;  // converged: uniform early exit
    PipeSync *const ps = a.pipe;
    const PipePlan &pl = a.plan;
    const int queue = blockIdx.x & (kPipeQueues - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mu = wave + (4 - kW);
    const int lane = threadIdx.x & 63;
    const uint32_t launch = a.pipe_launch;  // even flags of this launch hold it
    int odd;
    int64_t idx;
    pipe_item(pl, blockIdx.x >> 3, odd, idx);
    const uint32_t qb = fdiv((uint32_t)idx, a.pipe_tdiv);
    const int t0 = queue * pl.T;
    const int t = t0 + (int)((uint32_t)idx - qb * (uint32_t)pl.T);
    const int64_t blk = (int64_t)t * pl.Q + qb;
    const bool moved = xcc_id() != (uint32_t)queue;

    // what the item reads that the launch does not write: its links and its own g
    bool active;
    const int64_t p = block_site(a.geo, blk, lane, active);
    const int64_t site = p + (odd ? a.geo.Vh : 0);
    int c[4], par;
    site_coords(a.geo, site, c, par);
    const uint32_t site_off = site_boff<LF>(site);
    const uint32_t y_off = site_boff<LF>(forward(a.geo, c, par, mu));
    const uint32_t z_off = site_boff<LF>(backward(a.geo, c, par, mu));
    const Field fg = g_field<LF>(a);
    const Field fu = links_field<LF>(a.U0, a.geo.V, mu);
    LinkWords lx, lz;
    load_link_words<LF>(fu, site_off, lx);  // U0_mu(x)
    load_link_words<LF>(fu, z_off, lz);     // U0_mu(z)
    M3<A> Gx;
    if (wave == 0) load_g<A, LF>(fg, site_off, Gx);

    bool coh = moved;
    if (odd) {
        const int span = 2 * pl.R + 1;
        int seen_moved = 0;
        for (int j = threadIdx.x; j < kSlices * span; j += 64 * kW) {
            const int dtj = j / span;
            int tt = t + dtj - (kSlices / 2);
            if (tt < 0) tt += a.geo.N[0];
            else if (tt >= a.geo.N[0]) tt -= a.geo.N[0];
            int64_t qq = (int64_t)qb + (j - dtj * span) - pl.R;
            if (qq < 0) qq += pl.Q;
            else if (qq >= pl.Q) qq -= pl.Q;
            const uint32_t *f = pipe_flags(ps) + (int64_t)tt * pl.Q + qq;
            for (uint32_t it = 0;; it++) {
                const uint32_t v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((v & ~kPipeMovedBit) == launch) {
                    seen_moved |= (v & kPipeMovedBit) != 0;
                    break;
                }
                if ((it & 63) == 63) {
                    if (it >= kPipeSpin) {
                        __hip_atomic_store(&ps->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    if (__hip_atomic_load(&ps->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        coh = __syncthreads_or(seen_moved) || moved;
    }

    Philox rng;
    if (ALGO == GF_SA) {  // the key of the site's (mu = 0, up) reference thread; odd sites: counter + 1
        const uint32_t tid = (uint32_t)((p / a.nsb) * 8 * a.nsb + p % a.nsb);
        philox_init(rng, tid, a.seed, odd ? a.counter + 1 : a.counter);
    }
    {
        // the direction-mu terms as direction_terms forms them; the neighbours
        // of an odd site are even sites written in this launch, by this XCD
        // except across the queue's first and last timeslices
        const bool coh_y = odd && (coh || (mu == 0 && t == t0 + pl.T - 1));
        const bool coh_z = odd && (coh || (mu == 0 && t == t0));
        M3<A> X, Z, T, W;
        decode_link<A, LF>(lx, W);
        if (coh_y) load_g<A, LF, kSc1>(fg, y_off, T);
        else load_g<A, LF>(fg, y_off, T);
        third_row(T);
        prod_adag(X, W, T);  // U0_mu(x) g(y)^dagger
        if (coh_z) load_g<A, LF, kSc1>(fg, z_off, T);
        else load_g<A, LF>(fg, z_off, T);
        decode_link<A, LF>(lz, W);
        third_row(W);
        prod_ab(Z, T, W);  // g(z) U0_mu(z)
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * (3 * i + j);
                lds[((k + 0) * kW + wave) * 64 + lane] = X.e[j][i].x + Z.e[i][j].x;
                lds[((k + 1) * kW + wave) * 64 + lane] = Z.e[i][j].y - X.e[j][i].y;
            }
    }
    __syncthreads();
    if (wave != 0) return;
    const auto sum_s = [&](M3<A> &S) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * (3 * i + j);
                A v[2];
#pragma unroll
                for (int r = 0; r < 2; r++) {  // in the order of mu, as k_gsweep adds them
                    const A *b = lds + (k + r) * kW * 64 + lane;
                    v[r] = b[0] + b[64];
#pragma unroll
                    for (int w = 2; w < kW; w++) v[r] += b[64 * w];
                }
                S.e[i][j] = cx(v[0], v[1]);
            }
    };
    // even sites: g stored coherently at device scope (read by odd items of this launch)
    site_tail<A, G, ALGO, 0, LF>(a, fg, site_off, active, blk, lane, sum_s, Gx, rng, !odd);
    if (odd) return;
    // the stores are complete before the flag says so
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    if (lane == 0) {
        __hip_atomic_store(pipe_flags(ps) + blk, launch | (moved ? kPipeMovedBit : 0u), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        if (moved) atomicAdd(&ps->moved, 1u);
    }
}

// MAG on the gauge transformation.  MAG's subgroup quaternion is quadratic in
// the links (GaugeFixingSubgroupStep.hxx:354-575), so there is no single matrix
// M to update; instead wave mu keeps its two links U_up = g(x) U0_mu(x)
// g(x+mu)^dagger and U_dn = g(x-mu) U0_mu(x-mu) g(x)^dagger in registers and
// applies every SU(2) step to them (U_up <- G U_up, U_dn <- U_dn G^dagger), as
// the link sweep does.  Per subgroup each wave exchanges its direction's
// (D, Im c, Re c), up and down link added, through LDS; wave 0 adds the four
// in the order mu = 0..3, computes the update (drawing the site's random
// numbers, where the update takes any) and hands it to the other waves
// through LDS, so that they do not repeat its square root, division and two
// inverse square roots.  Wave 0 also updates and stores g(x).
// theta (THETA = 1 before, 2 after the update) comes from the links as in the
// link sweep's MAG check.
template <typename A, int ALGO, int THETA, int LF>
__global__ __launch_bounds__(kGBlock) void k_gsweep_mag(GaugeSweepArgs a)
{
    constexpr int kX = 4 * 3 * 64;  // one exchange: [direction 4][D, Im c, Re c][lane 64]
    __shared__ A lds[2 * kX + 4 * 64 + (THETA ? kThetaVals * 4 * 64 : 0)];
    A *const qlds = lds + 2 * kX;  // the update: [q0..q3][lane 64]
    A *const tlds = qlds + 4 * 64;
    static_assert(LF != 2 || sizeof(A) == 8, "FP64 storage takes FP64 arithmetic");
    if (a.state && a.state->done) return;  // converged: uniform early exit
    if (THETA == 3 && !a.state->near) return;
    const int mu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int parity = a.parity;
    int64_t blk;
    if (a.wf) {
        if (!wave_block(a, blk, parity)) return;
    } else {
        blk = logical_block(a);
    }
    bool active;
    const int64_t p = block_site(a.geo, blk, lane, active);
    const int64_t site = (active ? p : 0) + (parity ? a.geo.Vh : 0);
    int c[4], par;
    site_coords(a.geo, site, c, par);
    const uint32_t site_off = site_boff<LF>(site);
    const Field fg = g_field<LF>(a);

    M3<A> Gx;
    load_g<A, LF>(fg, site_off, Gx);
    third_row(Gx);
    Philox rng;  // wave 0 computes the updates
    if ((ALGO == GF_SA || ALGO == GF_SR || ALGO == GF_RANDOM) && mu == 0) site_rng(a, p, parity, rng);

    M3<A> up, dn;
    {
        M3<A> X, Z;
        direction_terms<A, LF>(a, fg, mu, c, par, site_off, X, Z);
        prod_ab(up, Gx, X);    // g(x) U0_mu(x) g(y)^dagger
        prod_adag(dn, Z, Gx);  // g(z) U0_mu(z) g(x)^dagger
    }

    double th = 0, gf = 0;
    auto theta_pass = [&]() {
        A v[kThetaVals];
        theta_contrib<A, GF_MAG>(up, dn, active, v);
#pragma unroll
        for (int k = 0; k < kThetaVals; k++) tlds[(k * 4 + mu) * 64 + lane] = v[k];
        __syncthreads();
        if (mu == 0) {
            double s3[kThetaVals];
#pragma unroll
            for (int k = 0; k < kThetaVals; k++) {
                const A *b = tlds + k * 4 * 64 + lane;
                s3[k] = (double)(((b[0] + b[64]) + b[128]) + b[192]);
            }
            if (active) theta_site<GF_MAG>(s3, th, gf);
        }
    };
    if (THETA == 1 || THETA == 3) theta_pass();
    if constexpr (THETA == 3) {  // read-only pass: theta only
        if (mu != 0) return;
        th = wave64_sum(th);
        gf = wave64_sum(gf);
        if (lane == 0) {
            a.partial[2 * blk] = th;
            a.partial[2 * blk + 1] = gf;
        }
        return;
    }

    int buf = 0;
    auto subgroup = [&](auto I_, auto J_) {
        constexpr int I = decltype(I_)::value, J = decltype(J_)::value;
        // each wave exchanges its direction's share: its up and down links'
        // terms added (the reference's order of the eight terms is not kept)
        A vu[4], vd[4], pd[4];
        link_contrib<A, GF_MAG, I, J, true>(up, vu);
        link_contrib<A, GF_MAG, I, J, false>(dn, vd);
        direction_share<A, GF_MAG>(vu, vd, false, pd);
        A *x = lds + buf * kX;
#pragma unroll
        for (int k = 0; k < 3; k++) x[(mu * 3 + k) * 64 + lane] = pd[k];
        __syncthreads();
        buf ^= 1;  // the next exchange uses the other buffer
        A q[4];
        if (mu == 0) {
            A cs[4][4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
#pragma unroll
                for (int k = 0; k < 3; k++) cs[m][k] = x[(m * 3 + k) * 64 + lane];
                cs[m][3] = 0;
            }
            site_update_dirs<A, GF_MAG, ALGO, false>(cs, q, a.param, rng);
#pragma unroll
            for (int k = 0; k < 4; k++) qlds[k * 64 + lane] = q[k];
        }
        // the other waves read it after this barrier; wave 0 writes the next
        // one only after the next exchange's barrier, which they reach after
        // reading this one
        __syncthreads();
        if (mu != 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) q[k] = qlds[k * 64 + lane];
        }
        left_su2<A, I, J>(up, q);
        const Hd<A> h = su2_dag(q);
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const Cx<A> ci = col_i(h, dn.e[r][I], dn.e[r][J]), cj = col_j(h, dn.e[r][I], dn.e[r][J]);
            dn.e[r][I] = ci;
            dn.e[r][J] = cj;
        }
        if (mu == 0) left_su2<A, I, J>(Gx, q);
    };
    subgroup(std::integral_constant<int, 0>(), std::integral_constant<int, 2>());
    subgroup(std::integral_constant<int, 1>(), std::integral_constant<int, 2>());
    subgroup(std::integral_constant<int, 0>(), std::integral_constant<int, 1>());

    if (THETA == 2) theta_pass();
    if (mu != 0) return;
    project_rows01(Gx);
    if (active) store_g<A, LF>(fg, site_off, Gx);
    if (THETA != 0) {
        th = wave64_sum(th);
        gf = wave64_sum(gf);
        if (lane == 0) {
            a.partial[2 * blk] = th;
            a.partial[2 * blk + 1] = gf;
        }
    }
}

// blocks of a launch: one per logical block, or a wavefront launch's even and odd grids
inline int64_t gsweep_grid(const GaugeSweepArgs &a)
{
    if (!a.wf) return sweep_grid(a.geo);
    return a.wf_even_grid + 8 * ((a.wf_odd_n + a.wf_odd1_n + 7) / 8);
}

template <typename A, int ALGO, int THETA, int LF> void launch_k_mag(const GaugeSweepArgs &a, hipStream_t s)
{
    const dim3 grid((unsigned)gsweep_grid(a)), block(kGBlock);
    hipLaunchKernelGGL((k_gsweep_mag<A, ALGO, THETA, LF>), grid, block, 0, s, a);
}

template <typename A, int LF> hipError_t launch_mag_lf(int algo, int theta, const GaugeSweepArgs &a, hipStream_t s)
{
    if constexpr (LF == 2) {  // the FP64 phase: OR with the check, or the read-only theta pass
        if constexpr (sizeof(A) != 8) {
            return hipErrorInvalidValue;
        } else {
            if (algo != GF_OR) return hipErrorInvalidValue;
            if (theta == 1) launch_k_mag<A, GF_OR, 1, 2>(a, s);
            else if (theta == 2) launch_k_mag<A, GF_OR, 2, 2>(a, s);
            else if (theta == 3) launch_k_mag<A, GF_OR, 3, 2>(a, s);
            else launch_k_mag<A, GF_OR, 0, 2>(a, s);
            return hipGetLastError();
        }
    } else {
        if (theta == 3) return hipErrorInvalidValue;
        if (algo == GF_OR) {
            if (theta == 1) launch_k_mag<A, GF_OR, 1, LF>(a, s);
            else if (theta == 2) launch_k_mag<A, GF_OR, 2, LF>(a, s);
            else launch_k_mag<A, GF_OR, 0, LF>(a, s);
        } else if (theta != 0) {
            return hipErrorInvalidValue;
        } else if (algo == GF_MICRO) {
            launch_k_mag<A, GF_MICRO, 0, LF>(a, s);
        } else if (algo == GF_SA) {
            launch_k_mag<A, GF_SA, 0, LF>(a, s);
        } else if (algo == GF_SR) {
            launch_k_mag<A, GF_SR, 0, LF>(a, s);
        } else if (algo == GF_RANDOM) {  // MAG's random step keeps q3 = 0 (GaugeFixingSubgroupStep.hxx:566-575)
            launch_k_mag<A, GF_RANDOM, 0, LF>(a, s);
        } else {
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
}

template <typename A, int G, int ALGO, int THETA, int LF>
void launch_k(const GaugeSweepArgs &a, hipStream_t s)
{
    const dim3 grid((unsigned)gsweep_grid(a)), block(64 * gsweep_waves<G>());
    hipLaunchKernelGGL((k_gsweep<A, G, ALGO, THETA, LF>), grid, block, 0, s, a);
}

template <typename A, int G> hipError_t launch_g(int algo, int theta, const GaugeSweepArgs &a, hipStream_t s)
{
    if (a.links64) {  // FP64 links and g: the FP64 phase, OR with the check or the read-only theta pass
        if constexpr (sizeof(A) == 8) {
            if (algo != GF_OR) return hipErrorInvalidValue;
            if (theta == 1) launch_k<A, G, GF_OR, 1, 2>(a, s);
            else if (theta == 2) launch_k<A, G, GF_OR, 2, 2>(a, s);
            else if (theta == 3) launch_k<A, G, GF_OR, 3, 2>(a, s);
            else launch_k<A, G, GF_OR, 0, 2>(a, s);
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    if (theta == 3) return hipErrorInvalidValue;
    if (a.links24) {  // fixed-point links: OR of the mixed driver, SA and microcanonical of the annealing
        if (algo == GF_OR) {
            if (theta == 1) launch_k<A, G, GF_OR, 1, 1>(a, s);
            else if (theta == 2) launch_k<A, G, GF_OR, 2, 1>(a, s);
            else launch_k<A, G, GF_OR, 0, 1>(a, s);
        } else if (theta != 0) {
            return hipErrorInvalidValue;
        } else if (algo == GF_MICRO) {
            launch_k<A, G, GF_MICRO, 0, 1>(a, s);
        } else if (algo == GF_SA) {
            launch_k<A, G, GF_SA, 0, 1>(a, s);
        } else {
            return hipErrorInvalidValue;
        }
    } else if (algo == GF_OR) {
        if (theta == 1) launch_k<A, G, GF_OR, 1, 0>(a, s);
        else if (theta == 2) launch_k<A, G, GF_OR, 2, 0>(a, s);
        else launch_k<A, G, GF_OR, 0, 0>(a, s);
    } else if (theta != 0) {
        return hipErrorInvalidValue;
    } else if (algo == GF_MICRO) {
        launch_k<A, G, GF_MICRO, 0, 0>(a, s);
    } else if (algo == GF_SA) {
        launch_k<A, G, GF_SA, 0, 0>(a, s);
    } else if (algo == GF_RANDOM) {
        launch_k<A, G, GF_RANDOM, 0, 0>(a, s);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// links (type R, layout SI) -> rows 0,1 in the 24-bit fixed-point form, in
// tiles of 64 sites per direction: word w of link (i, mu) at
// P[mu links24_words(V) / 4 + (i & ~63) 9 + 64 w + (i & 63)]
template <typename R, int SI> __global__ __launch_bounds__(256) void k_pack24(const R *U, uint32_t *P, int64_t V)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= 4 * V) return;
    const int mu = (int)(t / V);
    const int64_t i = t - (int64_t)mu * V;
    const LinkRef<R> l = link_ref(const_cast<R *>(U), V, SI, i, mu);
    M3<R> u;
    load12(l, u);
    uint32_t q[12];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int k = 2 * (c + 3 * r);
            q[k] = (uint32_t)(int)rint(fmin(fmax((double)u.e[r][c].x, -1.0), 1.0) * kQ24) & 0xFFFFFFu;
            q[k + 1] = (uint32_t)(int)rint(fmin(fmax((double)u.e[r][c].y, -1.0), 1.0) * kQ24) & 0xFFFFFFu;
        }
    uint32_t *o = P + links24_words(V) / 4 * mu + (i & ~int64_t(63)) * 9 + (i & 63);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t a = q[4 * k], b = q[4 * k + 1], c = q[4 * k + 2], d = q[4 * k + 3];
        o[(3 * k) * 64] = a | (b << 24);
        o[(3 * k + 1) * 64] = (b >> 8) | (c << 16);
        o[(3 * k + 2) * 64] = (c >> 16) | (d << 8);
    }
}

template <typename A> hipError_t launch_a(int gauge, int algo, int theta, const GaugeSweepArgs &a, hipStream_t s)
{
    if (gauge == GF_MAG) {
        if (a.links64) return launch_mag_lf<A, 2>(algo, theta, a, s);
        return a.links24 ? launch_mag_lf<A, 1>(algo, theta, a, s) : launch_mag_lf<A, 0>(algo, theta, a, s);
    }
    if (gauge == GF_LANDAU) return launch_g<A, GF_LANDAU>(algo, theta, a, s);
    if (gauge == GF_COULOMB) return launch_g<A, GF_COULOMB>(algo, theta, a, s);
    if (gauge == GF_U1XU1) return launch_g<A, GF_U1XU1>(algo, theta, a, s);
    return hipErrorInvalidValue;
}

}  // namespace


namespace {
// 24-bit links -> FP32 rows 0,1 (compact layout), through the sweep kernels' decode
__global__ __launch_bounds__(256) void k_unpack24(const uint32_t *P, float *U32, int64_t V)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= 4 * V) return;
    const int mu = (int)(t / V);
    const int64_t i = t - (int64_t)mu * V;
    const Field f = links_field<1>((const float *)P, V, mu);
    M3<float> u;
    load_rows01_q24<float>(f, tile_off((uint32_t)i * 4u, 9), u);
    float *o = U32 + (int64_t)12 * V * mu + i;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            o[(2 * (c + 3 * r)) * V] = u.e[r][c].x;
            o[(2 * (c + 3 * r) + 1) * V] = u.e[r][c].y;
        }
}
}  // namespace

hipError_t launch_unpack24(const uint32_t *P, float *U32, int64_t V, hipStream_t s)
{
    if (links24_words(V) >= (int64_t(1) << 31)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((4 * V + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_unpack24, grid, block, 0, s, P, U32, V);
    return hipGetLastError();
}

hipError_t launch_pack24(int ti, int si, const void *U, uint32_t *P, int64_t V, hipStream_t s)
{
    if (links24_words(V) >= (int64_t(1) << 31)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((4 * V + 255) / 256)), block(256);
    const double *d = (const double *)U;
    const float *f = (const float *)U;
    if (ti == GF_F64 && si == 18) hipLaunchKernelGGL((k_pack24<double, 18>), grid, block, 0, s, d, P, V);
    else if (ti == GF_F64 && si == 12) hipLaunchKernelGGL((k_pack24<double, 12>), grid, block, 0, s, d, P, V);
    else if (ti == GF_F32 && si == 12) hipLaunchKernelGGL((k_pack24<float, 12>), grid, block, 0, s, f, P, V);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_gsweep(int gauge, int algo, int theta, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s)
{
    if (!gauge_sweep_fits(a.geo.V)) return hipErrorInvalidValue;
    if (a.links64 && (!fp64_arith || !gauge_sweep64_fits(a.geo.V))) return hipErrorInvalidValue;
    if (theta == 3 && !a.state) return hipErrorInvalidValue;
    GaugeSweepArgs b = a;
    const int64_t half_slice = a.geo.Vh / a.geo.N[0];  // sites of one parity in a timeslice
    if (b.remap && a.geo.N[0] % 8 == 0 && half_slice % 64 == 0) {
        b.remap = 2;
        b.slice_blocks = half_slice / 64;
        b.slices_per_xcd = a.geo.N[0] / 8;
        if (b.tpass <= 0 || b.slices_per_xcd % b.tpass != 0) b.tpass = b.slices_per_xcd;
        b.tpass_div = make_fastdiv((uint32_t)b.tpass);
        b.pass_div = make_fastdiv((uint32_t)(b.tpass * b.slice_blocks));
    } else if (b.remap && sweep_grid(a.geo) % 8 == 0) {
        b.remap = 1;
    } else {
        b.remap = 0;
    }
    if (b.nsb <= 0) b.nsb = 32;
    b.wf = 0;
    return fp64_arith ? launch_a<double>(gauge, algo, theta, b, s) : launch_a<float>(gauge, algo, theta, b, s);
}

hipError_t launch_gsweep_wavefront(int gauge, int algo, bool fp64_arith, const GaugeSweepArgs &a, int slab,
                                   hipStream_t s)
{
    if (slab == kPipelinedSlab) return launch_gsweep_pipe(gauge, algo, fp64_arith, a, s);
    if (!gauge_sweep_fits(a.geo.V)) return hipErrorInvalidValue;
    if (a.links64 && (!fp64_arith || !gauge_sweep64_fits(a.geo.V))) return hipErrorInvalidValue;
    const int N = a.geo.N[0];
    const int64_t grid = sweep_grid(a.geo);
    // blocks per timeslice: whole 64-site blocks, or a Coulomb lattice's
    // per-slice blocks (each slice's last block padded)
    if (N < 4 || slab < 1 || grid % N != 0 || (!a.geo.slice_bps && (a.geo.Vh / N) % 64 != 0))
        return hipErrorNotSupported;
    const int64_t bps = grid / N;
    const int K = (N + slab - 1) / slab;
    auto e = [&](int k) { return (int64_t)std::min(k * slab, N); };
    GaugeSweepArgs b = a;
    if (b.nsb <= 0) b.nsb = 32;
    b.wf = 1;
    b.parity = 0;
    b.counter_odd = a.counter + 1;
    for (int k = 0; k <= K; k++) {
        // even timeslices [e(k), e(k+1)); odd timeslices [max(1, e(k-1) - 1), e(k) - 1),
        // and in the last launch the remaining ones and timeslice 0
        const int64_t even_lo = k < K ? e(k) : 0, even_hi = k < K ? e(k + 1) : 0;
        int64_t odd_lo = 0, odd_hi = 0;
        if (k > 0) {
            odd_lo = std::max<int64_t>(1, e(k - 1) - 1);
            odd_hi = k < K ? e(k) - 1 : N;
            if (odd_hi < odd_lo) odd_hi = odd_lo;
        }
        b.wf_even0 = even_lo * bps;
        b.wf_even_n = (even_hi - even_lo) * bps;
        b.wf_even_grid = 8 * ((b.wf_even_n + 7) / 8);
        b.wf_odd0 = odd_lo * bps;
        b.wf_odd_n = (odd_hi - odd_lo) * bps;
        b.wf_odd1 = 0;
        b.wf_odd1_n = k == K ? bps : 0;
        if (gsweep_grid(b) == 0) continue;
        const hipError_t err = fp64_arith ? launch_a<double>(gauge, algo, 0, b, s) : launch_a<float>(gauge, algo, 0, b, s);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

namespace {
template <typename A, int G, int ALGO, int LF> hipError_t launch_k_pipe(const GaugeSweepArgs &a, hipStream_t s)
{
    const dim3 grid((unsigned)(kPipeQueues * 2 * a.plan.n)), block(64 * gsweep_waves<G>());
    hipLaunchKernelGGL((k_gsweep_pipe<A, G, ALGO, LF>), grid, block, 0, s, a);
    return hipGetLastError();
}
template <int G, int LF> hipError_t launch_pipe_algo(int algo, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s)
{
    if (fp64_arith) return algo == GF_OR ? launch_k_pipe<double, G, GF_OR, LF>(a, s) : hipErrorNotSupported;
    if (algo == GF_OR) return launch_k_pipe<float, G, GF_OR, LF>(a, s);
    if (algo == GF_MICRO) return launch_k_pipe<float, G, GF_MICRO, LF>(a, s);
    if (algo == GF_SA) return launch_k_pipe<float, G, GF_SA, LF>(a, s);
    return hipErrorNotSupported;
}
template <int G> hipError_t launch_pipe_lf(int algo, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s)
{
    return a.links24 ? launch_pipe_algo<G, 1>(algo, fp64_arith, a, s) : launch_pipe_algo<G, 0>(algo, fp64_arith, a, s);
}
}  // namespace

namespace {
std::atomic<long long> g_pipe_launches{0};
}
long long pipe_launches() { return g_pipe_launches.load(); }

bool gsweep_pipe_applies(const Geom &geo, int gauge, bool links64)
{
    if (links64 || gauge == GF_MAG) return false;
    // Coulomb's per-timeslice blocks: whole 64-site blocks only (then they are the plain ones)
    if (geo.slice_bps != 0 && geo.slice_bps * 64 != geo.slice_half) return false;
    if (!gauge_sweep_fits(geo.V)) return false;
    PipePlan p;
    return pipe_plan(geo.N, p) && kPipeQueues * 2 * p.n < (int64_t(1) << 31);
}

hipError_t launch_gsweep_pipe(int gauge, int algo, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s)
{
    if (!a.pipe || !gsweep_pipe_applies(a.geo, gauge, a.links64)) return hipErrorNotSupported;
    // FP64 arithmetic: the OR sweeps of the mixed driver's stage B only
    if (algo != GF_OR && algo != GF_MICRO && algo != GF_SA) return hipErrorNotSupported;
    if (fp64_arith && algo != GF_OR) return hipErrorNotSupported;
    GaugeSweepArgs b = a;
    if (!a.pipe_count || !pipe_plan(a.geo.N, b.plan, a.pipe_slack)) return hipErrorNotSupported;
    // launch numbers 1, 2, ... below the flags' bit 31 (the flags start at 0)
    if (*a.pipe_count >= kPipeMovedBit - 1) return hipErrorNotSupported;
    b.pipe_launch = ++*a.pipe_count;
    b.pipe_tdiv = make_fastdiv((uint32_t)b.plan.T);
    b.wf = 0;
    b.remap = 0;
    if (b.nsb <= 0) b.nsb = 32;
    const hipError_t e = gauge == GF_LANDAU    ? launch_pipe_lf<GF_LANDAU>(algo, fp64_arith, b, s)
                         : gauge == GF_COULOMB ? launch_pipe_lf<GF_COULOMB>(algo, fp64_arith, b, s)
                                               : launch_pipe_lf<GF_U1XU1>(algo, fp64_arith, b, s);
    if (e == hipSuccess) g_pipe_launches++;
    return e;
}

}  // namespace su3gf

namespace su3gf {
namespace {
template <int G, int LF> hipError_t launch_multi_algo(int algo, const MultiSweepArgs &m, unsigned grid, hipStream_t s)
{
    const dim3 block(64 * gsweep_waves<G>());
    if (algo == GF_SA) hipLaunchKernelGGL((k_gsweep_multi<float, G, GF_SA, LF>), dim3(grid), block, 0, s, m);
    else if (algo == GF_MICRO) hipLaunchKernelGGL((k_gsweep_multi<float, G, GF_MICRO, LF>), dim3(grid), block, 0, s, m);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int G> hipError_t launch_multi_g(int algo, int links24, const MultiSweepArgs &m, unsigned grid, hipStream_t s)
{
    return links24 ? launch_multi_algo<G, 1>(algo, m, grid, s) : launch_multi_algo<G, 0>(algo, m, grid, s);
}
}  // namespace

hipError_t launch_gsweep_multi(int gauge, int algo, int links24, MultiSweepArgs &m, hipStream_t s)
{
    if (m.n < 1 || m.n > kMultiMax || (m.parity & ~1)) return hipErrorInvalidValue;
    int64_t total = 0;
    for (int k = 0; k < m.n; k++) {
        const Geom &g = m.item[k].geo;
        if (!gauge_sweep_fits(g.V) || g.slice_bps != 0 || !m.item[k].U0 || !m.item[k].g) return hipErrorInvalidValue;
        m.first[k] = (int32_t)total;
        total += sweep_grid(g);
        if (total >= (int64_t(1) << 30)) return hipErrorInvalidValue;
    }
    m.first[m.n] = (int32_t)total;
    m.per_xcd = (int32_t)((total + 7) / 8);
    const unsigned grid = 8u * (unsigned)m.per_xcd;
    if (gauge == GF_LANDAU) return launch_multi_g<GF_LANDAU>(algo, links24, m, grid, s);
    if (gauge == GF_COULOMB) return launch_multi_g<GF_COULOMB>(algo, links24, m, grid, s);
    if (gauge == GF_U1XU1) return launch_multi_g<GF_U1XU1>(algo, links24, m, grid, s);
    return hipErrorInvalidValue;
}

}  // namespace su3gf
