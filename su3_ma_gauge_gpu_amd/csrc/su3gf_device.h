// su3gf_device.h -- register-resident SU(3) / SU(2) algebra, link addressing,
// lattice geometry and the Philox4x32-10 generator for the gfx950 sweep kernels.
//
// Everything here is __device__ code operating on values held in VGPRs.  The
// formulas follow the reference's lattice classes (paths relative to its src/):
//   quaternion of a 2x2 subgroup  lattice/SU3.hxx:193-217
//   quaternion <-> SU(2) matrix   lattice/Quaternion.hxx:100-132
//   left / right subgroup product lattice/SU3.hxx:682-714
//   third-row reconstruction      lattice/SU3.hxx:655-664
//   Gram-Schmidt projection       lattice/SU3.hxx:536-583
//   parity-split site index       lattice/SiteIndex.hxx:289-313, SiteCoord.hxx:168-196
//   link component addressing     lattice/access_pattern/GpuPattern.hxx:53-55
//   Philox4x32-10 and u01         external/Random123/philox.h:206-326, u01.h:111,129
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace su3gf {

// ----------------------------------------------------------------- complex
template <typename R> struct Cx { R x, y; };

template <typename R> __device__ __forceinline__ Cx<R> cx(R x, R y) { Cx<R> c; c.x = x; c.y = y; return c; }
template <typename R> __device__ __forceinline__ Cx<R> conj(Cx<R> a) { return cx(a.x, -a.y); }
template <typename R> __device__ __forceinline__ Cx<R> operator+(Cx<R> a, Cx<R> b) { return cx(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ Cx<R> operator-(Cx<R> a, Cx<R> b) { return cx(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ Cx<R> operator*(Cx<R> a, Cx<R> b)
{
    return cx(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __device__ __forceinline__ Cx<R> operator*(Cx<R> a, R s) { return cx(a.x * s, a.y * s); }
template <typename R> __device__ __forceinline__ R abs2(Cx<R> a) { return a.x * a.x + a.y * a.y; }

// a * conj(b) without forming conj(b)
template <typename R> __device__ __forceinline__ Cx<R> mul_conj(Cx<R> a, Cx<R> b)
{
    return cx(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// ------------------------------------------------------------------- math
__device__ __forceinline__ float rsq(float v) { return __frsqrt_rn(v); }
__device__ __forceinline__ double rsq(double v) { return rsqrt(v); }
__device__ __forceinline__ float sqrt_(float v) { return sqrtf(v); }
__device__ __forceinline__ double sqrt_(double v) { return sqrt(v); }

// ------------------------------------------------------------ 3x3 matrix
template <typename R> struct M3 { Cx<R> e[3][3]; };

// third row = conj(row0 x row1), normalised (SU3.hxx:655-664)
template <typename R> __device__ __forceinline__ void reconstruct(M3<R> &u)
{
    Cx<R> a = conj(u.e[0][1] * u.e[1][2] - u.e[0][2] * u.e[1][1]);
    Cx<R> b = conj(u.e[0][2] * u.e[1][0] - u.e[0][0] * u.e[1][2]);
    Cx<R> c = conj(u.e[0][0] * u.e[1][1] - u.e[0][1] * u.e[1][0]);
    R n = rsq(abs2(a) + abs2(b) + abs2(c));
    u.e[2][0] = a * n;
    u.e[2][1] = b * n;
    u.e[2][2] = c * n;
}

// Gram-Schmidt on rows 0,1 then third row (SU3.hxx:536-583)
template <typename R> __device__ __forceinline__ void project(M3<R> &u)
{
    R n0 = rsq(abs2(u.e[0][0]) + abs2(u.e[0][1]) + abs2(u.e[0][2]));
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[0][k] = u.e[0][k] * n0;
    Cx<R> sp = mul_conj(u.e[1][0], u.e[0][0]) + mul_conj(u.e[1][1], u.e[0][1]) + mul_conj(u.e[1][2], u.e[0][2]);
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[1][k] = u.e[1][k] - u.e[0][k] * sp;
    R n1 = rsq(abs2(u.e[1][0]) + abs2(u.e[1][1]) + abs2(u.e[1][2]));
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[1][k] = u.e[1][k] * n1;
    reconstruct(u);
}

// q(M) for subgroup (i,j): (Re Mii + Re Mjj, Im Mij + Im Mji, Re Mij - Re Mji, Im Mii - Im Mjj)
template <typename R, int I, int J> __device__ __forceinline__ void quat(const M3<R> &u, R q[4])
{
    q[0] = u.e[I][I].x + u.e[J][J].x;
    q[1] = u.e[I][J].y + u.e[J][I].y;
    q[2] = u.e[I][J].x - u.e[J][I].x;
    q[3] = u.e[I][I].y - u.e[J][J].y;
}

// rows I,J <- G(q) rows I,J with G = [[q0+iq3, q2+iq1], [-q2+iq1, q0-iq3]]
template <typename R, int I, int J> __device__ __forceinline__ void left_mult(M3<R> &u, const R q[4])
{
    const Cx<R> g00 = cx(q[0], q[3]), g01 = cx(q[2], q[1]), g10 = cx(-q[2], q[1]), g11 = cx(q[0], -q[3]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Cx<R> ik = g00 * u.e[I][k] + g01 * u.e[J][k];
        Cx<R> jk = g10 * u.e[I][k] + g11 * u.e[J][k];
        u.e[I][k] = ik;
        u.e[J][k] = jk;
    }
}

// cols I,J <- cols I,J G(q)^dagger
template <typename R, int I, int J> __device__ __forceinline__ void right_mult_dag(M3<R> &u, const R q[4])
{
    // G(q)^dagger = G(q0, -q1, -q2, -q3)
    const Cx<R> h00 = cx(q[0], -q[3]), h01 = cx(-q[2], -q[1]), h10 = cx(q[2], -q[1]), h11 = cx(q[0], q[3]);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        Cx<R> ki = h00 * u.e[k][I] + h10 * u.e[k][J];
        Cx<R> kj = h01 * u.e[k][I] + h11 * u.e[k][J];
        u.e[k][I] = ki;
        u.e[k][J] = kj;
    }
}

// -------------------------------------------------------- link addressing
// Links live in a component-major, parity-split array: component `comp` of
// link (site idx, direction mu) is at base[idx + V*(comp + stride*mu)].
// comp = c + 2*(j + 3*i) for row i, column j, c = 0 real / 1 imaginary.
// stride = 18 is the reference's GpuPattern layout; stride = 12 the compact
// layout that stores rows 0,1 only.
template <typename R> struct LinkRef {
    R *p;        // &base[idx + V*stride*mu]
    int64_t V;
};

// stride = kTiled12: the drivers' internal FP64 layout, rows 0,1 of each
// direction in tiles of 64 sites -- component c of the link (idx, mu) at
// base[mu * tiled12_reals(V) + (idx - idx % 64) * 12 + 64 * c + idx % 64] -- so
// that a block's 64 sites read one contiguous 6 KB tile per direction
// instead of twelve 512-byte pieces V * 8 bytes apart.
constexpr int kTiled12 = -12;
__host__ __device__ inline int64_t tiled12_reals(int64_t V) { return 12 * 64 * ((V + 63) / 64); }
// The FP64 gauge-transformation sweep addresses one direction of such links,
// and g, as one buffer of 32-bit byte offsets: 12 doubles per site in whole
// tiles stay below 2^31 bytes up to 22,369,600 sites (68^4 fits).
__host__ __device__ inline bool gauge_sweep64_fits(int64_t V) { return 8 * tiled12_reals(V) < (int64_t(1) << 31); }

template <typename R> __device__ __forceinline__ LinkRef<R> link_ref(R *base, int64_t V, int stride, int64_t idx, int mu)
{
    LinkRef<R> l;
    if (stride == kTiled12) {
        l.p = base + mu * tiled12_reals(V) + (idx & ~int64_t(63)) * 12 + (idx & 63);
        l.V = 64;
        return l;
    }
    l.p = base + (int64_t)stride * V * mu + idx;
    l.V = V;
    return l;
}

// The gauge transformation g of the mixed-precision and annealing drivers:
// FP32 rows 0,1 per site in tiles of 64 sites, component c (0..11) of site i
// at g[(i & ~63) * 12 + c * 64 + (i & 63)].  A block of 64 sites reads one
// contiguous 3 KB tile instead of twelve planes V floats apart.  The buffer
// holds gauge_floats(V) floats (whole tiles).
__host__ __device__ inline int64_t gauge_floats(int64_t V) { return 12 * ((V + 63) / 64) * 64; }
// The gauge-transformation sweep addresses g, and each direction of the links,
// as one buffer with 32-bit byte offsets: 12 FP32 words per site in whole
// tiles must stay below 2^31 bytes, i.e. at most 44,739,200 sites (80^4 fits,
// 84^4 does not).
__host__ __device__ inline bool gauge_sweep_fits(int64_t V) { return 4 * gauge_floats(V) < (int64_t(1) << 31); }
__device__ __forceinline__ LinkRef<float> gauge_ref(float *g, int64_t i)
{
    LinkRef<float> l;
    l.p = g + (i & ~int64_t(63)) * 12 + (i & 63);
    l.V = 64;  // component stride
    return l;
}

template <typename R> __device__ __forceinline__ void load12(const LinkRef<R> &l, M3<R> &u)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            u.e[i][j] = cx(l.p[l.V * comp], l.p[l.V * (comp + 1)]);
        }
}

template <typename R> __device__ __forceinline__ void store12(const LinkRef<R> &l, const M3<R> &u)
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            l.p[l.V * comp] = u.e[i][j].x;
            l.p[l.V * (comp + 1)] = u.e[i][j].y;
        }
}

template <typename R> __device__ __forceinline__ void store_row2(const LinkRef<R> &l, const M3<R> &u)
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int comp = 2 * (j + 6);
        l.p[l.V * comp] = u.e[2][j].x;
        l.p[l.V * (comp + 1)] = u.e[2][j].y;
    }
}

// ----------------------------------------------------------- geometry
// Division by a runtime constant d < 2^31 for dividends n < 2^31:
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = 2^32 (2^s - d)/d + 1.
struct FastDiv {
    uint32_t d, m, s;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f)
{
    return (__umulhi(n, f.m) + n) >> f.s;
}

// the multiplier and shift of fdiv for divisor d (1 <= d < 2^31)
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    f.s = 0;
    while ((1ull << f.s) < d) f.s++;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << f.s) - d)) / d + 1);
    return f;
}

struct Geom {
    int N[4];          // {Nt, Nx, Ny, Nz}
    int64_t V, Vh;
    FastDiv fx, fy, fz;
    // 0, or (Coulomb checked per timeslice) the blocks of 64 sites that each
    // timeslice's slice_half sites of one parity get: no block spans two slices
    int64_t slice_half, slice_bps;
    FastDiv slice_div;  // division by slice_bps
};

// Blocks of a parity sweep: 64 consecutive sites of the parity each, or
// slice_bps per timeslice (the last of a slice padded with idle lanes).
__host__ __device__ inline int64_t sweep_grid(const Geom &g)
{
    return g.slice_bps ? (int64_t)g.N[0] * g.slice_bps : (g.Vh + 63) / 64;
}

// Index p (within the parity) of the site of lane `lane` of block b, and
// whether the lane has a site; an idle lane's p is not a site of the block.
__device__ __forceinline__ int64_t block_site(const Geom &g, int64_t b, int lane, bool &active)
{
    if (g.slice_bps == 0) {
        const int64_t p = b * 64 + lane;
        active = p < g.Vh;
        return p;
    }
    const int64_t t = fdiv((uint32_t)b, g.slice_div);  // block indices stay below 2^31
    const int64_t k = (b - t * g.slice_bps) * 64 + lane;
    active = k < g.slice_half;
    return t * g.slice_half + k;
}

// parity-split index -> coordinates (SiteCoord.hxx:168-196)
__device__ __forceinline__ void site_coords(const Geom &g, int64_t idx, int c[4], int &par)
{
    par = idx >= g.Vh;
    uint32_t n = (uint32_t)(2 * (idx - (par ? g.Vh : 0)));
    uint32_t q = fdiv(n, g.fz);
    c[3] = (int)(n - q * (uint32_t)g.N[3]);
    n = q;
    q = fdiv(n, g.fy);
    c[2] = (int)(n - q * (uint32_t)g.N[2]);
    n = q;
    q = fdiv(n, g.fx);
    c[1] = (int)(n - q * (uint32_t)g.N[1]);
    c[0] = (int)q;
    if (((c[0] + c[1] + c[2] + c[3]) & 1) != par) c[3]++;
}

// coordinates of a site of parity `par` -> parity-split index (SiteIndex.hxx:289-313)
__device__ __forceinline__ int64_t site_index(const Geom &g, const int c[4], int par)
{
    int64_t n = (((int64_t)c[0] * g.N[1] + c[1]) * g.N[2] + c[2]) * g.N[3] + c[3];
    return (n >> 1) + (par ? g.Vh : 0);
}

// backward neighbour x - mu of a site with coordinates c and parity par
__device__ __forceinline__ int64_t backward(const Geom &g, const int c[4], int par, int mu)
{
    int d[4] = {c[0], c[1], c[2], c[3]};
    d[mu] = (d[mu] == 0) ? g.N[mu] - 1 : d[mu] - 1;
    return site_index(g, d, par ^ 1);
}

// forward neighbour x + mu
__device__ __forceinline__ int64_t forward(const Geom &g, const int c[4], int par, int mu)
{
    int d[4] = {c[0], c[1], c[2], c[3]};
    d[mu] = (d[mu] + 1 == g.N[mu]) ? 0 : d[mu] + 1;
    return site_index(g, d, par ^ 1);
}

// ---------------------------------------------------------------- Philox
// The words of the current block are named, not an array: an array indexed
// by the running word count is kept in scratch memory by the compiler.
struct Philox {
    uint32_t k0, k1, c0, c1;
    uint32_t w0, w1, w2, w3;
    int local;
};

__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; round++) {
        // each 32 x 32 -> 64-bit product in one multiply-add (both halves at once)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// key (tid, seed), counter (n, launch counter, 0x12345678, 0xabcdef09), first
// block uses n = 1 (PhiloxWrapper.hxx:93-140)
__device__ __forceinline__ void philox_init(Philox &p, uint32_t tid, uint32_t seed, uint32_t counter)
{
    p.k0 = tid;
    p.k1 = seed;
    p.c0 = 0;
    p.c1 = counter;
    p.local = 0;
}

__device__ __forceinline__ void philox_block(Philox &p)
{
    p.c0++;
    uint32_t c[4] = {p.c0, p.c1, 0x12345678u, 0xabcdef09u};
    philox10(c, p.k0, p.k1);
    p.w0 = c[0]; p.w1 = c[1]; p.w2 = c[2]; p.w3 = c[3];
}

// uniform in (0,1): FP64 hands out the 64-bit words [1] then [0] of a block,
// FP32 the 32-bit words 3, 2, 1, 0 (u01_open_open_64_53 / _32_24)
__device__ __forceinline__ double philox_rand(Philox &p, double)
{
    if (p.local == 0) {
        philox_block(p);
        p.local = 1;
        const uint64_t w = (uint64_t)p.w2 | ((uint64_t)p.w3 << 32);
        return (double)(1 + (w >> 11)) * 0x1.fffffffffffffp-54;
    }
    p.local = 0;
    const uint64_t w = (uint64_t)p.w0 | ((uint64_t)p.w1 << 32);
    return (double)(1 + (w >> 11)) * 0x1.fffffffffffffp-54;
}

__device__ __forceinline__ float philox_rand(Philox &p, float)
{
    if (p.local == 0) {
        philox_block(p);
        p.local = 3;
        return (float)(1 + (p.w3 >> 8)) * 0x1.fffffep-25f;
    }
    // words 2, 1, 0 in turn: shift the next one into w3
    p.local--;
    p.w3 = p.w2;
    p.w2 = p.w1;
    p.w1 = p.w0;
    return (float)(1 + (p.w3 >> 8)) * 0x1.fffffep-25f;
}

// ------------------------------------------------------ quad reductions
// The four lanes of a quad hold the four directions of one site.  A butterfly
// over quad_perm [1,0,3,2] then [2,3,0,1] leaves the same, bit-identical sum
// in all four lanes (floating-point addition is commutative).
template <int CTRL> __device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v)
{
    const int64_t b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

constexpr int QP_XOR1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int QP_XOR2 = 0x4E;  // quad_perm [2,3,0,1]

template <typename R> __device__ __forceinline__ R quad_sum(R v)
{
    v = v + dpp<QP_XOR1>(v);
    v = v + dpp<QP_XOR2>(v);
    return v;
}

}  // namespace su3gf
