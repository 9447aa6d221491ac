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
#include "su3gf_sweep.h"

namespace su3gf {
namespace {

constexpr int kGBlock = 256;  // 64 sites x 4 directions

// acc += A B^dagger
template <typename A> __device__ __forceinline__ void mac_adag(M3<A> &acc, const M3<A> &X, const M3<A> &Y)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Cx<A> s = acc.e[i][k];
#pragma unroll
            for (int j = 0; j < 3; j++) s = s + mul_conj(X.e[i][j], Y.e[k][j]);
            acc.e[i][k] = s;
        }
}

// acc += A B
template <typename A> __device__ __forceinline__ void mac_ab(M3<A> &acc, const M3<A> &X, const M3<A> &Y)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Cx<A> s = acc.e[i][k];
#pragma unroll
            for (int j = 0; j < 3; j++) s = s + X.e[i][j] * Y.e[j][k];
            acc.e[i][k] = s;
        }
}

template <typename A> __device__ __forceinline__ void zero(M3<A> &m)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) m.e[i][j] = cx((A)0, (A)0);
}

// rows 0,1 from the compact FP32 array, row 2 = conj(row0 x row1)
template <typename A>
__device__ __forceinline__ void load_su3(const float *base, int64_t V, int64_t idx, int mu, M3<A> &u)
{
    const float *q = base + (int64_t)12 * V * mu + idx;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int comp = 2 * (j + 3 * i);
            u.e[i][j] = cx((A)q[V * comp], (A)q[V * (comp + 1)]);
        }
    u.e[2][0] = conj(u.e[0][1] * u.e[1][2] - u.e[0][2] * u.e[1][1]);
    u.e[2][1] = conj(u.e[0][2] * u.e[1][0] - u.e[0][0] * u.e[1][2]);
    u.e[2][2] = conj(u.e[0][0] * u.e[1][1] - u.e[0][1] * u.e[1][0]);
}

// theta_x and Re tr of the up and down links from sum U_up = G P^dagger and
// sum U_down = Q G^dagger
template <int G, typename A>
__device__ __forceinline__ void site_theta(const M3<A> &Gx, const M3<A> &P, const M3<A> &Q, double &th, double &gf)
{
    M3<A> up, dn;
    zero(up);
    zero(dn);
    mac_adag(up, Gx, P);
    mac_adag(dn, Q, Gx);
    A v[kThetaVals];
    theta_contrib<A, G>(up, dn, true, v);
    double s[kThetaVals];
#pragma unroll
    for (int k = 0; k < kThetaVals; k++) s[k] = (double)v[k];
    theta_site<G>(s, th, gf);
}

// Block order: with `remap` the blocks are dealt out so that each of the eight
// XCDs (which take blocks round-robin) works through one contiguous eighth of
// the sites, keeping the neighbours' g in that XCD's L2.
__device__ __forceinline__ int64_t logical_block(int remap)
{
    const int64_t b = blockIdx.x, n = gridDim.x;
    if (!remap) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
}

// Work mapping: a block of 256 threads covers 64 sites as four waves, wave mu
// forming the direction-mu terms g(y) U0_mu(x)^dagger and g(z) U0_mu(z) of its
// 64 sites (four 12-real loads per lane, all coalesced).  The terms go through
// LDS; wave 0 adds them in the order mu = 0..3, builds M, runs the three
// subgroup updates on it and stores g(x).
template <typename A, int G, int ALGO, int THETA>
__global__ __launch_bounds__(kGBlock) void k_gsweep(GaugeSweepArgs a)
{
    constexpr int kMats = THETA ? 2 : 1;  // P and Q apart for theta, else their sum
    __shared__ A lds[kMats * 18 * 4 * 64];
    if (a.state && a.state->done) return;  // converged: uniform early exit
    const int mu = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t blk = logical_block(a.remap);
    const int64_t p = blk * 64 + lane;
    const bool active = p < a.geo.Vh;
    const int64_t site = (active ? p : 0) + (a.parity ? a.geo.Vh : 0);
    const int64_t V = a.geo.V;
    int c[4], par;
    site_coords(a.geo, site, c, par);

    M3<A> Gx;
    if (mu == 0) load_su3(a.g, V, site, 0, Gx);

    Philox rng;
    if (ALGO == GF_SA || ALGO == GF_RANDOM) {
        // key word 0 = global id of the reference's (mu = 0, up) thread of this
        // site in its V/2/NSB x 8*NSB launch (PhiloxWrapper.hxx:93-110)
        const uint32_t tid = (uint32_t)((p / a.nsb) * 8 * a.nsb + p % a.nsb);
        philox_init(rng, tid, a.seed, a.counter);
    }

    if (ALGO == GF_RANDOM) {
        // the random transformation does not depend on the links
        if (mu != 0) return;
    } else {
        M3<A> P, Q;
        zero(P);
        if (THETA) zero(Q);
        if (!(G == GF_COULOMB && mu == 0)) {
            M3<A> T, W;
            const int64_t y = forward(a.geo, c, par, mu);
            const int64_t z = backward(a.geo, c, par, mu);
            load_su3(a.g, V, y, 0, T);
            load_su3(a.U0, V, site, mu, W);
            mac_adag(P, T, W);
            load_su3(a.g, V, z, 0, T);
            load_su3(a.U0, V, z, mu, W);
            if (THETA) mac_ab(Q, T, W);
            else mac_ab(P, T, W);
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * (3 * i + j);
                lds[((k + 0) * 4 + mu) * 64 + lane] = P.e[i][j].x;
                lds[((k + 1) * 4 + mu) * 64 + lane] = P.e[i][j].y;
                if (THETA) {
                    lds[((18 + k + 0) * 4 + mu) * 64 + lane] = Q.e[i][j].x;
                    lds[((18 + k + 1) * 4 + mu) * 64 + lane] = Q.e[i][j].y;
                }
            }
        __syncthreads();
        if (mu != 0) return;
    }

    M3<A> P, Q;
    if (ALGO != GF_RANDOM) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int k = 2 * (3 * i + j);
                A v[2], w[2];
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const A *b = lds + (k + r) * 4 * 64 + lane;
                    v[r] = ((b[0] + b[64]) + b[128]) + b[192];
                    if (THETA) {
                        const A *bq = lds + (18 + k + r) * 4 * 64 + lane;
                        w[r] = ((bq[0] + bq[64]) + bq[128]) + bq[192];
                    }
                }
                P.e[i][j] = cx(v[0], v[1]);
                if (THETA) Q.e[i][j] = cx(w[0], w[1]);
            }
    }

    double th = 0, gf = 0;
    if (THETA == 1 && active) site_theta<G>(Gx, P, Q, th, gf);

    // M = (P + Q) Gx^dagger
    M3<A> M;
    zero(M);
    if (ALGO != GF_RANDOM) {
        mac_adag(M, P, Gx);
        if (THETA) mac_adag(M, Q, Gx);
    }
#define SU3GF_GSTEP(I, J)                                                 \
    {                                                                     \
        A q[4];                                                           \
        quat<A, I, J>(M, q);                                              \
        if (G == GF_U1XU1) {                                              \
            q[1] = 0;                                                     \
            q[2] = 0;                                                     \
        }                                                                 \
        calc_update<A, ALGO, false>(q, a.param, rng);                     \
        if (G == GF_U1XU1) {                                              \
            q[1] = 0;                                                     \
            q[2] = 0;                                                     \
            const A n = rsq(q[0] * q[0] + q[3] * q[3]);                   \
            q[0] *= n;                                                    \
            q[3] *= n;                                                    \
        }                                                                 \
        right_mult_dag<A, I, J>(M, q);                                    \
        left_mult<A, I, J>(Gx, q);                                        \
    }
    SU3GF_GSTEP(0, 2)
    SU3GF_GSTEP(1, 2)
    SU3GF_GSTEP(0, 1)
#undef SU3GF_GSTEP

    project(Gx);  // keep g unitary at the precision it is stored in
    if (active) {
        float *q = a.g + site;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int comp = 2 * (j + 3 * i);
                q[V * comp] = (float)Gx.e[i][j].x;
                q[V * (comp + 1)] = (float)Gx.e[i][j].y;
            }
    }

    if (THETA == 2 && active) site_theta<G>(Gx, P, Q, th, gf);
    if (THETA != 0) {
        th = wave64_sum(th);
        gf = wave64_sum(gf);
        if (lane == 0) {  // indexed by the block's sites, so a timeslice's partials are contiguous
            a.partial[2 * blk] = th;
            a.partial[2 * blk + 1] = gf;
        }
    }
}

template <typename A, int G> hipError_t launch_g(int algo, int theta, const GaugeSweepArgs &a, hipStream_t s)
{
    const dim3 grid((unsigned)gsweep_blocks(a.geo.Vh)), block(kGBlock);
    if (algo == GF_OR) {
        if (theta == 1) hipLaunchKernelGGL((k_gsweep<A, G, GF_OR, 1>), grid, block, 0, s, a);
        else if (theta == 2) hipLaunchKernelGGL((k_gsweep<A, G, GF_OR, 2>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_gsweep<A, G, GF_OR, 0>), grid, block, 0, s, a);
    } else if (theta != 0) {
        return hipErrorInvalidValue;
    } else if (algo == GF_MICRO) {
        hipLaunchKernelGGL((k_gsweep<A, G, GF_MICRO, 0>), grid, block, 0, s, a);
    } else if (algo == GF_SA) {
        hipLaunchKernelGGL((k_gsweep<A, G, GF_SA, 0>), grid, block, 0, s, a);
    } else if (algo == GF_RANDOM) {
        hipLaunchKernelGGL((k_gsweep<A, G, GF_RANDOM, 0>), grid, block, 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename A> hipError_t launch_a(int gauge, int algo, int theta, const GaugeSweepArgs &a, hipStream_t s)
{
    if (gauge == GF_LANDAU) return launch_g<A, GF_LANDAU>(algo, theta, a, s);
    if (gauge == GF_COULOMB) return launch_g<A, GF_COULOMB>(algo, theta, a, s);
    if (gauge == GF_U1XU1) return launch_g<A, GF_U1XU1>(algo, theta, a, s);
    return hipErrorInvalidValue;
}

}  // namespace

int64_t gsweep_blocks(int64_t Vh) { return (Vh + 63) / 64; }

hipError_t launch_gsweep(int gauge, int algo, int theta, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s)
{
    GaugeSweepArgs b = a;
    if (gsweep_blocks(a.geo.Vh) % 8 != 0) b.remap = 0;
    if (b.nsb <= 0) b.nsb = 32;
    return fp64_arith ? launch_a<double>(gauge, algo, theta, b, s) : launch_a<float>(gauge, algo, theta, b, s);
}

}  // namespace su3gf
