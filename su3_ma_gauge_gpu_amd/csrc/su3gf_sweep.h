// su3gf_sweep.h -- the checkerboard SU(2)-subgroup sweep kernel for gfx950.
//
// Included by two translation units: su3gf_exact.hip (compiled with
// -ffp-contract=off, EXACT = true: every multiply and add rounds separately,
// in the reference's order, and 1/sqrt replaces rsqrt) and su3gf_fast.hip
// (EXACT = false: FMA contraction, hardware rsqrt, optional fused theta and
// gauge-transformation accumulation).
//
// Work mapping: a 256-thread block is four waves; wave w handles direction
// mu = w and lane l handles site p = 64*blockIdx + l of the parity, so every
// load instruction reads 64 consecutive sites of one link component (512 B in
// FP64).  Lane (mu, p) owns the up link U_mu(x) and the down link U_mu(x-mu).
// Per SU(2) subgroup each lane writes the contributions of its two links to
// LDS; after one barrier every lane of the site reads them and adds them --
// the exact kernels all eight in the reference's order, up links mu = 0..3,
// then down links mu = 0..3 (GaugeFixingSubgroupStep.hxx:126-208), the fast
// kernels each direction's combined share -- so the four waves hold
// bit-identical sums and compute the same update, which each lane applies to
// its own two links.  Two LDS buffers alternate between subgroups, so one
// barrier per subgroup suffices.
//
// Reference (paths relative to its src/): GaugeFixingSubgroupStep.hxx:74-705,
// algorithms/{Or,Micro,Sa,Sr,Random}Update.hxx, {Landau,Coulomb,MAG,U1xU1}KernelsSU3.hxx
// apply(), LandauKernelsSU3.hxx:114-191 (theta), CommonKernelsSU3.hxx:71-90.
#pragma once
#include <type_traits>

#include "su3gf_kernels.h"

namespace su3gf {

template <bool EXACT, typename R> __device__ __forceinline__ R inv_sqrt(R v)
{
    if (EXACT) return (R)1 / sqrt_(v);
    return rsq(v);
}

// third row = conj(row0 x row1) / |.|, written as the reference does
// (SU3.hxx:655-664)
template <bool EXACT, typename R> __device__ __forceinline__ void reconstruct_e(M3<R> &u)
{
    const Cx<R> a = conj(u.e[0][1]) * conj(u.e[1][2]) - conj(u.e[0][2]) * conj(u.e[1][1]);
    const Cx<R> b = conj(u.e[0][2]) * conj(u.e[1][0]) - conj(u.e[0][0]) * conj(u.e[1][2]);
    const Cx<R> c = conj(u.e[0][0]) * conj(u.e[1][1]) - conj(u.e[0][1]) * conj(u.e[1][0]);
    const R n = inv_sqrt<EXACT>(abs2(a) + abs2(b) + abs2(c));
    u.e[2][0] = a * n;
    u.e[2][1] = b * n;
    u.e[2][2] = c * n;
}

// projectSU3 (SU3.hxx:536-583): normalise row 0, orthogonalise and normalise
// row 1 (divisions as in the reference), rebuild row 2
template <bool EXACT, typename R> __device__ __forceinline__ void project_e(M3<R> &u)
{
    R abs_u = abs2(u.e[0][0]);
    abs_u = abs_u + abs2(u.e[0][1]);
    abs_u = abs_u + abs2(u.e[0][2]);
    abs_u = sqrt_(abs_u);
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[0][k] = cx(u.e[0][k].x / abs_u, u.e[0][k].y / abs_u);
    Cx<R> sp = u.e[1][0] * conj(u.e[0][0]);
    sp = sp + u.e[1][1] * conj(u.e[0][1]);
    sp = sp + u.e[1][2] * conj(u.e[0][2]);
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[1][k] = u.e[1][k] - u.e[0][k] * sp;
    R abs_v = abs2(u.e[1][0]);
    abs_v = abs_v + abs2(u.e[1][1]);
    abs_v = abs_v + abs2(u.e[1][2]);
    abs_v = sqrt_(abs_v);
#pragma unroll
    for (int k = 0; k < 3; k++) u.e[1][k] = cx(u.e[1][k].x / abs_v, u.e[1][k].y / abs_v);
    reconstruct_e<EXACT>(u);
}

// ------------------------------------------------------------ update rules
__device__ __forceinline__ double cospi_(double x) { return cospi(x); }
__device__ __forceinline__ float cospi_(float x) { return cospif(x); }
__device__ __forceinline__ double sinpi_(double x) { return sinpi(x); }
__device__ __forceinline__ float sinpi_(float x) { return sinpif(x); }
__device__ __forceinline__ double log_(double x) { return log(x); }
__device__ __forceinline__ float log_(float x) { return logf(x); }

// OrUpdate.hxx:52-78
template <bool E, typename R> __device__ __forceinline__ void upd_or(R a[4], float omega)
{
    const R ai = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    const R a0 = a[0] * a[0];
    R b;
    if constexpr (!E && sizeof(R) == 4)  // FP32 without the reference's order: reciprocal to 1 ulp
        b = ((R)omega * a0 + ai) * __builtin_amdgcn_rcpf(a0 + ai);
    else
        b = ((R)omega * a0 + ai) / (a0 + ai);
    const R c = inv_sqrt<E>(a0 + b * b * ai);
    a[0] *= c;
    a[1] *= b * c;
    a[2] *= b * c;
    a[3] *= b * c;
}

// MicroUpdate.hxx:43-66
template <bool E, typename R> __device__ __forceinline__ void upd_micro(R a[4])
{
    const R ai = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    const R a0 = a[0] * a[0];
    if (E) {
        const R b = (R)2 * a[0] / (a0 + ai);
        a[0] = (a0 - ai) / (a0 + ai);
        a[1] *= b;
        a[2] *= b;
        a[3] *= b;
    } else {
        const R inv = (R)1 / (a0 + ai);
        const R b = (R)2 * a[0] * inv;
        a[0] = (a0 - ai) * inv;
        a[1] *= b;
        a[2] *= b;
        a[3] *= b;
    }
}

// SrUpdate.hxx:54-96
template <bool E, typename R> __device__ __forceinline__ void upd_sr(R a[4], float p, Philox &rng)
{
    const R r = philox_rand(rng, R(0));
    const R a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    if (r < (R)p) {
        a[0] = a0 * a0 - a1 * a1 - a2 * a2 - a3 * a3;
        a[1] = (R)2 * a0 * a1;
        a[2] = (R)2 * a0 * a2;
        a[3] = (R)2 * a0 * a3;
    }
    const R c = inv_sqrt<E>(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
    a[0] *= c;
    a[1] *= c;
    a[2] *= c;
    a[3] *= c;
}

// RandomUpdate.hxx:47-59
template <typename R> __device__ __forceinline__ void upd_random(R a[4], Philox &rng)
{
    const R alpha = philox_rand(rng, R(0));
    const R phi = (R)2 * philox_rand(rng, R(0));
    const R cos_theta = (R)2 * philox_rand(rng, R(0)) - (R)1;
    const R sin_theta = sqrt_((R)1 - cos_theta * cos_theta);
    const R sin_alpha = sinpi_(alpha);
    a[0] = cospi_(alpha);
    a[1] = sin_alpha * sin_theta * cospi_(phi);
    a[2] = sin_alpha * sin_theta * sinpi_(phi);
    a[3] = sin_alpha * cos_theta;
}

// SaUpdate.hxx:52-147 (Kennedy-Pendleton heatbath, temperature = 1/beta).
// The reference writes its constants as double literals (0.0001, 2.0, 1.0,
// 0.5), so with Real = float the exact kernels evaluate those expressions in
// double as it does -- the acceptance test, cospi(2.0*rand), 1.0-delta,
// 2.0*rand-1.0 and sqrt(1.0-cos^2) -- with float operands rounded first
// (r4*r4, cos_theta*cos_theta) and float results; W is that wider type.
// The fast kernels keep everything in R.
template <bool E, typename R, typename RAND = R>
__device__ __forceinline__ void upd_sa(R a[4], float temperature, Philox &rng)
{
    using W = std::conditional_t<E, double, R>;
    R e0 = a[0], e1 = -a[1], e2 = -a[2], e3 = -a[3];
    const R dk = inv_sqrt<E>(e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3);
    const R p0 = dk * (R)temperature;
    R delta;
    for (;;) {
        R r1, r2;
        do { r1 = (R)philox_rand(rng, RAND(0)); } while ((W)r1 < (W)0.0001);
        r1 = -log_(r1) * p0;
        do { r2 = (R)philox_rand(rng, RAND(0)); } while ((W)r2 < (W)0.0001);
        r2 = -log_(r2) * p0;
        R r3 = (R)cospi_((W)2 * (W)(R)philox_rand(rng, RAND(0)));
        r3 = r3 * r3;
        delta = r2 + r1 * r3;
        const R r4 = (R)philox_rand(rng, RAND(0));
        const R r4sq = r4 * r4;
        if (!((W)r4sq > (W)1 - (W)0.5 * (W)delta)) break;
    }
    const R b0 = (R)((W)1 - (W)delta);
    const R cos_theta = (R)((W)2 * (W)(R)philox_rand(rng, RAND(0)) - (W)1);
    const R cos_sq = cos_theta * cos_theta;
    const R sin_theta = (R)sqrt_((W)1 - (W)cos_sq);
    const R sin_alpha = sqrt_((R)1 - b0 * b0);
    const R phi = (R)((W)2 * (W)(R)philox_rand(rng, RAND(0)));
    const R b1 = sin_alpha * sin_theta * cospi_(phi);
    const R b2 = sin_alpha * sin_theta * sinpi_(phi);
    const R b3 = sin_alpha * cos_theta;
    e0 *= dk;
    e1 *= dk;
    e2 *= dk;
    e3 *= dk;
    a[0] = b0 * e0 + b3 * e3 + b2 * e2 + e1 * b1;
    a[3] = e0 * b3 - e3 * b0 + b1 * e2 - b2 * e1;
    a[2] = b3 * e1 - b0 * e2 + b2 * e0 - b1 * e3;
    a[1] = b2 * e3 + b1 * e0 - b3 * e2 - e1 * b0;
}

// The reference's double-precision updates on FP32 links (USE_DP_ORUPDATE,
// USE_DP_MICROUPDATE, USE_DP_SAUPDATE: OrUpdate.hxx:54-64, MicroUpdate.hxx:45-54,
// SaUpdate.hxx:54-98).  The subgroup sums stay float: the squares are float
// expressions assigned to doubles, as the reference's shA is Real; the rest
// is double, and each result is rounded to float as it is stored.
template <bool E> __device__ __forceinline__ void upd_or_dp(float a[4], float omega)
{
    const double ai = (double)(a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
    const double a0 = (double)(a[0] * a[0]);
    const double b = ((double)omega * a0 + ai) / (a0 + ai);
    const double c = inv_sqrt<E>(a0 + b * b * ai);
    a[0] = (float)((double)a[0] * c);
    a[1] = (float)((double)a[1] * (b * c));
    a[2] = (float)((double)a[2] * (b * c));
    a[3] = (float)((double)a[3] * (b * c));
}

__device__ __forceinline__ void upd_micro_dp(float a[4])
{
    const double ai = (double)(a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
    const double a0 = (double)(a[0] * a[0]);
    const double b = 2. * (double)a[0] / (a0 + ai);
    a[0] = (float)((a0 - ai) / (a0 + ai));
    a[1] = (float)((double)a[1] * b);
    a[2] = (float)((double)a[2] * b);
    a[3] = (float)((double)a[3] * b);
}

// the heatbath in double; the generator still hands out float uniforms
// (PhiloxWrapper<Real>::rand with Real = float)
template <bool E> __device__ __forceinline__ void upd_sa_dp(float a[4], float temperature, Philox &rng)
{
    double d[4] = {(double)a[0], (double)a[1], (double)a[2], (double)a[3]};
    upd_sa<E, double, float>(d, temperature, rng);
#pragma unroll
    for (int k = 0; k < 4; k++) a[k] = (float)d[k];
}

// A: a gf_algo, or GF_OR / GF_MICRO / GF_SA + kDpUpdate (FP32 links only)
template <typename R, int A, bool E> __device__ __forceinline__ void calc_update(R a[4], float param, Philox &rng)
{
    if constexpr (A >= kDpUpdate) {
        static_assert(sizeof(R) == 4, "double-precision updates are for FP32 links");
        constexpr int B = A - kDpUpdate;
        if constexpr (B == GF_OR) upd_or_dp<E>(a, param);
        else if constexpr (B == GF_MICRO) upd_micro_dp(a);
        else upd_sa_dp<E>(a, param, rng);
    } else {
        if (A == GF_OR) upd_or<E>(a, param);
        else if (A == GF_MICRO) upd_micro<E>(a);
        else if (A == GF_SA) upd_sa<E>(a, param, rng);
        else if (A == GF_SR) upd_sr<E>(a, param, rng);
        else upd_random(a, rng);
    }
}

// ---------------------------------------------------- per-link contributions
// Landau / Coulomb / U1xU1: the raw subgroup quaternion q(U) (SU3.hxx:193-217).
// MAG: D = |a00|^2 - |a01|^2 - |a10|^2 + |a11|^2 and c = a00 conj(a10) - a01 conj(a11)
// (up) or conj(a00) a01 - conj(a10) a11 (down), stored as (D, Im c, Re c)
// (GaugeFixingSubgroupStep.hxx:387-390,457-460).
template <typename R, int G, int I, int J, bool UP>
__device__ __forceinline__ void link_contrib(const M3<R> &u, R v[4])
{
    if (G == GF_MAG) {
        const Cx<R> a00 = u.e[I][I], a01 = u.e[I][J], a10 = u.e[J][I], a11 = u.e[J][J];
        v[0] = abs2(a00) - abs2(a01) - abs2(a10) + abs2(a11);
        const Cx<R> c = UP ? a00 * conj(a10) - a01 * conj(a11) : conj(a00) * a01 - conj(a10) * a11;
        v[1] = c.y;
        v[2] = c.x;
        v[3] = 0;
    } else {
        quat<R, I, J>(u, v);
    }
}

// The update of a site from its summed subgroup quaternion s (the update
// rule, then the gauge-specific restriction and normalisation).
template <typename R, int G, int A, bool E>
__device__ __forceinline__ void finish_update(R s0, R s1, R s2, R s3, R q[4], float param, Philox &rng)
{
    if (G == GF_LANDAU || G == GF_COULOMB) {
        q[0] = s0; q[1] = s1; q[2] = s2; q[3] = s3;
        calc_update<R, A, E>(q, param, rng);
    } else if (G == GF_U1XU1) {
        q[0] = s0; q[1] = 0; q[2] = 0; q[3] = s3;
        calc_update<R, A, E>(q, param, rng);
        q[1] = 0;
        q[2] = 0;
        const R n = inv_sqrt<E>(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        q[0] *= n; q[1] *= n; q[2] *= n; q[3] *= n;
    } else {  // MAG (GaugeFixingSubgroupStep.hxx:529-575)
        s1 *= (R)2;
        s2 *= (R)2;
        s0 = s0 + sqrt_(s0 * s0 + s1 * s1 + s2 * s2);
        q[0] = s0; q[1] = s1; q[2] = s2; q[3] = 0;
        calc_update<R, A, E>(q, param, rng);
        q[3] = 0;
        const R n = inv_sqrt<E>(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        q[0] *= n; q[1] *= n; q[2] *= n; q[3] *= n;
    }
}

// Reference accumulation of the eight contributions of one site in ladder
// order (up mu = 0..3, then down); q[] receives the SU(2) element to apply.
template <typename R, int G, int A, bool E>
__device__ __forceinline__ void site_update(const R (&c)[8][4], R q[4], float param, Philox &rng)
{
    R s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (G == GF_LANDAU || G == GF_COULOMB) {
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (G == GF_COULOMB && m == 0) continue;
            s0 += c[m][0]; s1 -= c[m][1]; s2 -= c[m][2]; s3 -= c[m][3];
        }
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (G == GF_COULOMB && m == 0) continue;
            s0 += c[4 + m][0]; s1 += c[4 + m][1]; s2 += c[4 + m][2]; s3 += c[4 + m][3];
        }
    } else if (G == GF_U1XU1) {
#pragma unroll
        for (int m = 0; m < 4; m++) { s0 += c[m][0]; s3 -= c[m][3]; }
#pragma unroll
        for (int m = 0; m < 4; m++) { s0 += c[4 + m][0]; s3 += c[4 + m][3]; }
    } else {
#pragma unroll
        for (int m = 0; m < 8; m++) { s0 += c[m][0]; s1 += c[m][1]; s2 += c[m][2]; }
    }
    finish_update<R, G, A, E>(s0, s1, s2, s3, q, param, rng);
}

// One direction's share of the summed quaternion: its up link's contribution
// u and its down link's d combined with the signs of the ladder sum above
// (Coulomb's time direction contributes nothing).
template <typename R, int G>
__device__ __forceinline__ void direction_share(const R u[4], const R d[4], bool time_link, R p[4])
{
    if ((G == GF_COULOMB && time_link)) {
        p[0] = p[1] = p[2] = p[3] = 0;
    } else if (G == GF_MAG) {
        p[0] = u[0] + d[0]; p[1] = u[1] + d[1]; p[2] = u[2] + d[2]; p[3] = 0;
    } else {  // Landau, Coulomb, U1xU1 (which reads p[0] and p[3] only)
        p[0] = u[0] + d[0]; p[1] = d[1] - u[1]; p[2] = d[2] - u[2]; p[3] = d[3] - u[3];
    }
}

// The same update from the four directions' shares, added in the order
// mu = 0..3 (the reference's order of the eight terms is not kept).
template <typename R, int G, int A, bool E>
__device__ __forceinline__ void site_update_dirs(const R (&c)[4][4], R q[4], float param, Philox &rng)
{
    const R s0 = ((c[0][0] + c[1][0]) + c[2][0]) + c[3][0];
    const R s1 = ((c[0][1] + c[1][1]) + c[2][1]) + c[3][1];
    const R s2 = ((c[0][2] + c[1][2]) + c[2][2]) + c[3][2];
    const R s3 = ((c[0][3] + c[1][3]) + c[2][3]) + c[3][3];
    finish_update<R, G, A, E>(s0, s1, s2, s3, q, param, rng);
}

// ------------------------------------------------------- fused theta / gff
// Per lane: the traceless anti-hermitian part of (U_up - U_dn) as 9 reals
// (Im M00, Im M11, Im M22, M01, M02, M12) with M = (D - tr D/3) - h.c., and
// Re tr U_up + Re tr U_dn -- a kernel of one parity touches every link once
// (as the up link of a site of that parity or as the down link, i.e. the up
// link of its neighbour), so these sum to the whole lattice's functional.
// MAG: the three complex X_k parts and the sum of |U_jj|^2 of both links.
constexpr int kThetaVals = 10;

template <typename R, int G>
__device__ __forceinline__ void theta_contrib(const M3<R> &up, const M3<R> &dn, bool use, R v[kThetaVals])
{
#pragma unroll
    for (int k = 0; k < kThetaVals; k++) v[k] = 0;
    if (!use) return;
    if (G == GF_MAG) {
        int k = 0;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = i + 1; j < 3; j++) {
                const Cx<R> cu = up.e[i][i] * conj(up.e[j][i]) - up.e[i][j] * conj(up.e[j][j]);
                const Cx<R> cd = conj(dn.e[i][i]) * dn.e[i][j] - conj(dn.e[j][i]) * dn.e[j][j];
                v[k++] = cu.x + cd.x;
                v[k++] = cu.y + cd.y;
            }
        v[9] = (abs2(up.e[0][0]) + abs2(up.e[1][1]) + abs2(up.e[2][2])) +
               (abs2(dn.e[0][0]) + abs2(dn.e[1][1]) + abs2(dn.e[2][2]));
    } else {
        M3<R> d;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) d.e[i][j] = up.e[i][j] - dn.e[i][j];
        const R imtr3 = (d.e[0][0].y + d.e[1][1].y + d.e[2][2].y) / (R)3;
        v[0] = (R)2 * (d.e[0][0].y - imtr3);
        v[1] = (R)2 * (d.e[1][1].y - imtr3);
        v[2] = (R)2 * (d.e[2][2].y - imtr3);
        v[3] = d.e[0][1].x - d.e[1][0].x;
        v[4] = d.e[0][1].y + d.e[1][0].y;
        v[5] = d.e[0][2].x - d.e[2][0].x;
        v[6] = d.e[0][2].y + d.e[2][0].y;
        v[7] = d.e[1][2].x - d.e[2][1].x;
        v[8] = d.e[1][2].y + d.e[2][1].y;
        v[9] = (up.e[0][0].x + up.e[1][1].x + up.e[2][2].x) + (dn.e[0][0].x + dn.e[1][1].x + dn.e[2][2].x);
    }
}

template <int G> __device__ __forceinline__ void theta_site(const double s[kThetaVals], double &th, double &gf)
{
    if (G == GF_MAG) {
        th = sqrt(s[0] * s[0] + s[1] * s[1]) + sqrt(s[2] * s[2] + s[3] * s[3]) + sqrt(s[4] * s[4] + s[5] * s[5]);
    } else {
        th = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] +
             2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5] + s[6] * s[6] + s[7] * s[7] + s[8] * s[8]);
    }
    gf = s[9];
}

__device__ __forceinline__ double wave64_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------- kernel
// THETA: 0 none, 1 theta of the even sites before the update (even kernel),
// 2 theta of the odd sites after the update (odd kernel), 3 theta of the
// parity's sites and no update (a read-only pass, run only while
// a.state->near is set); each block writes its (theta, gff) sums to
// a.partial[2*block .. 2*block+1].
// GACC: also left-multiply the site's gauge transformation g(x) (FP32 rows
// 0,1 in 64-site tiles at a.gacc) by the same SU(2) elements.
// EXACT exchanges the eight links' contributions and adds them in the
// reference's ladder order; otherwise each wave exchanges its direction's
// combined share (half the LDS traffic and registers), and the fused check
// has an LDS area of its own, so that the even kernel's theta shares the first
// exchange's barrier.
template <typename R, int G, int A, int S, bool EXACT, int THETA, bool GACC>
__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs<R> a)
{
    static_assert(!EXACT || THETA == 0, "the reference-order kernels carry no fused check");
    constexpr int kTerms = EXACT ? 8 : 4;   // exchanged per site: links, or directions
    constexpr int kQ = kTerms * 4 * 64;     // one exchange buffer: [term][comp 4][lane 64]
    constexpr int kT = kThetaVals * 4 * 64;
    __shared__ R lds[2 * kQ + (THETA ? kT : 0)];
    R *const tlds = lds + 2 * kQ;

    if (a.state && a.state->done) return;  // converged: uniform early exit
    if (THETA == 3 && !a.state->near) return;

    const int mu = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    bool active;
    const int64_t p = block_site(a.g, blockIdx.x, lane, active);
    const int64_t site = (active ? p : 0) + (a.parity ? a.g.Vh : 0);
    int c[4], par;
    site_coords(a.g, site, c, par);
    // Coulomb: the time links are transformed but do not enter the sums.  On a
    // timeslice ({1,Nx,Ny,Nz}) the down time link is U_0(x, t-1) in the slice
    // t-1 buffer at the same index (CoulombKernelsSU3.hxx:208-250); on a 4D
    // lattice (all slices at once) it is the ordinary backward link.
    const bool time_link = (G == GF_COULOMB) && (mu == 0);
    const bool slice_dw = time_link && a.g.N[0] == 1;
    const int64_t nb = slice_dw ? site : backward(a.g, c, par, mu);
    const LinkRef<R> lu = link_ref(a.U, a.g.V, S, site, mu);
    const LinkRef<R> ld = link_ref(slice_dw ? a.Udw : a.U, a.g.V, S, nb, mu);
    M3<R> up, dn;
    if (active) {
        load12(lu, up);
        load12(ld, dn);
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) { up.e[i][j] = cx((R)(i == j), (R)0); dn.e[i][j] = up.e[i][j]; }
    }
    reconstruct_e<EXACT>(up);
    reconstruct_e<EXACT>(dn);

    M3<float> gt;
    LinkRef<float> lg;
    const bool own_g = GACC && mu == 0 && active;
    if (GACC) {
        lg = gauge_ref(a.gacc, site);
        if (own_g) {
            load12(lg, gt);
            reconstruct_e<false>(gt);
        }
    }

    Philox rng;
    constexpr int AB = A % kDpUpdate;  // the update rule without its precision
    if (AB == GF_SA || AB == GF_SR || AB == GF_RANDOM) {
        // key word 0 = global thread id of the reference's (mu=0, down=0) thread
        const uint32_t tid = (uint32_t)((p / a.nsb) * 8 * a.nsb + p % a.nsb);
        philox_init(rng, tid, a.seed, a.counter);
    }

    // fused check: each wave writes its direction's terms to the check's own
    // LDS area, and wave 0 adds them at the end of the kernel, after a barrier
    // that follows the writes (the area is written once per launch, so no
    // barrier guards its reuse; the even kernel's writes precede the first
    // exchange's barrier)
    double th = 0, gf = 0;
    auto theta_write = [&]() {
        R v[kThetaVals];
        theta_contrib<R, G>(up, dn, active && !time_link, v);
#pragma unroll
        for (int k = 0; k < kThetaVals; k++) tlds[(k * 4 + mu) * 64 + lane] = v[k];
    };
    auto theta_read = [&]() {
        if (mu != 0) return;
        double s[kThetaVals];
#pragma unroll
        for (int k = 0; k < kThetaVals; k++)
            s[k] = ((double)tlds[(k * 4 + 0) * 64 + lane] + (double)tlds[(k * 4 + 1) * 64 + lane]) +
                   ((double)tlds[(k * 4 + 2) * 64 + lane] + (double)tlds[(k * 4 + 3) * 64 + lane]);
        double t, g;
        theta_site<G>(s, t, g);
        th = wave64_sum(active ? t : 0.0);
        gf = wave64_sum(active ? g : 0.0);
    };

    if (THETA == 1) theta_write();
    if constexpr (THETA == 3) {
        theta_write();
        __syncthreads();
        theta_read();
        if (threadIdx.x == 0) {
            a.partial[2 * blockIdx.x] = th;
            a.partial[2 * blockIdx.x + 1] = gf;
        }
        return;
    }

    auto subgroup = [&](auto I_, auto J_, int buf) {
        constexpr int I = decltype(I_)::value, J = decltype(J_)::value;
        R *const bq = lds + buf * kQ;
        R vu[4], vd[4];
        link_contrib<R, G, I, J, true>(up, vu);
        link_contrib<R, G, I, J, false>(dn, vd);
        R q[4];
        if constexpr (EXACT) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                bq[((0 + mu) * 4 + k) * 64 + lane] = vu[k];
                bq[((4 + mu) * 4 + k) * 64 + lane] = vd[k];
            }
            __syncthreads();
            R cc[8][4];
#pragma unroll
            for (int m = 0; m < 8; m++)
#pragma unroll
                for (int k = 0; k < 4; k++) cc[m][k] = bq[(m * 4 + k) * 64 + lane];
            site_update<R, G, A, EXACT>(cc, q, a.param, rng);
        } else {
            R pd[4];
            direction_share<R, G>(vu, vd, time_link, pd);
#pragma unroll
            for (int k = 0; k < 4; k++) bq[(mu * 4 + k) * 64 + lane] = pd[k];
            __syncthreads();
            R cc[4][4];
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int k = 0; k < 4; k++) cc[m][k] = bq[(m * 4 + k) * 64 + lane];
            site_update_dirs<R, G, A, EXACT>(cc, q, a.param, rng);
        }
        left_mult<R, I, J>(up, q);
        right_mult_dag<R, I, J>(dn, q);
        if (GACC && own_g) {
            const float qf[4] = {(float)q[0], (float)q[1], (float)q[2], (float)q[3]};
            left_mult<float, I, J>(gt, qf);
        }
    };
    subgroup(std::integral_constant<int, 0>(), std::integral_constant<int, 2>(), 0);
    subgroup(std::integral_constant<int, 1>(), std::integral_constant<int, 2>(), 1);
    subgroup(std::integral_constant<int, 0>(), std::integral_constant<int, 1>(), 0);

    if (active) {
        store12(lu, up);
        store12(ld, dn);
    }
    if (GACC && own_g) store12(lg, gt);

    if (THETA == 1) theta_read();
    if (THETA == 2) {
        theta_write();
        __syncthreads();
        theta_read();
    }
    if (THETA != 0 && threadIdx.x == 0) {
        // thread 0 (wave 0) holds the block's sums; a separate check kernel
        // reduces the partials of both parities in a fixed order
        a.partial[2 * blockIdx.x] = th;
        a.partial[2 * blockIdx.x + 1] = gf;
    }
}

}  // namespace su3gf
