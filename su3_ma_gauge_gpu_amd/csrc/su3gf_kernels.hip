// su3gf_kernels.hip -- gfx950 kernels of the SU(3) gauge-fixing sweep.
//
// Work mapping (sweep and quality): one quad of lanes per lattice site.  Lane
// mu of the quad owns the up link U_mu(x) and the down link U_mu(x - mu), so a
// wave64 updates 16 sites and every link of a parity is touched by exactly one
// lane.  Each lane loads rows 0,1 of its two links (12 reals each, coalesced:
// 16 consecutive sites per direction per load instruction), rebuilds row 2,
// and the subgroup sums over the 8 links of a site are formed with two DPP
// quad_perm steps -- no LDS, no barriers.  The SU(2) update is computed
// redundantly (bit-identically) in the four lanes and each lane applies it to
// its own two links.  Only rows 0,1 are stored back, as the reference does.
//
// Reference (paths relative to its src/): the per-site arithmetic follows
// gaugefixing/GaugeFixingSubgroupStep.hxx:74-705, the update rules
// gaugefixing/algorithms/{Or,Micro,Sa,Sr,Random}Update.hxx, the apply()
// bodies gaugefixing/{Landau,Coulomb,MAG,U1xU1}KernelsSU3.hxx, the quality
// kernels gaugefixing/{Landau,Coulomb,MAG}KernelsSU3.hxx generateGaugeQualityPerSite
// and gaugefixing/CommonKernelsSU3.hxx:71-130.
#include "su3gf_kernels.h"

namespace su3gf {

// ------------------------------------------------------------ update rules
// OrUpdate.hxx:52-78
template <typename R> __device__ __forceinline__ void upd_or(R a[4], float omega)
{
    const R ai = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    const R a0 = a[0] * a[0];
    const R b = ((R)omega * a0 + ai) / (a0 + ai);
    const R c = rsq(a0 + b * b * ai);
    const R bc = b * c;
    a[0] *= c;
    a[1] *= bc;
    a[2] *= bc;
    a[3] *= bc;
}

// MicroUpdate.hxx:43-66
template <typename R> __device__ __forceinline__ void upd_micro(R a[4])
{
    const R ai = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
    const R a0 = a[0] * a[0];
    const R inv = (R)1 / (a0 + ai);
    const R b = (R)2 * a[0] * inv;
    a[0] = (a0 - ai) * inv;
    a[1] *= b;
    a[2] *= b;
    a[3] *= b;
}

__device__ __forceinline__ double cospi_(double x) { return cospi(x); }
__device__ __forceinline__ float cospi_(float x) { return cospif(x); }
__device__ __forceinline__ double sinpi_(double x) { return sinpi(x); }
__device__ __forceinline__ float sinpi_(float x) { return sinpif(x); }
__device__ __forceinline__ double log_(double x) { return log(x); }
__device__ __forceinline__ float log_(float x) { return logf(x); }

// SrUpdate.hxx:54-96
template <typename R> __device__ __forceinline__ void upd_sr(R a[4], float p, Philox &rng)
{
    const R r = philox_rand(rng, R(0));
    const R a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    if (r < (R)p) {
        a[0] = a0 * a0 - a1 * a1 - a2 * a2 - a3 * a3;
        a[1] = (R)2 * a0 * a1;
        a[2] = (R)2 * a0 * a2;
        a[3] = (R)2 * a0 * a3;
    }
    const R c = rsq(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
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

// SaUpdate.hxx:52-147 (Kennedy-Pendleton heatbath, temperature = 1/beta)
template <typename R> __device__ __forceinline__ void upd_sa(R a[4], float temperature, Philox &rng)
{
    R e0 = a[0], e1 = -a[1], e2 = -a[2], e3 = -a[3];
    const R dk = rsq(e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3);
    const R p0 = dk * (R)temperature;
    R delta;
    for (;;) {
        R r1, r2;
        do { r1 = philox_rand(rng, R(0)); } while (r1 < (R)0.0001);
        r1 = -log_(r1) * p0;
        do { r2 = philox_rand(rng, R(0)); } while (r2 < (R)0.0001);
        r2 = -log_(r2) * p0;
        R r3 = cospi_((R)2 * philox_rand(rng, R(0)));
        r3 = r3 * r3;
        delta = r2 + r1 * r3;
        const R r4 = philox_rand(rng, R(0));
        if (!(r4 * r4 > (R)1 - (R)0.5 * delta)) break;
    }
    const R b0 = (R)1 - delta;
    const R cos_theta = (R)2 * philox_rand(rng, R(0)) - (R)1;
    const R sin_theta = sqrt_((R)1 - cos_theta * cos_theta);
    const R sin_alpha = sqrt_((R)1 - b0 * b0);
    const R phi = (R)2 * philox_rand(rng, R(0));
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

template <typename R, int A> __device__ __forceinline__ void calc_update(R a[4], float param, Philox &rng)
{
    if (A == GF_OR) upd_or(a, param);
    else if (A == GF_MICRO) upd_micro(a);
    else if (A == GF_SA) upd_sa(a, param, rng);
    else if (A == GF_SR) upd_sr(a, param, rng);
    else upd_random(a, rng);
}

template <typename R> __device__ __forceinline__ void normalize4(R q[4])
{
    const R n = rsq(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[0] *= n;
    q[1] *= n;
    q[2] *= n;
    q[3] *= n;
}

// One SU(2) subgroup step of one site, seen from the lane that owns the up
// link `up` = U_mu(x) and the down link `dn` = U_mu(x-mu)
// (GaugeFixingSubgroupStep.hxx:74-705).  `w` = 0 masks this lane's
// contribution to the site sums (Coulomb time direction).
template <typename R, int G, int A, int I, int J>
__device__ __forceinline__ void subgroup(M3<R> &up, M3<R> &dn, R w, float param, Philox &rng)
{
    R q[4];
    if (G == GF_LANDAU || G == GF_COULOMB) {
        R qu[4], qd[4];
        quat<R, I, J>(up, qu);
        quat<R, I, J>(dn, qd);
        // up links enter as q(U^dagger) = (q0,-q1,-q2,-q3), down links as q(U)
        q[0] = quad_sum(w * (qu[0] + qd[0]));
        q[1] = quad_sum(w * (qd[1] - qu[1]));
        q[2] = quad_sum(w * (qd[2] - qu[2]));
        q[3] = quad_sum(w * (qd[3] - qu[3]));
        calc_update<R, A>(q, param, rng);
    } else if (G == GF_U1XU1) {
        R qu[4], qd[4];
        quat<R, I, J>(up, qu);
        quat<R, I, J>(dn, qd);
        q[0] = quad_sum(qu[0] + qd[0]);
        q[1] = 0;
        q[2] = 0;
        q[3] = quad_sum(qd[3] - qu[3]);
        calc_update<R, A>(q, param, rng);
        q[1] = 0;
        q[2] = 0;
        normalize4(q);
    } else {  // MAG
        const Cx<R> u00 = up.e[I][I], u01 = up.e[I][J], u10 = up.e[J][I], u11 = up.e[J][J];
        const Cx<R> d00 = dn.e[I][I], d01 = dn.e[I][J], d10 = dn.e[J][I], d11 = dn.e[J][J];
        const R D = (abs2(u00) - abs2(u01) - abs2(u10) + abs2(u11)) + (abs2(d00) - abs2(d01) - abs2(d10) + abs2(d11));
        // up: c = a00 conj(a10) - a01 conj(a11); down: c = conj(a00) a01 - conj(a10) a11
        const Cx<R> cu = mul_conj(u00, u10) - mul_conj(u01, u11);
        const Cx<R> cd = mul_conj(d01, d00) - mul_conj(d11, d10);
        q[0] = quad_sum(D);
        q[1] = quad_sum(cu.y + cd.y) * (R)2;
        q[2] = quad_sum(cu.x + cd.x) * (R)2;
        q[3] = 0;
        q[0] = q[0] + sqrt_(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        calc_update<R, A>(q, param, rng);
        q[3] = 0;
        normalize4(q);
    }
    left_mult<R, I, J>(up, q);
    right_mult_dag<R, I, J>(dn, q);
}

// ------------------------------------------------------------------ sweep
template <typename R, int G, int A, int S>
__global__ __launch_bounds__(kBlock) void k_sweep(SweepArgs<R> a)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t p = t >> 2;
    const int mu = (int)(t & 3);
    if (p >= a.g.Vh) return;
    const int64_t site = p + (a.parity ? a.g.Vh : 0);
    int c[4], par;
    site_coords(a.g, site, c, par);
    const bool time_link = (G == GF_COULOMB) && (mu == 0);
    const int64_t nb = time_link ? site : backward(a.g, c, par, mu);
    const LinkRef<R> lu = link_ref(a.U, a.g.V, S, site, mu);
    const LinkRef<R> ld = link_ref(time_link ? a.Udw : a.U, a.g.V, S, nb, mu);
    M3<R> up, dn;
    load12(lu, up);
    load12(ld, dn);
    reconstruct(up);
    reconstruct(dn);
    Philox rng;
    if (A == GF_SA || A == GF_SR || A == GF_RANDOM) {
        // key word 0 = global thread id of the reference's (mu=0, down=0) thread
        const uint32_t tid = (uint32_t)((p / a.nsb) * 8 * a.nsb + p % a.nsb);
        philox_init(rng, tid, a.seed, a.counter);
    }
    const R w = time_link ? (R)0 : (R)1;
    subgroup<R, G, A, 0, 2>(up, dn, w, a.param, rng);
    subgroup<R, G, A, 1, 2>(up, dn, w, a.param, rng);
    subgroup<R, G, A, 0, 1>(up, dn, w, a.param, rng);
    store12(lu, up);
    store12(ld, dn);
}

// -------------------------------------------------------------- project
// CommonKernelsSU3.hxx:71-90 -- one thread per link.
template <typename R, int S>
__global__ __launch_bounds__(kBlock) void k_project(R *U, int64_t V)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * V) return;
    const int mu = (int)(t / V);
    const int64_t idx = t - (int64_t)mu * V;
    const LinkRef<R> l = link_ref(U, V, S, idx, mu);
    M3<R> u;
    load12(l, u);
    project(u);
    store12(l, u);
    if (S == 18) store_row2(l, u);
}

// ---------------------------------------------------------- convert
template <typename RI, int SI, typename RO, int SO>
__global__ __launch_bounds__(kBlock) void k_convert(const RI *src, RO *dst, int64_t V, int reunitarize)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * V) return;
    const int mu = (int)(t / V);
    const int64_t idx = t - (int64_t)mu * V;
    const LinkRef<RI> ls = link_ref(const_cast<RI *>(src), V, SI, idx, mu);
    const LinkRef<RO> ld = link_ref(dst, V, SO, idx, mu);
    M3<RI> a;
    load12(ls, a);
    M3<RO> b;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) b.e[i][j] = cx((RO)a.e[i][j].x, (RO)a.e[i][j].y);
    if (reunitarize) project(b);
    else reconstruct(b);
    store12(ld, b);
    if (SO == 18) store_row2(ld, b);
}

// ------------------------------------------------------------- set hot
// CommonKernelsSU3.hxx:92-130 -- one thread per site, 18-real layout.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_set_hot(R *U, int64_t V, uint32_t seed, uint32_t counter)
{
    const int64_t site = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (site >= V) return;
    Philox rng;
    philox_init(rng, (uint32_t)site, seed, counter);
    for (int mu = 0; mu < 4; mu++) {
        M3<R> u;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) u.e[i][j] = cx((R)(i == j ? 1 : 0), (R)0);
        R q[4];
#define SU3GF_HOT_SUB(I, J)                                                                    \
        q[0] = philox_rand(rng, R(0)) * (R)2 - (R)1;                                           \
        q[1] = philox_rand(rng, R(0)) * (R)2 - (R)1;                                           \
        q[2] = philox_rand(rng, R(0)) * (R)2 - (R)1;                                           \
        q[3] = philox_rand(rng, R(0)) * (R)2 - (R)1;                                           \
        {                                                                                      \
            const R n = rsq(q[0] * q[0] + q[3] * q[3] + q[2] * q[2] + q[1] * q[1]);            \
            R h[4] = {q[0] * n, -q[1] * n, -q[2] * n, -q[3] * n};                              \
            right_mult_dag<R, I, J>(u, h);                                                     \
        }
        SU3GF_HOT_SUB(0, 1)
        SU3GF_HOT_SUB(0, 2)
        SU3GF_HOT_SUB(1, 2)
#undef SU3GF_HOT_SUB
        const LinkRef<R> l = link_ref(U, V, 18, site, mu);
        store12(l, u);
        store_row2(l, u);
    }
}

// ------------------------------------------------------------- quality
// Per site (quad): Landau/U1xU1 (LandauKernelsSU3.hxx:114-191), Coulomb
// (mu >= 1 only, CoulombKernelsSU3.hxx:110-183), MAG (MAGKernelsSU3.hxx:116-223).
// Writes per-site values (if dGff != null) and one (theta, gff) partial per block.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

template <typename R, int G, int S>
__global__ __launch_bounds__(kBlock) void k_quality(QualityArgs<R> a)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t site = t >> 2;
    const int mu = (int)(t & 3);
    double th = 0, gf = 0;
    if (site < a.g.V) {
        int c[4], par;
        site_coords(a.g, site, c, par);
        const int64_t nb = backward(a.g, c, par, mu);
        M3<R> up, dn;
        load12(link_ref(const_cast<R *>(a.U), a.g.V, S, site, mu), up);
        load12(link_ref(const_cast<R *>(a.U), a.g.V, S, nb, mu), dn);
        reconstruct(up);
        reconstruct(dn);
        if (G == GF_MAG) {
            double gl = (double)abs2(up.e[0][0]) + (double)abs2(up.e[1][1]) + (double)abs2(up.e[2][2]);
            double X[6];
            int k = 0;
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = i + 1; j < 3; j++) {
                    const Cx<R> cu = mul_conj(up.e[i][i], up.e[j][i]) - mul_conj(up.e[i][j], up.e[j][j]);
                    const Cx<R> cd = mul_conj(dn.e[i][j], dn.e[i][i]) - mul_conj(dn.e[j][j], dn.e[j][i]);
                    X[k++] = (double)cu.x + (double)cd.x;
                    X[k++] = (double)cu.y + (double)cd.y;
                }
            gl = quad_sum(gl);
#pragma unroll
            for (int m = 0; m < 6; m++) X[m] = quad_sum(X[m]);
            // i+j-1 ordering of the reference: (0,1)->0, (0,2)->1, (1,2)->2
            th = sqrt(X[0] * X[0] + X[1] * X[1]) + sqrt(X[2] * X[2] + X[3] * X[3]) + sqrt(X[4] * X[4] + X[5] * X[5]);
            gf = gl;
        } else {
            const R w = (G == GF_COULOMB && mu == 0) ? (R)0 : (R)1;
            M3<R> d;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    d.e[i][j] = cx(quad_sum(w * (up.e[i][j].x - dn.e[i][j].x)), quad_sum(w * (up.e[i][j].y - dn.e[i][j].y)));
            const R trx = (d.e[0][0].x + d.e[1][1].x + d.e[2][2].x) / (R)3;
            const R try_ = (d.e[0][0].y + d.e[1][1].y + d.e[2][2].y) / (R)3;
#pragma unroll
            for (int i = 0; i < 3; i++) d.e[i][i] = cx(d.e[i][i].x - trx, d.e[i][i].y - try_);
            double prec = 0;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    // (D - D^dagger)_ij = D_ij - conj(D_ji)
                    const R x = d.e[i][j].x - d.e[j][i].x;
                    const R y = d.e[i][j].y + d.e[j][i].y;
                    prec += (double)(x * x + y * y);
                }
            th = prec;
            gf = quad_sum((double)(w * (up.e[0][0].x + up.e[1][1].x + up.e[2][2].x)));
        }
        if (mu == 0 && a.dGff) {
            a.dGff[site] = gf;
            a.dA[site] = th;
        }
    }
    if (mu != 0) { th = 0; gf = 0; }
    if (a.partial) {
        __shared__ double sh[2][kBlock / 64];
        double ts = a.crit_max ? wave_max(th) : wave_sum(th);
        double gs = wave_sum(gf);
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { sh[0][wv] = ts; sh[1][wv] = gs; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double T = sh[0][0], Gs = sh[1][0];
            for (int k = 1; k < kBlock / 64; k++) {
                T = a.crit_max ? fmax(T, sh[0][k]) : T + sh[0][k];
                Gs += sh[1][k];
            }
            a.partial[2 * blockIdx.x] = T;
            a.partial[2 * blockIdx.x + 1] = Gs;
        }
    }
}

// Fixed-order final reduction of the block partials: thread k sums partials
// k, k + 1024, ... in order, then a fixed tree -> out[0] = theta, out[1] = gff.
__global__ __launch_bounds__(1024) void k_reduce_final(const double *partial, int64_t n, int crit_max, double *out)
{
    __shared__ double st[1024], sg[1024];
    double T = 0, Gs = 0;
    for (int64_t k = threadIdx.x; k < n; k += 1024) {
        T = crit_max ? fmax(T, partial[2 * k]) : T + partial[2 * k];
        Gs += partial[2 * k + 1];
    }
    st[threadIdx.x] = T;
    sg[threadIdx.x] = Gs;
    __syncthreads();
    for (int s = 512; s >= 1; s >>= 1) {
        if ((int)threadIdx.x < s) {
            st[threadIdx.x] = crit_max ? fmax(st[threadIdx.x], st[threadIdx.x + s]) : st[threadIdx.x] + st[threadIdx.x + s];
            sg[threadIdx.x] += sg[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = st[0];
        out[1] = sg[0];
    }
}

// ---------------------------------------------------------- launchers
static inline unsigned blocks_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

template <typename R, int G, int A, int S>
static hipError_t launch_sweep_t(const SweepArgs<R> &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_sweep<R, G, A, S>), dim3(blocks_for(4 * a.g.Vh)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

template <typename R, int G, int S>
static hipError_t launch_sweep_a(int algo, const SweepArgs<R> &a, hipStream_t s)
{
    switch (algo) {
    case GF_OR: return launch_sweep_t<R, G, GF_OR, S>(a, s);
    case GF_MICRO: return launch_sweep_t<R, G, GF_MICRO, S>(a, s);
    case GF_SA: return launch_sweep_t<R, G, GF_SA, S>(a, s);
    case GF_SR: return launch_sweep_t<R, G, GF_SR, S>(a, s);
    default: return launch_sweep_t<R, G, GF_RANDOM, S>(a, s);
    }
}

template <typename R, int S>
static hipError_t launch_sweep_g(int gauge, int algo, const SweepArgs<R> &a, hipStream_t s)
{
    switch (gauge) {
    case GF_LANDAU: return launch_sweep_a<R, GF_LANDAU, S>(algo, a, s);
    case GF_COULOMB: return launch_sweep_a<R, GF_COULOMB, S>(algo, a, s);
    case GF_MAG: return launch_sweep_a<R, GF_MAG, S>(algo, a, s);
    default: return launch_sweep_a<R, GF_U1XU1, S>(algo, a, s);
    }
}

template <typename R>
hipError_t launch_sweep(int stride, int gauge, int algo, const SweepArgs<R> &a, hipStream_t s)
{
    if (stride == 18) return launch_sweep_g<R, 18>(gauge, algo, a, s);
    return launch_sweep_g<R, 12>(gauge, algo, a, s);
}
template hipError_t launch_sweep<float>(int, int, int, const SweepArgs<float> &, hipStream_t);
template hipError_t launch_sweep<double>(int, int, int, const SweepArgs<double> &, hipStream_t);

template <typename R>
hipError_t launch_project(int stride, R *U, int64_t V, hipStream_t s)
{
    if (stride == 18) hipLaunchKernelGGL((k_project<R, 18>), dim3(blocks_for(4 * V)), dim3(kBlock), 0, s, U, V);
    else hipLaunchKernelGGL((k_project<R, 12>), dim3(blocks_for(4 * V)), dim3(kBlock), 0, s, U, V);
    return hipGetLastError();
}
template hipError_t launch_project<float>(int, float *, int64_t, hipStream_t);
template hipError_t launch_project<double>(int, double *, int64_t, hipStream_t);

template <typename RI, typename RO>
static hipError_t launch_convert_t(int si, int so, const RI *src, RO *dst, int64_t V, int reun, hipStream_t s)
{
    const dim3 grid(blocks_for(4 * V)), block(kBlock);
    if (si == 18 && so == 18) hipLaunchKernelGGL((k_convert<RI, 18, RO, 18>), grid, block, 0, s, src, dst, V, reun);
    else if (si == 18 && so == 12) hipLaunchKernelGGL((k_convert<RI, 18, RO, 12>), grid, block, 0, s, src, dst, V, reun);
    else if (si == 12 && so == 18) hipLaunchKernelGGL((k_convert<RI, 12, RO, 18>), grid, block, 0, s, src, dst, V, reun);
    else hipLaunchKernelGGL((k_convert<RI, 12, RO, 12>), grid, block, 0, s, src, dst, V, reun);
    return hipGetLastError();
}

hipError_t launch_convert(int ti, int si, const void *src, int to, int so, void *dst, int64_t V, int reun,
                          hipStream_t s)
{
    if (ti == GF_F64 && to == GF_F64) return launch_convert_t(si, so, (const double *)src, (double *)dst, V, reun, s);
    if (ti == GF_F64 && to == GF_F32) return launch_convert_t(si, so, (const double *)src, (float *)dst, V, reun, s);
    if (ti == GF_F32 && to == GF_F64) return launch_convert_t(si, so, (const float *)src, (double *)dst, V, reun, s);
    return launch_convert_t(si, so, (const float *)src, (float *)dst, V, reun, s);
}

template <typename R>
hipError_t launch_set_hot(R *U, int64_t V, uint32_t seed, uint32_t counter, hipStream_t s)
{
    hipLaunchKernelGGL((k_set_hot<R>), dim3(blocks_for(V)), dim3(kBlock), 0, s, U, V, seed, counter);
    return hipGetLastError();
}
template hipError_t launch_set_hot<float>(float *, int64_t, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_set_hot<double>(double *, int64_t, uint32_t, uint32_t, hipStream_t);

int64_t quality_blocks(int64_t V) { return (int64_t)blocks_for(4 * V); }

template <typename R, int S>
static hipError_t launch_quality_s(int gauge, const QualityArgs<R> &a, hipStream_t s)
{
    const dim3 grid(blocks_for(4 * a.g.V)), block(kBlock);
    switch (gauge) {
    case GF_COULOMB: hipLaunchKernelGGL((k_quality<R, GF_COULOMB, S>), grid, block, 0, s, a); break;
    case GF_MAG: hipLaunchKernelGGL((k_quality<R, GF_MAG, S>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((k_quality<R, GF_LANDAU, S>), grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}

template <typename R>
hipError_t launch_quality(int stride, int gauge, const QualityArgs<R> &a, hipStream_t s)
{
    if (stride == 18) return launch_quality_s<R, 18>(gauge, a, s);
    return launch_quality_s<R, 12>(gauge, a, s);
}
template hipError_t launch_quality<float>(int, int, const QualityArgs<float> &, hipStream_t);
template hipError_t launch_quality<double>(int, int, const QualityArgs<double> &, hipStream_t);

hipError_t launch_reduce_final(const double *partial, int64_t n, int crit_max, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, s, partial, n, crit_max, out);
    return hipGetLastError();
}

}  // namespace su3gf
