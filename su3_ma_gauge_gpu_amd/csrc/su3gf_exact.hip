// su3gf_exact.hip -- kernels of the drop-in operator API, compiled with
// -ffp-contract=off: every multiply and add rounds separately and in the
// reference's order, sums over a site's links follow the reference's ladder
// and 1/sqrt stands for rsqrt, so an FP64 half-sweep reproduces the CPU
// restatement of the reference bit for bit.
//
// Kernels: the sweep (su3gf_sweep.h, EXACT = true), projectSU3
// (CommonKernelsSU3.hxx:71-90), setHot (CommonKernelsSU3.hxx:92-130) and the
// per-site quality (LandauKernelsSU3.hxx:114-191, CoulombKernelsSU3.hxx:110-183,
// MAGKernelsSU3.hxx:116-223) with a deterministic block reduction.
#include "su3gf_sweep.h"

namespace su3gf {
namespace exact {

static inline unsigned blocks_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------- sweep
template <typename R, int G, int A, int S>
static hipError_t launch_t(const SweepArgs<R> &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_sweep<R, G, A, S, true, 0, false>), dim3((unsigned)sweep_grid(a.g)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

template <typename R, int G, int S>
static hipError_t launch_a(int algo, const SweepArgs<R> &a, hipStream_t s)
{
    if constexpr (sizeof(R) == 4) {  // FP32 links, update rule in double
        switch (algo) {
        case GF_OR + kDpUpdate: return launch_t<R, G, GF_OR + kDpUpdate, S>(a, s);
        case GF_MICRO + kDpUpdate: return launch_t<R, G, GF_MICRO + kDpUpdate, S>(a, s);
        case GF_SA + kDpUpdate: return launch_t<R, G, GF_SA + kDpUpdate, S>(a, s);
        default: break;
        }
    }
    switch (algo) {
    case GF_OR: return launch_t<R, G, GF_OR, S>(a, s);
    case GF_MICRO: return launch_t<R, G, GF_MICRO, S>(a, s);
    case GF_SA: return launch_t<R, G, GF_SA, S>(a, s);
    case GF_SR: return launch_t<R, G, GF_SR, S>(a, s);
    case GF_RANDOM: return launch_t<R, G, GF_RANDOM, S>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <typename R, int S>
static hipError_t launch_g(int gauge, int algo, const SweepArgs<R> &a, hipStream_t s)
{
    switch (gauge) {
    case GF_LANDAU: return launch_a<R, GF_LANDAU, S>(algo, a, s);
    case GF_COULOMB: return launch_a<R, GF_COULOMB, S>(algo, a, s);
    case GF_MAG: return launch_a<R, GF_MAG, S>(algo, a, s);
    default: return launch_a<R, GF_U1XU1, S>(algo, a, s);
    }
}

template <typename R>
hipError_t launch_sweep(int stride, int gauge, int algo, const SweepArgs<R> &a, hipStream_t s)
{
    if (stride == 18) return launch_g<R, 18>(gauge, algo, a, s);
    return launch_g<R, 12>(gauge, algo, a, s);
}
template hipError_t launch_sweep<float>(int, int, int, const SweepArgs<float> &, hipStream_t);
template hipError_t launch_sweep<double>(int, int, int, const SweepArgs<double> &, hipStream_t);

// ----------------------------------------------------------------- project
// one thread per link of `nfields` fields (4 link directions, or 1 for a
// gauge-transformation array)
template <typename R, int S>
__global__ __launch_bounds__(kBlock) void k_project(R *U, int64_t V, int nfields, const DevState *st)
{
    if (st && st->done) return;  // a driver's run has converged: the checked state stays as it is
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nfields * V) return;
    const int mu = (int)(t / V);
    const int64_t idx = t - (int64_t)mu * V;
    const LinkRef<R> l = link_ref(U, V, S, idx, mu);
    M3<R> u;
    load12(l, u);
    project_e<true>(u);
    store12(l, u);
    if (S == 18) store_row2(l, u);
}

template <typename R>
hipError_t launch_project(int stride, int nfields, R *U, int64_t V, hipStream_t s, const DevState *st)
{
    const dim3 grid(blocks_for(nfields * V)), block(kBlock);
    if (stride == 18) hipLaunchKernelGGL((k_project<R, 18>), grid, block, 0, s, U, V, nfields, st);
    else if (stride == kTiled12 && sizeof(R) == 8)
        hipLaunchKernelGGL((k_project<R, kTiled12>), grid, block, 0, s, U, V, nfields, st);
    else if (stride == 12) hipLaunchKernelGGL((k_project<R, 12>), grid, block, 0, s, U, V, nfields, st);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}
template hipError_t launch_project<float>(int, int, float *, int64_t, hipStream_t, const DevState *);
template hipError_t launch_project<double>(int, int, double *, int64_t, hipStream_t, const DevState *);

// ----------------------------------------------------------------- set hot
// per site: identity times three random SU(2) subgroup elements from the right,
// (0,1), (0,2), (1,2), each from four uniforms in (-1,1) normalised
// (Quaternion::projectSU2, Quaternion.hxx:220-228); all 18 reals written.
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
#define SU3GF_HOT_SUB(I, J)                                                           \
        {                                                                             \
            R q[4];                                                                   \
            _Pragma("unroll") for (int k = 0; k < 4; k++) q[k] = philox_rand(rng, R(0)) * (R)2 - (R)1; \
            const R n = (R)1 / sqrt_(q[0] * q[0] + q[3] * q[3] + q[2] * q[2] + q[1] * q[1]); \
            _Pragma("unroll") for (int k = 0; k < 4; k++) q[k] *= n;                 \
            /* right multiplication by G(q) = right_mult_dag by G(q)^dagger */       \
            const R h[4] = {q[0], -q[1], -q[2], -q[3]};                               \
            right_mult_dag<R, I, J>(u, h);                                            \
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

template <typename R>
hipError_t launch_set_hot(R *U, int64_t V, uint32_t seed, uint32_t counter, hipStream_t s)
{
    hipLaunchKernelGGL((k_set_hot<R>), dim3(blocks_for(V)), dim3(kBlock), 0, s, U, V, seed, counter);
    return hipGetLastError();
}
template hipError_t launch_set_hot<float>(float *, int64_t, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_set_hot<double>(double *, int64_t, uint32_t, uint32_t, hipStream_t);

// ----------------------------------------------------------------- quality
// One thread per site, in the reference's order of operations
// (LandauKernelsSU3.hxx:114-191, CoulombKernelsSU3.hxx:110-183,
// MAGKernelsSU3.hxx:116-223): the per-site values are those of the reference
// bit for bit.  A block of 256 threads takes the 128 even and the 128 odd
// sites of parity-split index [128 L, 128 L + 128), i.e. 256 consecutive
// sites of the lexicographic order, so that most down links a site reads
// (its z and y neighbours') are the up links its own block reads, and the
// blocks are dealt out to the eight XCDs in contiguous ranges (t innermost
// where the timeslices allow), so that the x and t neighbours' links are
// still in that XCD's L2: each link comes from HBM about once.
template <typename R, int S>
__device__ __forceinline__ void load_site_link(const QualityArgs<R> &a, int64_t idx, int mu, M3<R> &u)
{
    load12(link_ref(const_cast<R *>(a.U), a.g.V, S, idx, mu), u);
    reconstruct_e<true>(u);
}

template <typename R, int G, int S>
__device__ __forceinline__ void quality_site(const QualityArgs<R> &a, int64_t site, double &dA, double &dGff)
{
    int c[4], par;
    site_coords(a.g, site, c, par);
    if (G == GF_MAG) {
        double gff = 0.0;
        double Xx[3] = {0, 0, 0}, Xy[3] = {0, 0, 0};
#pragma unroll 1
        for (int mu = 0; mu < 4; mu++) {
            M3<R> up, dw;
            load_site_link<R, S>(a, site, mu, up);
            load_site_link<R, S>(a, backward(a.g, c, par, mu), mu, dw);
#pragma unroll
            for (int j = 0; j < 3; j++) gff += (double)abs2(up.e[j][j]);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = i + 1; j < 3; j++) {
                    Cx<R> cc = up.e[i][i] * conj(up.e[j][i]) - up.e[i][j] * conj(up.e[j][j]);
                    Xx[i + j - 1] += (double)cc.x;
                    Xy[i + j - 1] += (double)cc.y;
                    // getSubgroupMatrixHermitian of the down link (SU3.hxx:231-241)
                    const Cx<R> b00 = conj(dw.e[i][i]), b01 = conj(dw.e[j][i]), b10 = conj(dw.e[i][j]),
                                b11 = conj(dw.e[j][j]);
                    cc = b00 * conj(b10) - b01 * conj(b11);
                    Xx[i + j - 1] += (double)cc.x;
                    Xy[i + j - 1] += (double)cc.y;
                }
        }
        dGff = gff;
        dA = sqrt(Xx[0] * Xx[0] + Xy[0] * Xy[0]) + sqrt(Xx[1] * Xx[1] + Xy[1] * Xy[1]) +
             sqrt(Xx[2] * Xx[2] + Xy[2] * Xy[2]);
        return;
    }
    const int mu0 = (G == GF_COULOMB) ? 1 : 0;
    M3<R> sum;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) sum.e[i][j] = cx((R)0, (R)0);
    // gff: Re tr of the up links, each trace summed from 0 in the order
    // (0,0), (1,1), (2,2), directions in order -- the reference's second loop
    // over the same links (LandauKernelsSU3.hxx:176-188), formed here from
    // the first loop's loads
    double result = 0;
    // one direction's two links at a time (140 VGPRs in FP64, 3 waves per
    // SIMD); unrolled, all eight links' loads are hoisted (256 VGPRs, 1 wave)
#pragma unroll 1
    for (int mu = mu0; mu < 4; mu++) {
        M3<R> t;
        load_site_link<R, S>(a, site, mu, t);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) sum.e[i][j] = sum.e[i][j] + t.e[i][j];
        Cx<R> tt = cx((R)0, (R)0);
        tt = tt + t.e[0][0];
        tt = tt + t.e[1][1];
        tt = tt + t.e[2][2];
        result += (double)tt.x;
        load_site_link<R, S>(a, backward(a.g, c, par, mu), mu, t);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) sum.e[i][j] = sum.e[i][j] - t.e[i][j];
    }
    // Sum -= Sum.trace()/3 on the diagonal; Sum -= Sum^dagger
    Cx<R> tr = cx((R)0, (R)0);
    tr = tr + sum.e[0][0];
    tr = tr + sum.e[1][1];
    tr = tr + sum.e[2][2];
    tr = cx(tr.x / (R)3, tr.y / (R)3);
#pragma unroll
    for (int i = 0; i < 3; i++) sum.e[i][i] = sum.e[i][i] - tr;
    M3<R> h;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) h.e[j][i] = conj(sum.e[i][j]);
    double prec = 0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const Cx<R> d = sum.e[i][j] - h.e[i][j];
            prec += (double)abs2(d);
        }
    dA = prec;
    dGff = result;
}

// logical block of this block, or -1 for a block past the last one
template <typename R> __device__ __forceinline__ int64_t quality_block(const QualityArgs<R> &a)
{
    const uint32_t b = blockIdx.x;
    if (a.remap == 0) return b;
    const uint32_t xcd = b & 7, k = b >> 3;
    if (a.remap == 1) {
        const int64_t per = (a.nblocks + 7) / 8;
        const int64_t L = (int64_t)xcd * per + k;
        return L < a.nblocks ? L : -1;
    }
    const uint32_t P = (uint32_t)a.tpass;
    const uint32_t pass = fdiv(k, a.pass_div);
    const uint32_t rem = k - pass * P * (uint32_t)a.slice_blocks;
    const uint32_t q = fdiv(rem, a.tpass_div);
    const int64_t t = (int64_t)xcd * a.slices_per_xcd + pass * P + (rem - q * P);
    return t * a.slice_blocks + q;
}

constexpr int kQualitySites = 128;  // sites of each parity per block

template <typename R, int G, int S>
__global__ __launch_bounds__(kBlock) void k_quality(QualityArgs<R> a)
{
    const int64_t L = quality_block(a);
    const int par = threadIdx.x >= kQualitySites;
    const int64_t p = L * kQualitySites + (threadIdx.x & (kQualitySites - 1));
    const bool active = L >= 0 && p < a.g.Vh;
    double th = 0, gf = 0;
    if (active) {
        const int64_t site = p + (par ? a.g.Vh : 0);
        quality_site<R, G, S>(a, site, th, gf);
        if (a.dGff) {
            a.dGff[site] = gf;
            a.dA[site] = th;
        }
    }
    if (!a.partial || L < 0) return;  // both uniform over the block
    __shared__ double sh[2][kBlock / 64];
    double ts = th, gs = gf;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(ts, off, 64);
        ts = a.crit_max ? fmax(ts, o) : ts + o;
        gs += __shfl_xor(gs, off, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][wv] = ts;
        sh[1][wv] = gs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double T = sh[0][0], Gs = sh[1][0];
        for (int k = 1; k < kBlock / 64; k++) {
            T = a.crit_max ? fmax(T, sh[0][k]) : T + sh[0][k];
            Gs += sh[1][k];
        }
        a.partial[2 * L] = T;
        a.partial[2 * L + 1] = Gs;
    }
}

int64_t quality_blocks(int64_t V) { return (V / 2 + kQualitySites - 1) / kQualitySites; }

template <typename R, int S>
static hipError_t launch_quality_s(int gauge, QualityArgs<R> a, hipStream_t s)
{
    static_assert(2 * kQualitySites == kBlock, "a block holds both parities' sites");
    a.nblocks = quality_blocks(a.g.V);
    const int64_t half_slice = a.g.Vh / a.g.N[0];
    unsigned grid;
    if (a.g.N[0] % 8 == 0 && half_slice % kQualitySites == 0 && a.nblocks % 8 == 0) {
        a.remap = 2;
        a.slice_blocks = half_slice / kQualitySites;
        a.slices_per_xcd = a.g.N[0] / 8;
        if (a.tpass <= 0 || a.slices_per_xcd % a.tpass != 0) a.tpass = a.slices_per_xcd;
        a.tpass_div = make_fastdiv((uint32_t)a.tpass);
        a.pass_div = make_fastdiv((uint32_t)(a.tpass * a.slice_blocks));
        grid = (unsigned)a.nblocks;
    } else {
        a.remap = 1;
        grid = (unsigned)(8 * ((a.nblocks + 7) / 8));
    }
    switch (gauge) {
    case GF_COULOMB: hipLaunchKernelGGL((k_quality<R, GF_COULOMB, S>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case GF_MAG: hipLaunchKernelGGL((k_quality<R, GF_MAG, S>), dim3(grid), dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL((k_quality<R, GF_LANDAU, S>), dim3(grid), dim3(kBlock), 0, s, a); break;
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

}  // namespace exact
}  // namespace su3gf
