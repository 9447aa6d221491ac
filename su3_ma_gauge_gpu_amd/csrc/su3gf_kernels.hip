// su3gf_kernels.hip -- utility kernels: layout/precision conversion, the
// final fixed-order reduction of quality partials, and the gauge-transformation
// kernels of the mixed-precision driver (identity, reprojection, and the FP64
// application out_mu(x) = G(x) U0_mu(x) G(x+mu)^dagger).
#include "su3gf_sweep.h"

namespace su3gf {

static inline unsigned blocks_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

// ----------------------------------------------------------------- convert
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
    if (reunitarize) project_e<false>(b);
    else reconstruct_e<false>(b);
    store12(ld, b);
    if (SO == 18) store_row2(ld, b);
}

template <typename RI, typename RO>
static hipError_t launch_convert_t(int si, int so, const RI *src, RO *dst, int64_t V, int reun, hipStream_t s)
{
    const dim3 grid(blocks_for(4 * V)), block(kBlock);
    if (si == kTiled12 || so == kTiled12) {
        // the drivers' FP64 tiles, to and from the caller's FP64 layouts
        if constexpr (sizeof(RI) == 8 && sizeof(RO) == 8) {
            if (si == 18 && so == kTiled12)
                hipLaunchKernelGGL((k_convert<RI, 18, RO, kTiled12>), grid, block, 0, s, src, dst, V, reun);
            else if (si == 12 && so == kTiled12)
                hipLaunchKernelGGL((k_convert<RI, 12, RO, kTiled12>), grid, block, 0, s, src, dst, V, reun);
            else if (si == kTiled12 && so == 18)
                hipLaunchKernelGGL((k_convert<RI, kTiled12, RO, 18>), grid, block, 0, s, src, dst, V, reun);
            else if (si == kTiled12 && so == 12)
                hipLaunchKernelGGL((k_convert<RI, kTiled12, RO, 12>), grid, block, 0, s, src, dst, V, reun);
            else
                return hipErrorInvalidValue;
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
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

// ---------------------------------------------------------- final reduction
// Fixed-order reduction of block partials: thread k sums partials k, k+1024,
// ... in order, then a fixed tree -> out[0] = theta, out[1] = gff.
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

hipError_t launch_reduce_final(const double *partial, int64_t n, int crit_max, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, s, partial, n, crit_max, out);
    return hipGetLastError();
}

// ------------------------------------------------- gauge transformations
// g: FP32 rows 0,1 per site in tiles of 64 sites (gauge_ref)
__global__ __launch_bounds__(kBlock) void k_gauge_identity(float *g, int64_t V)
{
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= V) return;
    M3<float> u;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) u.e[i][j] = cx((float)(i == j), 0.f);
    store12(gauge_ref(g, idx), u);
}

hipError_t launch_gauge_identity(float *g, int64_t V, hipStream_t s)
{
    hipLaunchKernelGGL(k_gauge_identity, dim3(blocks_for(V)), dim3(kBlock), 0, s, g, V);
    return hipGetLastError();
}

// g <- FP32 rounding of the FP64 Gram-Schmidt projection of g
__global__ __launch_bounds__(kBlock) void k_project_gauge(float *g, int64_t V)
{
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= V) return;
    const LinkRef<float> l = gauge_ref(g, idx);
    M3<float> f;
    load12(l, f);
    M3<double> d;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) d.e[i][j] = cx((double)f.e[i][j].x, (double)f.e[i][j].y);
    project_e<false>(d);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) f.e[i][j] = cx((float)d.e[i][j].x, (float)d.e[i][j].y);
    store12(l, f);
}

hipError_t launch_project_gauge(float *g, int64_t V, hipStream_t s)
{
    hipLaunchKernelGGL(k_project_gauge, dim3(blocks_for(V)), dim3(kBlock), 0, s, g, V);
    return hipGetLastError();
}

__device__ __forceinline__ void gauge_at(const float *g, int64_t idx, M3<double> &G)
{
    M3<float> f;
    load12(gauge_ref(const_cast<float *>(g), idx), f);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) G.e[i][j] = cx((double)f.e[i][j].x, (double)f.e[i][j].y);
    project_e<false>(G);
}

// out_mu(x) = G(x) U0_mu(x) G(x+mu)^dagger, one thread per link; in place when
// out == U0 (each thread reads and writes only its own link).
template <int S0, typename RO, int SO>
__global__ __launch_bounds__(kBlock) void k_gauge_apply(const double *U0, const float *g, RO *out, Geom geo)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * geo.V) return;
    const int mu = (int)(t / geo.V);
    const int64_t idx = t - (int64_t)mu * geo.V;
    int c[4], par;
    site_coords(geo, idx, c, par);
    const int64_t fw = forward(geo, c, par, mu);
    M3<double> Gx, Gf, U;
    gauge_at(g, idx, Gx);
    gauge_at(g, fw, Gf);
    load12(link_ref(const_cast<double *>(U0), geo.V, S0, idx, mu), U);
    reconstruct_e<false>(U);
    // T = Gx U ; M = T Gf^dagger (rows 0,1)
    M3<double> T;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cx<double> s = Gx.e[i][0] * U.e[0][j];
            s = s + Gx.e[i][1] * U.e[1][j];
            s = s + Gx.e[i][2] * U.e[2][j];
            T.e[i][j] = s;
        }
    M3<RO> M;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cx<double> s = mul_conj(T.e[i][0], Gf.e[j][0]);
            s = s + mul_conj(T.e[i][1], Gf.e[j][1]);
            s = s + mul_conj(T.e[i][2], Gf.e[j][2]);
            M.e[i][j] = cx((RO)s.x, (RO)s.y);
        }
    reconstruct_e<false>(M);
    const LinkRef<RO> lo = link_ref(out, geo.V, SO, idx, mu);
    store12(lo, M);
    if (SO == 18) store_row2(lo, M);
}

hipError_t launch_gauge_apply(const double *U0, int s0, const float *g, void *out, int to, int so, const Geom &geo,
                              hipStream_t s)
{
    const dim3 grid(blocks_for(4 * geo.V)), block(kBlock);
    if (to == GF_F64) {
        double *o = (double *)out;
        if (so == kTiled12) {  // into the drivers' FP64 tiles
            if (s0 == 18) hipLaunchKernelGGL((k_gauge_apply<18, double, kTiled12>), grid, block, 0, s, U0, g, o, geo);
            else hipLaunchKernelGGL((k_gauge_apply<12, double, kTiled12>), grid, block, 0, s, U0, g, o, geo);
            return hipGetLastError();
        }
        if (s0 == 18 && so == 18) hipLaunchKernelGGL((k_gauge_apply<18, double, 18>), grid, block, 0, s, U0, g, o, geo);
        else if (s0 == 18 && so == 12) hipLaunchKernelGGL((k_gauge_apply<18, double, 12>), grid, block, 0, s, U0, g, o, geo);
        else if (s0 == 12 && so == 18) hipLaunchKernelGGL((k_gauge_apply<12, double, 18>), grid, block, 0, s, U0, g, o, geo);
        else hipLaunchKernelGGL((k_gauge_apply<12, double, 12>), grid, block, 0, s, U0, g, o, geo);
    } else {
        if (so != 12) return hipErrorInvalidValue;
        float *o = (float *)out;
        if (s0 == 18) hipLaunchKernelGGL((k_gauge_apply<18, float, 12>), grid, block, 0, s, U0, g, o, geo);
        else hipLaunchKernelGGL((k_gauge_apply<12, float, 12>), grid, block, 0, s, U0, g, o, geo);
    }
    return hipGetLastError();
}

// The FP64 phase of the mixed driver sweeps an FP64 g (12 doubles per site
// in 64-site tiles, kTiled12's layout with one field) on read-only FP64 links.
__global__ __launch_bounds__(kBlock) void k_gauge_identity64(double *g, int64_t V)
{
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= V) return;
    M3<double> u;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) u.e[i][j] = cx((double)(i == j), 0.0);
    store12(link_ref(g, V, kTiled12, idx, 0), u);
}

hipError_t launch_gauge_identity64(double *g, int64_t V, hipStream_t s)
{
    hipLaunchKernelGGL(k_gauge_identity64, dim3(blocks_for(V)), dim3(kBlock), 0, s, g, V);
    return hipGetLastError();
}

__device__ __forceinline__ void gauge64_at(const double *g, int64_t V, int64_t idx, M3<double> &G)
{
    load12(link_ref(const_cast<double *>(g), V, kTiled12, idx, 0), G);
    project_e<false>(G);
}

// out_mu(x) = G(x) U1_mu(x) G(x+mu)^dagger, G the FP64 g (reprojected), U1
// the FP64 links in kTiled12; out in layout SO (18 or 12), reprojected when
// reun (the final projectSU3 of the run), one thread per link
template <int SO>
__global__ __launch_bounds__(kBlock) void k_gauge_apply64(const double *U1, const double *g, double *out, Geom geo,
                                                          int reun)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * geo.V) return;
    const int mu = (int)(t / geo.V);
    const int64_t idx = t - (int64_t)mu * geo.V;
    int c[4], par;
    site_coords(geo, idx, c, par);
    const int64_t fw = forward(geo, c, par, mu);
    M3<double> Gx, Gf, U;
    gauge64_at(g, geo.V, idx, Gx);
    gauge64_at(g, geo.V, fw, Gf);
    load12(link_ref(const_cast<double *>(U1), geo.V, kTiled12, idx, mu), U);
    reconstruct_e<false>(U);
    M3<double> T, M;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cx<double> s = Gx.e[i][0] * U.e[0][j];
            s = s + Gx.e[i][1] * U.e[1][j];
            s = s + Gx.e[i][2] * U.e[2][j];
            T.e[i][j] = s;
        }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Cx<double> s = mul_conj(T.e[i][0], Gf.e[j][0]);
            s = s + mul_conj(T.e[i][1], Gf.e[j][1]);
            s = s + mul_conj(T.e[i][2], Gf.e[j][2]);
            M.e[i][j] = s;
        }
    if (reun) project_e<false>(M);
    else reconstruct_e<false>(M);
    const LinkRef<double> lo = link_ref(out, geo.V, SO, idx, mu);
    store12(lo, M);
    if (SO == 18) store_row2(lo, M);
}

hipError_t launch_gauge_apply64(const double *U1, const double *g, double *out, int so, const Geom &geo, int reun,
                                hipStream_t s)
{
    const dim3 grid(blocks_for(4 * geo.V)), block(kBlock);
    if (so == 18) hipLaunchKernelGGL((k_gauge_apply64<18>), grid, block, 0, s, U1, g, out, geo, reun);
    else if (so == 12) hipLaunchKernelGGL((k_gauge_apply64<12>), grid, block, 0, s, U1, g, out, geo, reun);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// ------------------------------------------------------------ driver state
__global__ void k_state_init(DevState *st, double tol, double theta_scale, double gff_scale, int stall_limit,
                             double stall_below, int stall_rule, int exact_end, double near_below, int near)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->done = 0;
    st->converged_sweep = -1;
    st->checked_sweep = -1;
    st->checks = 0;
    st->stall = 0;
    st->stall_limit = stall_limit;
    st->stall_below = stall_below;
    st->stall_rule = stall_rule;
    st->best = 1e300;
    st->windows = 0;
    st->win_logsum = 0;
    st->win_mean = 0;
    st->win_drop = 0;
    st->tol = tol;
    st->theta_scale = theta_scale;
    st->gff_scale = gff_scale;
    st->theta = 0;
    st->gff = 0;
    st->exact_end = exact_end;
    st->near = exact_end && near;
    st->near_below = near_below;
}

hipError_t launch_state_init(DevState *st, double tol, double theta_scale, double gff_scale, int stall_limit,
                             double stall_below, int stall_rule, int exact_end, double near_below, int near,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st, tol, theta_scale, gff_scale, stall_limit,
                       stall_below, stall_rule, exact_end, near_below, near);
    return hipGetLastError();
}

// Fixed-order reduction of the even partials of sweep `sweep` and the odd
// partials of sweep `sweep - 1`, then the convergence and stagnation tests.
// With nseg == 1 the whole lattice is one problem: thread k sums blocks k,
// k + 1024, ... in order, then a fixed tree.  With nseg > 1 the blocks form
// nseg equal consecutive segments (the timeslices of a Coulomb lattice fixed
// all at once): wave w reduces segments w, w + 16, ... (lane k sums blocks
// k, k + 64, ... of the segment, then a fixed shuffle tree), theta is the
// largest segment theta -- every slice must meet the precision, as when the
// reference fixes them one after another -- and gff the average.
constexpr int kMaxSeg = 4096;

// updated: the even kernel that formed the even sites' theta has already
// applied its update (CHECK_FUSED*), so with exact_end a theta below tol does
// not end the phase (the state it describes is gone); it sets `near`, and the
// next check, made before its even update, decides
__device__ __forceinline__ void check_update(DevState *st, double theta, double gff, int sweep, bool updated)
{
    st->theta = theta;
    st->gff = gff;
    st->checked_sweep = sweep - 1;
    st->checks++;
    // stagnation, tested only for theta below stall_below (above it theta may
    // still rise for a while: overrelaxation is not monotone, and a Coulomb
    // lattice tests its worst slice)
    bool stalled = false;
    if (st->stall_rule == 0) {
        // stall_limit checks in a row without lowering the best theta by 10%:
        // single-precision sweeps at their rounding floor gain nothing
        if (theta < 0.9 * st->best) {
            st->best = theta;
            st->stall = 0;
        } else if (theta < st->stall_below) {
            st->stall++;
        } else {
            st->stall = 0;
        }
        stalled = st->stall_limit > 0 && st->stall >= st->stall_limit;
    } else if (theta < st->stall_below && theta > 0) {
        // windows of stall_limit checks, for runs that converge slowly (MAG at
        // large volume: a few percent per window) far above their rounding
        // floor.  A window's drop is the decrease of the mean of log(theta)
        // from the window before; the run has stagnated when a drop is below
        // log(1.1) and either below log(1.001) or below a quarter of the
        // previous drop, so a steady slow rate goes on.  Means rather than
        // minima keep the check-to-check scatter of theta from passing for a
        // change of rate.
        st->win_logsum += log(theta);
        if (++st->stall >= st->stall_limit && st->stall_limit > 0) {
            const double mean = st->win_logsum / st->stall;
            if (st->windows > 0) {
                const double drop = st->win_mean - mean;
                stalled = drop < log(1.1) && (drop < log(1.001) || (st->windows > 1 && drop < 0.25 * st->win_drop));
                st->win_drop = drop;
            }
            st->win_mean = mean;
            st->windows++;
            st->stall = 0;
            st->win_logsum = 0;
        }
    } else {
        st->stall = 0;
        st->win_logsum = 0;
        st->windows = 0;
    }
    if (theta < st->tol && !(updated && st->exact_end)) {
        st->done = 1;
        st->converged_sweep = sweep - 1;
    } else if (stalled) {
        st->done = 2;
        st->converged_sweep = sweep - 1;
    }
    if (st->exact_end && theta < st->near_below) st->near = 1;
}

__global__ __launch_bounds__(1024) void k_fused_check(DevState *st, const double *even, const double *odd,
                                                      int64_t nblocks, int nseg, int sweep, int mode)
{
    if (st->done) return;
    if (mode == CHECK_PRE && !st->near) return;
    if (mode == CHECK_FUSED_IF_FAR && st->near) return;
    const bool updated = mode != CHECK_PRE;
    __shared__ double sa[1024], sb[1024], sc[1024];
    if (nseg <= 1) {
        double te = 0, ge = 0, to = 0;
        for (int64_t k = threadIdx.x; k < nblocks; k += 1024) {
            te += even[2 * k];
            ge += even[2 * k + 1];
            to += odd[2 * k];
        }
        sa[threadIdx.x] = te;
        sb[threadIdx.x] = ge;
        sc[threadIdx.x] = to;
        __syncthreads();
        for (int s = 512; s >= 1; s >>= 1) {
            if ((int)threadIdx.x < s) {
                sa[threadIdx.x] += sa[threadIdx.x + s];
                sb[threadIdx.x] += sb[threadIdx.x + s];
                sc[threadIdx.x] += sc[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0)
            check_update(st, (sa[0] + sc[0]) * st->theta_scale, sb[0] * st->gff_scale, sweep, updated);
        return;
    }
    __shared__ double seg_theta[kMaxSeg];
    __shared__ double seg_gff[kMaxSeg];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t per = nblocks / nseg;
    for (int g = wave; g < nseg; g += 16) {
        double t = 0, f = 0;
        for (int64_t k = lane; k < per; k += 64) {
            const int64_t b = (int64_t)g * per + k;
            t += even[2 * b] + odd[2 * b];
            f += even[2 * b + 1];
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            t += __shfl_xor(t, off, 64);
            f += __shfl_xor(f, off, 64);
        }
        if (lane == 0) {
            seg_theta[g] = t;
            seg_gff[g] = f;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tmax = 0, ftot = 0;
        for (int g = 0; g < nseg; g++) {
            tmax = fmax(tmax, seg_theta[g]);
            ftot += seg_gff[g];
        }
        check_update(st, tmax * st->theta_scale * nseg, ftot * st->gff_scale, sweep, updated);
    }
}

hipError_t launch_fused_check(DevState *st, const double *even, const double *odd, int64_t nblocks, int nseg,
                              int sweep, int mode, hipStream_t s)
{
    if (nseg < 1 || nseg > kMaxSeg || nblocks % nseg) return hipErrorInvalidValue;
    if (mode < CHECK_FUSED || mode > CHECK_FUSED_IF_FAR) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fused_check, dim3(1), dim3(1024), 0, s, st, even, odd, nblocks, nseg, sweep, mode);
    return hipGetLastError();
}

}  // namespace su3gf
