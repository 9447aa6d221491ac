// su3gf_api.hip -- implementation of the C ABI declared in include/su3gf.h.
//
// Host-side validation, dispatch to the gfx950 kernels, the quality
// reduction bookkeeping of GaugeFixingStats and the OR / mixed-precision
// drivers.  Every HIP call is checked; failures come back as gf_status with
// a message retrievable through gf_last_error().
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "su3gf_kernels.h"

using namespace su3gf;

namespace {

thread_local std::string g_err;

gf_status fail(gf_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

gf_status hip_fail(hipError_t e, const char *what)
{
    return fail(GF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call)                                             \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return hip_fail(e_, #call);         \
    } while (0)

#define GF_TRY(call)                                              \
    do {                                                          \
        gf_status s_ = (call);                                    \
        if (s_ != GF_OK) return s_;                               \
    } while (0)

FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    f.d = d;
    f.s = 0;
    while ((1ull << f.s) < d) f.s++;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << f.s) - d)) / d + 1);
    return f;
}

gf_status check_dims(const int dims[4], int64_t *V)
{
    if (!dims) return fail(GF_ERR_BAD_ARG, "dims is null");
    int64_t v = 1;
    for (int i = 0; i < 4; i++) {
        if (dims[i] <= 0) return fail(GF_ERR_BAD_SHAPE, "extent %d of dimension %d is not positive", dims[i], i);
        if (i > 0 && dims[i] % 2) return fail(GF_ERR_BAD_SHAPE, "extent %d of dimension %d is odd", dims[i], i);
        if (i == 0 && dims[0] != 1 && dims[0] % 2)
            return fail(GF_ERR_BAD_SHAPE, "time extent %d is odd (only a timeslice may have extent 1)", dims[0]);
        v *= dims[i];
    }
    if (v >= (1ll << 31)) return fail(GF_ERR_BAD_SHAPE, "lattice volume %lld exceeds 2^31 sites", (long long)v);
    *V = v;
    return GF_OK;
}

gf_status check_lattice(const gf_lattice *lat, int64_t *V)
{
    if (!lat) return fail(GF_ERR_BAD_ARG, "lattice descriptor is null");
    GF_TRY(check_dims(lat->dims, V));
    if (lat->dtype != GF_F32 && lat->dtype != GF_F64) return fail(GF_ERR_BAD_ARG, "unknown dtype %d", lat->dtype);
    if (lat->layout != GF_LAYOUT_R12 && lat->layout != GF_LAYOUT_PATTERN18)
        return fail(GF_ERR_BAD_ARG, "unknown layout %d", lat->layout);
    return GF_OK;
}

Geom make_geom(const int dims[4], int64_t V)
{
    Geom g;
    for (int i = 0; i < 4; i++) g.N[i] = dims[i];
    g.V = V;
    g.Vh = V / 2;
    g.fx = make_fastdiv((uint32_t)dims[1]);
    g.fy = make_fastdiv((uint32_t)dims[2]);
    g.fz = make_fastdiv((uint32_t)dims[3]);
    return g;
}

inline hipStream_t S(void *stream) { return (hipStream_t)stream; }

}  // namespace

struct gf_stats {
    int dims[4];
    int gauge;
    int crit;
    int64_t V;
    int64_t nblocks;
    double *partial;   // device, 2 * nblocks
    double *out;       // device, 2
    double *host;      // pinned host, 2
};

extern "C" {

const char *gf_last_error(void) { return g_err.c_str(); }
const char *gf_version(void) { return "su3gf 0.1 (gfx950)"; }

// ------------------------------------------------------------ device/memory
gf_status gf_device_info(int *count, char *name, size_t name_len)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        if (count) *count = 0;
        return fail(GF_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
    }
    if (count) *count = n;
    if (name && name_len) {
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, 0));
        snprintf(name, name_len, "%s", prop.gcnArchName);
    }
    return GF_OK;
}

gf_status gf_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(GF_ERR_BAD_ARG, "ptr is null");
    HIP_TRY(hipMalloc(ptr, bytes));
    return GF_OK;
}
gf_status gf_free(void *ptr)
{
    HIP_TRY(hipFree(ptr));
    return GF_OK;
}
gf_status gf_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GF_OK;
}
gf_status gf_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GF_OK;
}
gf_status gf_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return GF_OK;
}
gf_status gf_memset(void *ptr, int value, size_t bytes)
{
    HIP_TRY(hipMemset(ptr, value, bytes));
    return GF_OK;
}
gf_status gf_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return GF_OK;
}
gf_status gf_stream_create(void **stream)
{
    if (!stream) return fail(GF_ERR_BAD_ARG, "stream is null");
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return GF_OK;
}
gf_status gf_stream_destroy(void *stream)
{
    HIP_TRY(hipStreamDestroy(S(stream)));
    return GF_OK;
}
gf_status gf_stream_synchronize(void *stream)
{
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return GF_OK;
}
gf_status gf_event_create(void **event)
{
    if (!event) return fail(GF_ERR_BAD_ARG, "event is null");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *event = e;
    return GF_OK;
}
gf_status gf_event_destroy(void *event)
{
    HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return GF_OK;
}
gf_status gf_event_record(void *event, void *stream)
{
    HIP_TRY(hipEventRecord((hipEvent_t)event, S(stream)));
    return GF_OK;
}
gf_status gf_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!ms) return fail(GF_ERR_BAD_ARG, "ms is null");
    HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GF_OK;
}

size_t gf_lattice_bytes(const gf_lattice *lat)
{
    int64_t V;
    if (check_lattice(lat, &V) != GF_OK) return 0;
    return (size_t)V * 4 * lat->layout * (lat->dtype == GF_F64 ? 8 : 4);
}

// ------------------------------------------------------------ operators
gf_status gf_sweep(const gf_lattice *lat, void *U, void *Udw, int parity, int gauge, int algo, float param,
                   int seed, uint32_t counter, int nsb, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    if (algo < GF_OR || algo > GF_RANDOM) return fail(GF_ERR_BAD_ARG, "unknown update %d", algo);
    if (parity != 0 && parity != 1) return fail(GF_ERR_BAD_ARG, "parity must be 0 or 1, got %d", parity);
    if (gauge == GF_COULOMB) {
        if (lat->dims[0] != 1) return fail(GF_ERR_BAD_SHAPE, "Coulomb sweeps act on a timeslice: dims[0] must be 1");
        if (!Udw) return fail(GF_ERR_BAD_ARG, "Coulomb sweep needs the slice t-1 buffer Udw");
    } else if (lat->dims[0] % 2) {
        return fail(GF_ERR_BAD_SHAPE, "time extent %d is odd", lat->dims[0]);
    }
    if (gauge == GF_MAG && algo == GF_SR) { /* SR is a MAG update in the reference */ }
    if (nsb <= 0) nsb = 32;
    const Geom g = make_geom(lat->dims, V);
    hipError_t e;
    if (lat->dtype == GF_F64) {
        SweepArgs<double> a;
        a.U = (double *)U;
        a.Udw = gauge == GF_COULOMB ? (double *)Udw : (double *)U;
        a.g = g;
        a.parity = parity;
        a.param = param;
        a.seed = (uint32_t)seed;
        a.counter = counter;
        a.nsb = nsb;
        e = launch_sweep<double>(lat->layout, gauge, algo, a, S(stream));
    } else {
        SweepArgs<float> a;
        a.U = (float *)U;
        a.Udw = gauge == GF_COULOMB ? (float *)Udw : (float *)U;
        a.g = g;
        a.parity = parity;
        a.param = param;
        a.seed = (uint32_t)seed;
        a.counter = counter;
        a.nsb = nsb;
        e = launch_sweep<float>(lat->layout, gauge, algo, a, S(stream));
    }
    if (e != hipSuccess) return hip_fail(e, "sweep kernel launch");
    return GF_OK;
}

gf_status gf_project(const gf_lattice *lat, void *U, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    hipError_t e = lat->dtype == GF_F64 ? launch_project<double>(lat->layout, (double *)U, V, S(stream))
                                        : launch_project<float>(lat->layout, (float *)U, V, S(stream));
    if (e != hipSuccess) return hip_fail(e, "project kernel launch");
    return GF_OK;
}

gf_status gf_set_hot(const gf_lattice *lat, void *U, int seed, uint32_t counter, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    if (lat->layout != GF_LAYOUT_PATTERN18) return fail(GF_ERR_BAD_ARG, "set_hot writes the 18-real layout");
    hipError_t e = lat->dtype == GF_F64 ? launch_set_hot<double>((double *)U, V, (uint32_t)seed, counter, S(stream))
                                        : launch_set_hot<float>((float *)U, V, (uint32_t)seed, counter, S(stream));
    if (e != hipSuccess) return hip_fail(e, "set_hot kernel launch");
    return GF_OK;
}

gf_status gf_convert(const gf_lattice *src_lat, const void *src, const gf_lattice *dst_lat, void *dst,
                     int reunitarize, void *stream)
{
    int64_t V, V2;
    GF_TRY(check_lattice(src_lat, &V));
    GF_TRY(check_lattice(dst_lat, &V2));
    for (int i = 0; i < 4; i++)
        if (src_lat->dims[i] != dst_lat->dims[i]) return fail(GF_ERR_BAD_SHAPE, "source and destination shapes differ");
    if (!src || !dst) return fail(GF_ERR_BAD_ARG, "null buffer");
    if (src == dst) return fail(GF_ERR_BAD_ARG, "convert needs distinct buffers");
    hipError_t e = launch_convert(src_lat->dtype, src_lat->layout, src, dst_lat->dtype, dst_lat->layout, dst, V,
                                  reunitarize, S(stream));
    if (e != hipSuccess) return hip_fail(e, "convert kernel launch");
    return GF_OK;
}

// -------------------------------------------------------------- stats
gf_status gf_stats_create(const int dims[4], int gauge, int crit, gf_stats **out)
{
    if (!out) return fail(GF_ERR_BAD_ARG, "out is null");
    *out = nullptr;
    int64_t V;
    GF_TRY(check_dims(dims, &V));
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    if (crit != GF_CRIT_MAX && crit != GF_CRIT_AVERAGE) return fail(GF_ERR_BAD_ARG, "unknown stopping criterion %d", crit);
    gf_stats *st = new gf_stats();
    memcpy(st->dims, dims, sizeof st->dims);
    st->gauge = gauge;
    st->crit = crit;
    st->V = V;
    st->nblocks = quality_blocks(V);
    hipError_t e = hipMalloc((void **)&st->partial, sizeof(double) * 2 * st->nblocks);
    if (e == hipSuccess) e = hipMalloc((void **)&st->out, sizeof(double) * 2);
    if (e == hipSuccess) e = hipHostMalloc((void **)&st->host, sizeof(double) * 2, hipHostMallocDefault);
    if (e != hipSuccess) {
        if (st->partial) (void)hipFree(st->partial);
        if (st->out) (void)hipFree(st->out);
        delete st;
        return hip_fail(e, "stats scratch allocation");
    }
    *out = st;
    return GF_OK;
}

gf_status gf_stats_destroy(gf_stats *st)
{
    if (!st) return GF_OK;
    hipError_t e1 = hipFree(st->partial), e2 = hipFree(st->out), e3 = hipHostFree(st->host);
    delete st;
    if (e1 != hipSuccess) return hip_fail(e1, "hipFree");
    if (e2 != hipSuccess) return hip_fail(e2, "hipFree");
    if (e3 != hipSuccess) return hip_fail(e3, "hipHostFree");
    return GF_OK;
}

static gf_status quality_launch(const gf_lattice *lat, int gauge, const void *U, double *dGff, double *dA,
                                double *partial, int crit_max, hipStream_t s)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    if (gauge == GF_COULOMB && lat->dims[0] != 1)
        return fail(GF_ERR_BAD_SHAPE, "Coulomb quality acts on a timeslice: dims[0] must be 1");
    const Geom g = make_geom(lat->dims, V);
    hipError_t e;
    if (lat->dtype == GF_F64) {
        QualityArgs<double> a{(const double *)U, g, dGff, dA, partial, crit_max};
        e = launch_quality<double>(lat->layout, gauge, a, s);
    } else {
        QualityArgs<float> a{(const float *)U, g, dGff, dA, partial, crit_max};
        e = launch_quality<float>(lat->layout, gauge, a, s);
    }
    if (e != hipSuccess) return hip_fail(e, "quality kernel launch");
    return GF_OK;
}

gf_status gf_stats_quality(gf_stats *st, const gf_lattice *lat, const void *U, double *gff, double *theta,
                           void *stream)
{
    if (!st) return fail(GF_ERR_BAD_ARG, "stats handle is null");
    if (!lat) return fail(GF_ERR_BAD_ARG, "lattice descriptor is null");
    for (int i = 0; i < 4; i++)
        if (lat->dims[i] != st->dims[i]) return fail(GF_ERR_BAD_SHAPE, "lattice shape differs from the stats shape");
    const int crit_max = st->crit == GF_CRIT_MAX;
    GF_TRY(quality_launch(lat, st->gauge, U, nullptr, nullptr, st->partial, crit_max, S(stream)));
    hipError_t e = launch_reduce_final(st->partial, st->nblocks, crit_max, st->out, S(stream));
    if (e != hipSuccess) return hip_fail(e, "reduction kernel launch");
    HIP_TRY(hipMemcpyAsync(st->host, st->out, 2 * sizeof(double), hipMemcpyDeviceToHost, S(stream)));
    HIP_TRY(hipStreamSynchronize(S(stream)));
    // GaugeFixingStats.hxx:203-206 with the per-gauge prefactors
    const double Nc = 3.0, Ndim = 4.0;
    const double pg = st->gauge == GF_COULOMB ? 1. / (Nc * (Ndim - 1)) : 1. / (Nc * Ndim);
    double a = st->host[0] * (1. / Nc);
    if (st->crit == GF_CRIT_AVERAGE) a = a / (double)st->V;
    if (gff) *gff = st->host[1] * pg / (double)st->V;
    if (theta) *theta = a;
    return GF_OK;
}

gf_status gf_quality_per_site(const gf_lattice *lat, int gauge, const void *U, double *dGff, double *dA, void *stream)
{
    if (!dGff || !dA) return fail(GF_ERR_BAD_ARG, "per-site output is null");
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    return quality_launch(lat, gauge, U, dGff, dA, nullptr, 0, S(stream));
}

// ------------------------------------------------------------ drivers
namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// OR phase with the reference loop semantics; `sweeps_done` counts sweeps.
gf_status or_phase(const gf_lattice *lat, void *U, int gauge, float omega, float precision, int maxiter,
                   int checkprecision, int reproject, gf_stats *st, hipStream_t s, gf_run_result *r)
{
    for (int i = 0; i < maxiter; i++) {
        GF_TRY(gf_sweep(lat, U, U, 0, gauge, GF_OR, omega, 0, 0, 32, s));
        GF_TRY(gf_sweep(lat, U, U, 1, gauge, GF_OR, omega, 0, 0, 32, s));
        r->sweeps++;
        if (reproject > 0 && i % reproject == 0) GF_TRY(gf_project(lat, U, s));
        if (checkprecision > 0 && i % checkprecision == 0) {
            GF_TRY(gf_stats_quality(st, lat, U, &r->gff, &r->theta, s));
            r->checks++;
            if (r->theta < precision) {
                r->converged = 1;
                break;
            }
        }
    }
    return GF_OK;
}

}  // namespace

gf_status gf_run_or(const gf_lattice *lat, void *U, int gauge, float omega, float precision, int maxiter,
                    int checkprecision, int reproject, gf_run_result *out, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !out) return fail(GF_ERR_BAD_ARG, "null argument");
    if (gauge == GF_COULOMB) return fail(GF_ERR_BAD_ARG, "gf_run_or drives 4D gauges; sweep Coulomb slices with gf_sweep");
    memset(out, 0, sizeof *out);
    gf_stats *st = nullptr;
    GF_TRY(gf_stats_create(lat->dims, gauge, GF_CRIT_AVERAGE, &st));
    EventPair ev;
    hipStream_t s = S(stream);
    gf_status rc = GF_OK;
    if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) rc = fail(GF_ERR_HIP, "event creation");
    if (rc == GF_OK && hipEventRecord(ev.a, s) != hipSuccess) rc = fail(GF_ERR_HIP, "event record");
    if (rc == GF_OK) rc = or_phase(lat, U, gauge, omega, precision, maxiter, checkprecision, reproject, st, s, out);
    if (rc == GF_OK && hipEventRecord(ev.b, s) != hipSuccess) rc = fail(GF_ERR_HIP, "event record");
    if (rc == GF_OK) {
        float ms = 0;
        if (hipEventSynchronize(ev.b) != hipSuccess || hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess)
            rc = fail(GF_ERR_HIP, "event timing");
        out->ms_total = ms;
    }
    gf_stats_destroy(st);
    return rc;
}

gf_status gf_run_mixed(const gf_lattice *lat, void *U, int gauge, float omega, float switch_precision, float precision,
                       int maxiter, int checkprecision, int reproject, gf_run_result *out, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !out) return fail(GF_ERR_BAD_ARG, "null argument");
    if (lat->dtype != GF_F64) return fail(GF_ERR_BAD_ARG, "the mixed driver takes an FP64 configuration");
    if (gauge == GF_COULOMB) return fail(GF_ERR_BAD_ARG, "gf_run_mixed drives 4D gauges");
    memset(out, 0, sizeof *out);
    gf_lattice l32 = *lat;
    l32.dtype = GF_F32;
    l32.layout = GF_LAYOUT_R12;
    void *U32 = nullptr;
    HIP_TRY(hipMalloc(&U32, gf_lattice_bytes(&l32)));
    gf_stats *st = nullptr;
    gf_status rc = gf_stats_create(lat->dims, gauge, GF_CRIT_AVERAGE, &st);
    EventPair ev;
    hipEvent_t mid = nullptr;
    hipStream_t s = S(stream);
    if (rc == GF_OK && (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess ||
                        hipEventCreate(&mid) != hipSuccess))
        rc = fail(GF_ERR_HIP, "event creation");
    if (rc == GF_OK && hipEventRecord(ev.a, s) != hipSuccess) rc = fail(GF_ERR_HIP, "event record");
    if (rc == GF_OK) rc = gf_convert(lat, U, &l32, U32, 0, s);
    if (rc == GF_OK)
        rc = or_phase(&l32, U32, gauge, omega, switch_precision, maxiter, checkprecision, reproject, st, s, out);
    out->sweeps_fp32 = out->sweeps;
    // back to FP64 with a reunitarization in FP64
    if (rc == GF_OK) rc = gf_convert(&l32, U32, lat, U, 1, s);
    if (rc == GF_OK && hipEventRecord(mid, s) != hipSuccess) rc = fail(GF_ERR_HIP, "event record");
    if (rc == GF_OK) {
        out->converged = 0;
        const int left = maxiter - out->sweeps;
        rc = or_phase(lat, U, gauge, omega, precision, left, checkprecision, reproject, st, s, out);
    }
    if (rc == GF_OK && hipEventRecord(ev.b, s) != hipSuccess) rc = fail(GF_ERR_HIP, "event record");
    if (rc == GF_OK) {
        float ms = 0, ms1 = 0;
        if (hipEventSynchronize(ev.b) != hipSuccess || hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess ||
            hipEventElapsedTime(&ms1, ev.a, mid) != hipSuccess)
            rc = fail(GF_ERR_HIP, "event timing");
        out->ms_total = ms;
        out->ms_fp32 = ms1;
    }
    if (mid) (void)hipEventDestroy(mid);
    gf_stats_destroy(st);
    (void)hipFree(U32);
    return rc;
}

}  // extern "C"
