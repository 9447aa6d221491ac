// su3gf_api.hip -- implementation of the C ABI declared in include/su3gf.h.
//
// Host-side validation, dispatch to the gfx950 kernels, the quality
// reduction bookkeeping of GaugeFixingStats and the OR / mixed-precision
// drivers.  Every HIP call is checked; failures come back as gf_status with
// a message retrievable through gf_last_error().
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <map>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "su3gf_kernels.h"
#include "su3gf_error.h"

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

// the gauge-transformation sweep's 32-bit buffer offsets bound its volume; the
// link sweeps (GF_MIXED_LINKS, run_sa mode 0, gf_run_fast) have no such bound
gf_status check_gauge_sweep_volume(int64_t V)
{
    if (gauge_sweep_fits(V)) return GF_OK;
    return fail(GF_ERR_BAD_SHAPE,
                "the gauge-transformation sweep takes at most 44739200 sites (%lld given); the link sweeps "
                "(GF_MIXED_LINKS, annealing mode 0, gf_run_fast) have no such limit",
                (long long)V);
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
    g.slice_half = 0;
    g.slice_bps = 0;
    g.slice_div = make_fastdiv(1);
    return g;
}

// Coulomb on a 4D lattice is checked per timeslice: the sweep kernels then
// give each timeslice's sites of a parity blocks of their own, so that a
// block's partial sums belong to one slice whatever the slice volume
Geom per_slice(Geom g, int gauge)
{
    if (gauge == GF_COULOMB && g.N[0] > 1) {
        g.slice_half = g.Vh / g.N[0];
        g.slice_bps = (g.slice_half + 63) / 64;
        g.slice_div = make_fastdiv((uint32_t)g.slice_bps);
    }
    return g;
}

inline hipStream_t S(void *stream) { return (hipStream_t)stream; }

// Driver scratch memory, stream-ordered: allocated and freed on the driver's
// stream (hipMallocAsync / hipFreeAsync), so a driver neither waits for nor
// stalls work that other streams run meanwhile, and repeated runs reuse the
// pool instead of the device allocator.
struct Scratch {
    void *p = nullptr;
    hipStream_t s = nullptr;
    Scratch() = default;
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    ~Scratch()
    {
        if (p) (void)hipFreeAsync(p, s);
    }
    hipError_t alloc(size_t bytes, hipStream_t stream)
    {
        s = stream;
        return hipMallocAsync(&p, bytes, stream);
    }
    template <typename T> T *as() const { return (T *)p; }
};

// The synchronisation words of a lattice's pipelined sweeps
// (launch_gsweep_pipe), allocated and zeroed only where that sweep applies.
// check() reports a dependency wait that ran out of its spin limit.
struct PipeBuf {
    Scratch mem;
    uint32_t count = 0;  // launches on these words (block-index order)
    hipError_t init(const Geom &geo, int gauge, bool links64, hipStream_t s)
    {
        if (!gsweep_pipe_applies(geo, gauge, links64)) return hipSuccess;
        const size_t bytes = sizeof(uint32_t) * pipe_sync_words(sweep_grid(geo));
        const hipError_t e = mem.alloc(bytes, s);
        return e != hipSuccess ? e : hipMemsetAsync(mem.p, 0, bytes, s);
    }
    PipeSync *get() const { return mem.as<PipeSync>(); }
    gf_status check(hipStream_t s) const
    {
        if (!mem.p) return GF_OK;
        uint32_t err = 0;
        if (hipMemcpyAsync(&err, &get()->err, sizeof err, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return fail(GF_ERR_HIP, "reading the pipelined sweep's state");
        if (err) return fail(GF_ERR_HIP, "a pipelined sweep's dependency wait ran out of its spin limit");
        return GF_OK;
    }
};

// Pinned host copies of the drivers' device state, one per host thread.  They
// are never freed (freeing pinned memory would synchronise the device): a
// thread that ends puts its copy on a process-wide free list, where the next
// new thread finds it, so threads that come and go (the batch calls' workers)
// reuse the same few.  Every copy into one is waited for before the driver
// returns, so a copy on the list is idle.
std::mutex g_pinned_mu;
std::vector<DevState *> g_pinned_free;

struct PinnedSlot {
    DevState *h = nullptr;
    ~PinnedSlot()
    {
        if (!h) return;
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        g_pinned_free.push_back(h);
    }
};

hipError_t pinned_state(DevState **out)
{
    static thread_local PinnedSlot slot;
    if (!slot.h) {
        {
            std::lock_guard<std::mutex> lock(g_pinned_mu);
            if (!g_pinned_free.empty()) {
                slot.h = g_pinned_free.back();
                g_pinned_free.pop_back();
            }
        }
        if (!slot.h) {
            DevState *h = nullptr;
            const hipError_t e = hipHostMalloc((void **)&h, sizeof(DevState), hipHostMallocDefault);
            if (e != hipSuccess) return e;
            slot.h = h;
        }
    }
    *out = slot.h;
    return hipSuccess;
}

}  // namespace

// the error state of the C ABI, for the host-only translation units
gf_status su3gf::set_error(gf_status st, const char *msg)
{
    g_err = msg;
    return st;
}

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
    // a device-to-device hipMemcpy may return before the copy is done; wait for
    // it so that work on any stream sees the copied data
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    HIP_TRY(hipDeviceSynchronize());
    return GF_OK;
}
gf_status gf_memcpy_async(void *dst, const void *src, size_t bytes, void *stream)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, S(stream)));
    return GF_OK;
}
gf_status gf_memset(void *ptr, int value, size_t bytes)
{
    HIP_TRY(hipMemset(ptr, value, bytes));
    HIP_TRY(hipDeviceSynchronize());
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
    const bool dp_update = (algo & GF_UPDATE_DP) != 0;
    algo &= ~GF_UPDATE_DP;
    if (algo < GF_OR || algo > GF_RANDOM) return fail(GF_ERR_BAD_ARG, "unknown update %d", algo);
    if (dp_update && algo != GF_OR && algo != GF_MICRO && algo != GF_SA)
        return fail(GF_ERR_BAD_ARG, "double-precision updates exist for OR, microcanonical and heatbath steps");
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
        SweepArgs<double> a{};
        a.U = (double *)U;
        a.Udw = gauge == GF_COULOMB ? (double *)Udw : (double *)U;
        a.g = g;
        a.parity = parity;
        a.param = param;
        a.seed = (uint32_t)seed;
        a.counter = counter;
        a.nsb = nsb;
        e = exact::launch_sweep<double>(lat->layout, gauge, algo, a, S(stream));
    } else {
        SweepArgs<float> a{};
        a.U = (float *)U;
        a.Udw = gauge == GF_COULOMB ? (float *)Udw : (float *)U;
        a.g = g;
        a.parity = parity;
        a.param = param;
        a.seed = (uint32_t)seed;
        a.counter = counter;
        a.nsb = nsb;
        e = exact::launch_sweep<float>(lat->layout, gauge, dp_update ? algo | kDpUpdate : algo, a, S(stream));
    }
    if (e != hipSuccess) return hip_fail(e, "sweep kernel launch");
    return GF_OK;
}

gf_status gf_project(const gf_lattice *lat, void *U, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    hipError_t e = lat->dtype == GF_F64 ? exact::launch_project<double>(lat->layout, 4, (double *)U, V, S(stream))
                                        : exact::launch_project<float>(lat->layout, 4, (float *)U, V, S(stream));
    if (e != hipSuccess) return hip_fail(e, "project kernel launch");
    return GF_OK;
}

gf_status gf_set_hot(const gf_lattice *lat, void *U, int seed, uint32_t counter, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    if (lat->layout != GF_LAYOUT_PATTERN18) return fail(GF_ERR_BAD_ARG, "set_hot writes the 18-real layout");
    hipError_t e = lat->dtype == GF_F64 ? exact::launch_set_hot<double>((double *)U, V, (uint32_t)seed, counter, S(stream))
                                        : exact::launch_set_hot<float>((float *)U, V, (uint32_t)seed, counter, S(stream));
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

size_t gf_links24_words(const int dims[4])
{
    int64_t V = 1;
    for (int i = 0; i < 4; i++) V *= dims[i] > 0 ? dims[i] : 0;
    return (size_t)links24_words(V);
}

gf_status gf_pack_links24(const gf_lattice *lat, const void *U, uint32_t *P, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !P) return fail(GF_ERR_BAD_ARG, "null buffer");
    if (lat->dtype == GF_F32 && lat->layout != GF_LAYOUT_R12)
        return fail(GF_ERR_BAD_ARG, "FP32 links are packed from the compact 12-real layout");
    hipError_t e = launch_pack24(lat->dtype, lat->layout, U, P, V, S(stream));
    if (e != hipSuccess) return hip_fail(e, "pack kernel launch");
    return GF_OK;
}

gf_status gf_unpack_links24(const int dims[4], const uint32_t *P, float *U32, void *stream)
{
    int64_t V;
    GF_TRY(check_dims(dims, &V));
    if (!P || !U32) return fail(GF_ERR_BAD_ARG, "null buffer");
    hipError_t e = launch_unpack24(P, U32, V, S(stream));
    if (e != hipSuccess) return hip_fail(e, "unpack kernel launch");
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
    st->nblocks = exact::quality_blocks(V);
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

namespace {
// the quality kernel's XCD block order for this lattice (QualityArgs::tpass),
// tuned on first use (defined with the sweeps' tuner below)
int quality_tpass(const gf_lattice *lat, int gauge, const void *U, hipStream_t s);

hipError_t quality_kernel(const gf_lattice *lat, int gauge, const void *U, double *dGff, double *dA,
                          double *partial, int crit_max, int tpass, hipStream_t s)
{
    int64_t V = 1;
    for (int d = 0; d < 4; d++) V *= lat->dims[d];
    // Coulomb on a 4D lattice: the spatial quality of every timeslice
    const Geom g = make_geom(lat->dims, V);
    if (lat->dtype == GF_F64) {
        QualityArgs<double> a{};
        a.U = (const double *)U; a.g = g; a.dGff = dGff; a.dA = dA; a.partial = partial; a.crit_max = crit_max;
        a.tpass = tpass;
        return exact::launch_quality<double>(lat->layout, gauge, a, s);
    }
    QualityArgs<float> a{};
    a.U = (const float *)U; a.g = g; a.dGff = dGff; a.dA = dA; a.partial = partial; a.crit_max = crit_max;
    a.tpass = tpass;
    return exact::launch_quality<float>(lat->layout, gauge, a, s);
}
}  // namespace

static gf_status quality_launch(const gf_lattice *lat, int gauge, const void *U, double *dGff, double *dA,
                                double *partial, int crit_max, hipStream_t s)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U) return fail(GF_ERR_BAD_ARG, "U is null");
    const hipError_t e = quality_kernel(lat, gauge, U, dGff, dA, partial, crit_max, quality_tpass(lat, gauge, U, s), s);
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
    // reconstruct the third rows (LandauGaugeFixingSU3_4D.cu:271)
    if (rc == GF_OK) rc = gf_project(lat, U, s);
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

}  // extern "C"

// ------------------------------------------- wavefront slab autotuner
// The gauge-transformation sweep can run a whole sweep as a wavefront over
// timeslices (launch_gsweep_wavefront), so that a timeslice's links come from
// HBM once per sweep instead of once per half-sweep; it pays only when the
// lattice does not fit the Infinity Cache, and the best slab depends on the
// lattice.  The first sweep of a (shape, gauge, link form, arithmetic) in the
// process times a few OR sweeps with each candidate slab -- and with the two
// kernels -- on a scratch copy of g and keeps the fastest; every choice gives
// bit-identical results, so only the time depends on it.
namespace {
std::mutex g_slab_mu;
// per (shape, gauge, link form, arithmetic): the wavefront slab of sweeps
// without a check (0: two kernels) and the XCD block order (tpass: passes of
// that many timeslices, 0: one pass)
struct GsweepTuning {
    int slab = 0;
    int tpass = 0;
    int pipe_slack = 0;  // the pipelined sweep's extra lag (slab == kPipelinedSlab)
};
std::map<std::array<int, 8>, GsweepTuning> g_slab_cache;
int g_slab_setting = -1;   // -1 tuned, 0 never, n > 0 fixed (gf_set_wavefront)
int g_tpass_setting = -1;  // -1 tuned, 0 one pass, n > 0 passes of n timeslices (gf_set_block_order)
int g_lag_setting = -1;    // -1 tuned, n >= 0 the pipelined sweep's extra lag (gf_set_pipelined_lag)
// extra lags the tuner tries for the pipelined sweep
constexpr int kPipeLags[] = {192, 256, 384, 512};

// links: 0 FP32, 1 24-bit, 2 FP64 (with an FP64 g)
std::array<int, 8> slab_key(const Geom &geo, int gauge, int links, bool fp64)
{
    return {geo.N[0], geo.N[1], geo.N[2], geo.N[3], gauge, links, fp64 ? 1 : 0, geo.slice_bps ? 1 : 0};
}

// timeslices each XCD walks in the launcher's timeslice order (launch_gsweep's
// remap = 2), or 0 where that order does not apply
int xcd_slices(const Geom &geo)
{
    const int64_t half_slice = geo.Vh / geo.N[0];
    return geo.N[0] % 8 == 0 && half_slice % 64 == 0 ? geo.N[0] / 8 : 0;
}

// average ms of one sweep over `n` sweeps (after one unmeasured), slab 0 = two kernels
float time_gsweep(const GaugeSweepArgs &t, int gauge, bool fp64, int slab, int n, hipStream_t s, hipEvent_t e0,
                  hipEvent_t e1)
{
    auto sweep = [&]() -> hipError_t {
        if (slab > 0) return launch_gsweep_wavefront(gauge, GF_OR, fp64, t, slab, s);
        for (int par = 0; par < 2; par++) {
            GaugeSweepArgs b = t;
            b.parity = par;
            const hipError_t e = launch_gsweep(gauge, GF_OR, 0, fp64, b, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    if (sweep() != hipSuccess) return -1.f;
    if (hipEventRecord(e0, s) != hipSuccess) return -1.f;
    for (int k = 0; k < n; k++)
        if (sweep() != hipSuccess) return -1.f;
    float ms = 0;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        return -1.f;
    return ms / n;
}

// SU3GF_TUNE_LOG=1 in the environment: the tuner prints every candidate's
// time per sweep and its choice to stderr
bool tune_log()
{
    static const bool on = [] {
        const char *e = std::getenv("SU3GF_TUNE_LOG");
        return e && *e && std::strcmp(e, "0") != 0;
    }();
    return on;
}

// times OR sweeps of each candidate on a scratch copy of g: first the
// block order of the two-kernel sweep (passes of each divisor of the XCD's
// timeslices), then, with that order, the wavefront slabs, then the
// pipelined sweep's lags; a candidate must beat the incumbent by 3% (launch
// count, noise), and a choice other than the default must beat it again in
// alternate timings.  Never fails (the defaults on any error).
GsweepTuning tune_gsweep(const GaugeSweepArgs &ga, int gauge, bool fp64, hipStream_t s)
{
    const Geom &geo = ga.geo;
    GsweepTuning tu;
    const bool log = tune_log();
    if (log)
        std::fprintf(stderr, "su3gf tune: %dx%dx%dx%d gauge %d links %s, %s arithmetic\n", geo.N[0], geo.N[1],
                     geo.N[2], geo.N[3], gauge, ga.links64 ? "fp64" : ga.links24 ? "24-bit" : "fp32",
                     fp64 ? "fp64" : "fp32");
    Scratch g;
    EventPair ev;
    const size_t gbytes = (ga.links64 ? sizeof(double) : sizeof(float)) * gauge_floats(geo.V);
    if (g.alloc(gbytes, s) == hipSuccess && hipMemcpyAsync(g.p, ga.g, gbytes, hipMemcpyDeviceToDevice, s) == hipSuccess &&
        hipEventCreate(&ev.a) == hipSuccess && hipEventCreate(&ev.b) == hipSuccess) {
        GaugeSweepArgs t = ga;
        t.g = g.as<float>();
        t.state = nullptr;
        t.partial = nullptr;
        t.param = 1.7f;
        const int base_tpass = g_tpass_setting >= 0 ? g_tpass_setting : 0;
        t.tpass = base_tpass;
        // about 20 ms of sweeps first, dropped: right after an allocation the
        // clocks and the Infinity Cache are not yet at a sustained run's
        // steady state, which would favour whatever is timed later; then
        // each candidate over about 8 ms of sweeps (at least 3)
        const float est = time_gsweep(t, gauge, fp64, 0, 3, s, ev.a, ev.b);
        if (est > 0) (void)time_gsweep(t, gauge, fp64, 0, std::min(400, std::max(3, (int)(20.f / est))), s, ev.a, ev.b);
        const int n = est > 0 ? std::min(100, std::max(3, (int)(8.f / est))) : 3;
        const float base = time_gsweep(t, gauge, fp64, 0, n, s, ev.a, ev.b);
        if (log) std::fprintf(stderr, "su3gf tune:   two kernels, one pass: %.4f ms per sweep (%d sweeps)\n", base, n);
        float two = base;
        const int spx = xcd_slices(geo);
        if (g_tpass_setting < 0 && spx > 1 && base > 0) {
            for (int P = 1; P < spx; P++) {
                if (spx % P != 0) continue;
                t.tpass = P;
                const float ms = time_gsweep(t, gauge, fp64, 0, n, s, ev.a, ev.b);
                if (log) std::fprintf(stderr, "su3gf tune:   two kernels, passes of %d timeslices: %.4f\n", P, ms);
                if (ms > 0 && ms < two && ms < 0.97f * base) {
                    two = ms;
                    tu.tpass = P;
                }
            }
            t.tpass = tu.tpass;
        }
        float best_ms = two;
        for (int c : {2, 3, 4, 6, 8, 12}) {
            if (c >= geo.N[0] || base <= 0) break;
            const float ms = time_gsweep(t, gauge, fp64, c, n, s, ev.a, ev.b);
            if (log) std::fprintf(stderr, "su3gf tune:   wavefront, slabs of %d: %.4f\n", c, ms);
            if (ms > 0 && ms < best_ms && ms < 0.97f * two) {
                best_ms = ms;
                tu.slab = c;
            }
        }
        // the pipelined sweep, on synchronisation words of its own, for each extra lag
        PipeBuf pb;
        if (base > 0 && pb.init(geo, gauge, ga.links64 != 0, s) == hipSuccess && pb.get()) {
            t.pipe = pb.get();
            t.pipe_count = &pb.count;
            for (int lag : kPipeLags) {
                if (g_lag_setting >= 0 && lag != kPipeLags[0]) break;
                t.pipe_slack = g_lag_setting >= 0 ? g_lag_setting : lag;
                const float ms = time_gsweep(t, gauge, fp64, kPipelinedSlab, n, s, ev.a, ev.b);
                if (log) std::fprintf(stderr, "su3gf tune:   pipelined, lag %d: %.4f\n", t.pipe_slack, ms);
                if (ms > 0 && ms < best_ms && ms < 0.97f * two && pb.check(s) == GF_OK) {
                    best_ms = ms;
                    tu.slab = kPipelinedSlab;
                    tu.pipe_slack = t.pipe_slack;
                }
            }
        }
        // confirm a choice other than the default (two kernels, the base
        // block order) by timing the two alternately, twice each: it must
        // still win by 3%, or the default stays
        if ((tu.slab != 0 || tu.tpass != base_tpass) && base > 0) {
            GaugeSweepArgs d = t, w = t;
            d.tpass = base_tpass;
            w.tpass = tu.tpass;
            w.pipe_slack = tu.pipe_slack;
            float d_ms = 0, w_ms = 0;
            for (int r = 0; r < 2; r++) {
                const float a1 = time_gsweep(d, gauge, fp64, 0, n, s, ev.a, ev.b);
                const float b1 = time_gsweep(w, gauge, fp64, tu.slab, n, s, ev.a, ev.b);
                if (a1 <= 0 || b1 <= 0) d_ms = -1e30f;
                d_ms += a1;
                w_ms += b1;
            }
            const bool keep = d_ms > 0 && w_ms < 0.97f * d_ms && (tu.slab != kPipelinedSlab || pb.check(s) == GF_OK);
            if (log)
                std::fprintf(stderr, "su3gf tune:   confirmation, alternately: default %.4f, choice %.4f: %s\n",
                             d_ms / 2, w_ms / 2, keep ? "kept" : "default kept");
            if (!keep) tu = GsweepTuning{};
        }
        t.pipe = nullptr;
    }
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    if (log) {
        char path[64];
        if (tu.slab == kPipelinedSlab) std::snprintf(path, sizeof path, "pipelined, lag %d", tu.pipe_slack);
        else if (tu.slab > 0) std::snprintf(path, sizeof path, "wavefront, slabs of %d", tu.slab);
        else std::snprintf(path, sizeof path, "two kernels");
        std::fprintf(stderr, "su3gf tune:   chosen: %s, block order %d\n", path, tu.tpass);
    }
    return tu;
}

// the slab for sweeps without a check of the gauge-transformation sweep
// described by ga (0: the two-kernel sweep), and its block order, set in
// ga.tpass; tuned on the first call for a kind of lattice unless both are set
// (gf_set_wavefront, gf_set_block_order).  Never fails.
int wavefront_slab(GaugeSweepArgs &ga, int gauge, bool fp64, hipStream_t s)
{
    const Geom &geo = ga.geo;
    std::lock_guard<std::mutex> lock(g_slab_mu);
    GsweepTuning tu;
    const auto key = slab_key(geo, gauge, ga.links64 ? 2 : ga.links24, fp64);
    const auto hit = g_slab_cache.find(key);
    if (hit != g_slab_cache.end()) {
        tu = hit->second;
    } else if (geo.N[0] >= 4 && (g_slab_setting < 0 || g_tpass_setting < 0)) {
        tu = tune_gsweep(ga, gauge, fp64, s);
        g_slab_cache[key] = tu;
    }
    ga.tpass = g_tpass_setting >= 0 ? g_tpass_setting : tu.tpass;
    ga.pipe_slack = g_lag_setting >= 0 ? g_lag_setting : tu.pipe_slack;
    return g_slab_setting >= 0 ? g_slab_setting : tu.slab;
}

// The drop-in quality kernel walks each XCD's timeslices in the same kind of
// passes (exact::launch_quality); it reads the links only, so candidates are
// timed on the caller's buffer with no outputs (three launches each after one
// unmeasured).  Same setting (gf_set_block_order) and 3% margin as the sweeps.
std::map<std::array<int, 7>, int> g_quality_cache;

int quality_tpass(const gf_lattice *lat, int gauge, const void *U, hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_slab_mu);
    if (g_tpass_setting >= 0) return g_tpass_setting;
    const std::array<int, 7> key{lat->dims[0], lat->dims[1], lat->dims[2], lat->dims[3], lat->dtype, lat->layout,
                                 gauge};
    const auto hit = g_quality_cache.find(key);
    if (hit != g_quality_cache.end()) return hit->second;
    int64_t V = 1;
    for (int d = 0; d < 4; d++) V *= lat->dims[d];
    const int64_t half_slice = V / 2 / lat->dims[0];
    const int spx = lat->dims[0] % 8 == 0 && half_slice % 128 == 0 && exact::quality_blocks(V) % 8 == 0
                        ? lat->dims[0] / 8 : 0;
    int best = 0;
    EventPair ev;
    if (spx > 1 && hipEventCreate(&ev.a) == hipSuccess && hipEventCreate(&ev.b) == hipSuccess) {
        auto time = [&](int P) -> float {
            if (quality_kernel(lat, gauge, U, nullptr, nullptr, nullptr, 0, P, s) != hipSuccess) return -1.f;
            if (hipEventRecord(ev.a, s) != hipSuccess) return -1.f;
            for (int k = 0; k < 3; k++)
                if (quality_kernel(lat, gauge, U, nullptr, nullptr, nullptr, 0, P, s) != hipSuccess) return -1.f;
            float ms = 0;
            if (hipEventRecord(ev.b, s) != hipSuccess || hipEventSynchronize(ev.b) != hipSuccess ||
                hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess)
                return -1.f;
            return ms;
        };
        const float one_pass = time(0);
        float best_ms = one_pass;
        for (int P = 1; P < spx && one_pass > 0; P++) {
            if (spx % P != 0) continue;
            const float ms = time(P);
            if (ms > 0 && ms < best_ms && ms < 0.97f * one_pass) {
                best_ms = ms;
                best = P;
            }
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    g_quality_cache[key] = best;
    return best;
}
}  // namespace

extern "C" gf_status gf_set_wavefront(int slab, int *previous)
{
    if (slab < -1) return fail(GF_ERR_BAD_ARG, "wavefront setting %d (-1 tuned, 0 never, n > 0 fixed)", slab);
    std::lock_guard<std::mutex> lock(g_slab_mu);
    if (previous) *previous = g_slab_setting;
    g_slab_setting = slab;
    return GF_OK;
}

extern "C" gf_status gf_set_block_order(int tpass, int *previous)
{
    if (tpass < -1) return fail(GF_ERR_BAD_ARG, "block order setting %d (-1 tuned, 0 one pass, n > 0 fixed)", tpass);
    std::lock_guard<std::mutex> lock(g_slab_mu);
    if (previous) *previous = g_tpass_setting;
    g_tpass_setting = tpass;
    return GF_OK;
}

namespace {
gf_status tuning_of(const gf_lattice *lat, int gauge, int mode, GsweepTuning *tu, bool *known)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (mode < GF_MIXED_GAUGE32 || mode > GF_MIXED_GAUGE24) return fail(GF_ERR_BAD_ARG, "not a gauge-sweep mode: %d", mode);
    const Geom geo = per_slice(make_geom(lat->dims, V), gauge);
    const auto hit = g_slab_cache.find(slab_key(geo, gauge, mode == GF_MIXED_GAUGE24, mode == GF_MIXED_GAUGE64));
    *known = hit != g_slab_cache.end();
    if (*known) *tu = hit->second;
    return GF_OK;
}
}  // namespace

extern "C" gf_status gf_wavefront_choice(const gf_lattice *lat, int gauge, int mode, int *slab)
{
    if (!slab) return fail(GF_ERR_BAD_ARG, "slab is null");
    std::lock_guard<std::mutex> lock(g_slab_mu);
    GsweepTuning tu;
    bool known = false;
    GF_TRY(tuning_of(lat, gauge, mode, &tu, &known));
    *slab = g_slab_setting >= 0 ? g_slab_setting : known ? tu.slab : -1;
    return GF_OK;
}

extern "C" gf_status gf_set_pipelined_lag(int lag, int *previous)
{
    if (lag < -1) return fail(GF_ERR_BAD_ARG, "pipelined lag setting %d (-1 tuned, n >= 0 fixed)", lag);
    std::lock_guard<std::mutex> lock(g_slab_mu);
    if (previous) *previous = g_lag_setting;
    g_lag_setting = lag;
    return GF_OK;
}

extern "C" gf_status gf_pipelined_launches(long long *n)
{
    if (!n) return fail(GF_ERR_BAD_ARG, "n is null");
    *n = pipe_launches();
    return GF_OK;
}

extern "C" gf_status gf_block_order_choice(const gf_lattice *lat, int gauge, int mode, int *tpass)
{
    if (!tpass) return fail(GF_ERR_BAD_ARG, "tpass is null");
    std::lock_guard<std::mutex> lock(g_slab_mu);
    GsweepTuning tu;
    bool known = false;
    GF_TRY(tuning_of(lat, gauge, mode, &tu, &known));
    *tpass = g_tpass_setting >= 0 ? g_tpass_setting : known ? tu.tpass : -1;
    return GF_OK;
}

// ------------------------------------------------------ simulated annealing
namespace {
// The reference apps' temperature schedules: heat(T) for each heatbath step,
// project() when i % reproject == 0 for temperature i; *n receives the
// number of temperatures.
template <typename Heat, typename Project>
gf_status sa_schedule(int gauge, float tmax, float tmin, int steps, int reproject, Heat &heat, Project &project,
                      int *n)
{
    float temperature = tmax;
    const float base = (tmax - tmin) / (float)steps;
    *n = 0;
    if (gauge == GF_LANDAU || gauge == GF_COULOMB) {
        // fixed number of temperatures (LandauGaugeFixingSU3_4D.cu:205-240,
        // CoulombGaugeFixingSU3_4D.cu:229-265)
        for (int i = 0; i < steps; i++, (*n)++) {
            GF_TRY(heat(temperature));
            if (reproject > 0 && i % reproject == 0) GF_TRY(project());
            temperature -= base;
        }
    } else {
        // down to tmin with steps 15 times finer in the critical window:
        // U1xU1 2 <= T <= 2.7 (U1xU1GaugeFixingSU3_4D.cu:219-251), MAG
        // 0.7 <= T <= 0.88 with five heatbath steps per temperature
        // (MAGaugeFixingSU3_4D.cu:228-265); float T against double bounds
        const double lo = gauge == GF_U1XU1 ? 2.0 : 0.7, hi = gauge == GF_U1XU1 ? 2.7 : 0.88;
        const int per_t = gauge == GF_MAG ? 5 : 1;
        int i = 0;
        do {
            const float step = (temperature <= hi && temperature >= lo) ? base / 15 : base;
            for (int j = 0; j < per_t; j++) GF_TRY(heat(temperature));
            if (reproject > 0 && i % reproject == 0) GF_TRY(project());
            temperature -= step;
            i++;
            (*n)++;
        } while (temperature >= tmin);
    }
    return GF_OK;
}

// GF_SA_GRAPH: the schedule's launches are captured on the caller's stream,
// or on a blocking stream of our own that follows the null stream, and
// replayed once on it
struct SaGraph {
    hipStream_t own = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    bool capturing = false;
    hipStream_t cs = nullptr;
    SaGraph() = default;
    SaGraph(const SaGraph &) = delete;
    SaGraph &operator=(const SaGraph &) = delete;
    ~SaGraph()
    {
        if (capturing) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(cs, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
        }
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        if (own) (void)hipStreamDestroy(own);
    }
    // starts the capture on *s (a stream of our own if *s is null), which becomes the capturing stream
    gf_status begin(hipStream_t *s)
    {
        cs = *s;
        if (!cs) {
            HIP_TRY(hipStreamCreate(&own));
            cs = own;
        }
        HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        capturing = true;
        *s = cs;
        return GF_OK;
    }
    gf_status replay()
    {
        capturing = false;
        HIP_TRY(hipStreamEndCapture(cs, &g));
        HIP_TRY(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
        HIP_TRY(hipGraphLaunch(x, cs));
        if (own) HIP_TRY(hipStreamSynchronize(own));
        return GF_OK;
    }
};
}  // namespace

extern "C" gf_status gf_run_sa(const gf_lattice *lat, void *U, void *Udw, int gauge, float tmax, float tmin, int steps,
                               int micro, int reproject, int mode, int seed, uint32_t *counter, int *temperatures,
                               void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !counter) return fail(GF_ERR_BAD_ARG, "null argument");
    if (steps < 0 || micro < 0) return fail(GF_ERR_BAD_ARG, "negative step count");
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    const bool graph = (mode & GF_SA_GRAPH) != 0;
    mode &= ~GF_SA_GRAPH;
    if (mode < 0 || mode > 2) return fail(GF_ERR_BAD_ARG, "unknown annealing mode %d", mode);
    if (mode > 0) GF_TRY(check_gauge_sweep_volume(V));
    const bool gform = mode != 0;  // the gauge transformation is annealed (modes 1, 2)
    if (gform) {
        if (lat->dtype != GF_F64) return fail(GF_ERR_BAD_ARG, "modes 1 and 2 take an FP64 configuration");
        if (gauge == GF_COULOMB && lat->dims[0] == 1)
            return fail(GF_ERR_BAD_ARG, "modes 1 and 2 anneal every timeslice of a 4D lattice at once");
    }
    hipStream_t s = S(stream);
    void *dw = gauge == GF_COULOMB ? Udw : U;
    // modes 1, 2: a read-only copy of U (FP32, or 24-bit fixed point), the gauge transformation swept
    gf_lattice l32 = *lat;
    l32.dtype = GF_F32;
    l32.layout = GF_LAYOUT_R12;
    Scratch bU32, bg;
    PipeBuf pipe;
    GaugeSweepArgs ga{};
    const Geom geo = make_geom(lat->dims, V);
    if (gform) {
        HIP_TRY(bU32.alloc(mode == 2 ? sizeof(uint32_t) * links24_words(V) : gf_lattice_bytes(&l32), s));
        HIP_TRY(bg.alloc(sizeof(float) * gauge_floats(V), s));
        HIP_TRY(launch_gauge_identity(bg.as<float>(), V, s));
        if (mode == 2) HIP_TRY(launch_pack24(lat->dtype, lat->layout, U, bU32.as<uint32_t>(), V, s));
        else HIP_TRY(launch_convert(lat->dtype, lat->layout, U, GF_F32, 12, bU32.p, V, 0, s));
        ga.links24 = mode == 2;
        ga.U0 = bU32.as<const float>();
        ga.g = bg.as<float>();
        ga.geo = geo;
        ga.remap = 1;
        ga.seed = (uint32_t)seed;
        ga.nsb = 32;
        // (a captured schedule keeps the two-kernel sweeps: the pipelined
        // sweep's launch numbers come from the host at each launch)
        if (!graph) HIP_TRY(pipe.init(geo, gauge, false, s));
        ga.pipe = pipe.get();
        ga.pipe_count = &pipe.count;
    }
    auto project = [&]() -> gf_status {
        if (gform) return GF_OK;  // the gauge-transformation sweep keeps g unitary itself
        GF_TRY(gf_project(lat, U, s));
        if (gauge == GF_COULOMB && Udw && Udw != U) GF_TRY(gf_project(lat, Udw, s));
        return GF_OK;
    };
    // one heatbath temperature: SA on both parities, then `micro`
    // microcanonical pairs (LandauGaugeFixingSU3_4D.cu:214-223)
    // modes 1, 2: a sweep as a wavefront over timeslices where that pays
    const int slab = gform ? wavefront_slab(ga, gauge, false, s) : 0;
    auto gsweep_pair = [&](int algo) -> hipError_t {
        if (slab > 0) {
            const hipError_t e = launch_gsweep_wavefront(gauge, algo, false, ga, slab, s);
            if (e != hipErrorNotSupported) return e;
        }
        for (int par = 0; par < 2; par++) {
            GaugeSweepArgs b = ga;
            b.parity = par;
            b.counter = ga.counter + (uint32_t)par;
            const hipError_t e = launch_gsweep(gauge, algo, 0, false, b, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    auto heat = [&](float T) -> gf_status {
        if (gform) {
            ga.param = T;
            ga.counter = *counter;
            HIP_TRY(gsweep_pair(GF_SA));
            *counter += 2;
            for (int m = 0; m < micro; m++) HIP_TRY(gsweep_pair(GF_MICRO));
            return GF_OK;
        }
        for (int par = 0; par < 2; par++) GF_TRY(gf_sweep(lat, U, dw, par, gauge, GF_SA, T, seed, (*counter)++, 32, s));
        for (int m = 0; m < micro; m++)
            for (int par = 0; par < 2; par++) GF_TRY(gf_sweep(lat, U, dw, par, gauge, GF_MICRO, 0.f, 0, 0, 32, s));
        return GF_OK;
    };
    SaGraph gr;
    if (graph) GF_TRY(gr.begin(&s));  // every launch below is captured
    int n = 0;
    GF_TRY(sa_schedule(gauge, tmax, tmin, steps, reproject, heat, project, &n));
    if (graph) GF_TRY(gr.replay());
    if (gform) {
        // U <- G U G^dagger in FP64 (G the FP64 projection of g), third rows rebuilt
        HIP_TRY(launch_project_gauge(bg.as<float>(), V, s));
        HIP_TRY(launch_gauge_apply((const double *)U, lat->layout, bg.as<const float>(), U, GF_F64, lat->layout, geo,
                                   s));
        GF_TRY(gf_project(lat, U, s));
        HIP_TRY(hipStreamSynchronize(s));
        GF_TRY(pipe.check(s));
    }
    if (temperatures) *temperatures = n;
    return GF_OK;
}

// ------------------------------------------- annealing of a batch in one launch
namespace {
std::mutex g_batch_mu;
int g_batch_launch_blocks = 8192;  // gf_set_batch_launch_blocks

// Modes 1, 2 of gf_run_sa on the lattices idx[0..n) (n <= kMultiMax) with
// every half-sweep of all of them in one launch (launch_gsweep_multi); the
// lattices have been checked.
gf_status run_sa_multi(const int *idx, int n, const gf_lattice *lats, void *const *U, int gauge, float tmax, float tmin,
                       int steps, int micro, int mode, bool graph, const int *seeds, uint32_t *counters,
                       int *temperatures, hipStream_t s)
{
    Scratch bU[kMultiMax], bg[kMultiMax];
    MultiSweepArgs m{};
    m.n = n;
    for (int j = 0; j < n; j++) {
        const int k = idx[j];
        const gf_lattice *lat = &lats[k];
        int64_t V;
        GF_TRY(check_lattice(lat, &V));
        gf_lattice l32 = *lat;
        l32.dtype = GF_F32;
        l32.layout = GF_LAYOUT_R12;
        HIP_TRY(bU[j].alloc(mode == 2 ? sizeof(uint32_t) * links24_words(V) : gf_lattice_bytes(&l32), s));
        HIP_TRY(bg[j].alloc(sizeof(float) * gauge_floats(V), s));
        HIP_TRY(launch_gauge_identity(bg[j].as<float>(), V, s));
        if (mode == 2) HIP_TRY(launch_pack24(lat->dtype, lat->layout, U[k], bU[j].as<uint32_t>(), V, s));
        else HIP_TRY(launch_convert(lat->dtype, lat->layout, U[k], GF_F32, 12, bU[j].p, V, 0, s));
        m.item[j] = MultiItem{bU[j].as<const float>(), bg[j].as<float>(), make_geom(lat->dims, V), (uint32_t)seeds[k],
                              counters[k]};
    }
    uint32_t add = 0;  // every lattice's counter advances by 2 per heatbath step, as in gf_run_sa
    auto pair = [&](int algo) -> hipError_t {
        for (int par = 0; par < 2; par++) {
            m.parity = par;
            const hipError_t e = launch_gsweep_multi(gauge, algo, mode == 2, m, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    auto heat = [&](float T) -> gf_status {
        m.param = T;
        m.counter_add = add;
        HIP_TRY(pair(GF_SA));
        add += 2;
        for (int k = 0; k < micro; k++) HIP_TRY(pair(GF_MICRO));
        return GF_OK;
    };
    auto project = []() -> gf_status { return GF_OK; };  // the sweep keeps g unitary itself
    SaGraph gr;
    if (graph) GF_TRY(gr.begin(&s));
    int nt = 0;
    GF_TRY(sa_schedule(gauge, tmax, tmin, steps, 0, heat, project, &nt));
    if (graph) GF_TRY(gr.replay());
    for (int j = 0; j < n; j++) {
        const int k = idx[j];
        const gf_lattice *lat = &lats[k];
        const Geom &geo = m.item[j].geo;
        HIP_TRY(launch_project_gauge(bg[j].as<float>(), geo.V, s));
        HIP_TRY(launch_gauge_apply((const double *)U[k], lat->layout, bg[j].as<const float>(), U[k], GF_F64,
                                   lat->layout, geo, s));
        GF_TRY(gf_project(lat, U[k], s));
        counters[k] += add;
        if (temperatures) temperatures[k] = nt;
    }
    HIP_TRY(hipStreamSynchronize(s));
    return GF_OK;
}
}  // namespace

extern "C" gf_status gf_set_batch_launch_blocks(int blocks, int *previous)
{
    if (blocks < 0) return fail(GF_ERR_BAD_ARG, "blocks per batch launch %d (0: no limit)", blocks);
    std::lock_guard<std::mutex> lock(g_batch_mu);
    if (previous) *previous = g_batch_launch_blocks;
    g_batch_launch_blocks = blocks;
    return GF_OK;
}

// gf_batch_run_sa with GF_SA_ONE_LAUNCH: the lattices are packed into groups
// (largest first, each into the first group with room: at most kMultiMax
// lattices and g_batch_launch_blocks blocks of a parity, unless the lattice
// alone has more) and the groups anneal one after another on s.
gf_status su3gf::batch_sa_one_launch(int n, const gf_lattice *lats, void *const *U, int gauge, float tmax, float tmin,
                                     int steps, int micro, int mode, const int *seeds, uint32_t *counters,
                                     int *temperatures, void *stream)
{
    const bool graph = (mode & GF_SA_GRAPH) != 0;
    mode &= ~(GF_SA_GRAPH | GF_SA_ONE_LAUNCH);
    if (mode != 1 && mode != 2) return fail(GF_ERR_BAD_ARG, "one-launch annealing takes modes 1 and 2, not %d", mode);
    if (gauge != GF_LANDAU && gauge != GF_COULOMB && gauge != GF_U1XU1)
        return fail(GF_ERR_BAD_ARG, "one-launch annealing: Landau, Coulomb and U1xU1 (gauge %d)", gauge);
    if (steps < 0 || micro < 0) return fail(GF_ERR_BAD_ARG, "negative step count");
    std::vector<int64_t> blocks(n);
    for (int k = 0; k < n; k++) {
        int64_t V;
        if (check_lattice(&lats[k], &V) != GF_OK || check_gauge_sweep_volume(V) != GF_OK) {
            const std::string msg = gf_last_error();
            return fail(GF_ERR_BAD_SHAPE, "batch problem %d: %s", k, msg.c_str());
        }
        if (lats[k].dtype != GF_F64) return fail(GF_ERR_BAD_ARG, "batch problem %d: modes 1 and 2 take an FP64 configuration", k);
        if (gauge == GF_COULOMB && lats[k].dims[0] == 1)
            return fail(GF_ERR_BAD_ARG, "batch problem %d: modes 1 and 2 anneal every timeslice of a 4D lattice at once", k);
        blocks[k] = sweep_grid(make_geom(lats[k].dims, V));
    }
    int cap;
    {
        std::lock_guard<std::mutex> lock(g_batch_mu);
        cap = g_batch_launch_blocks;
    }
    std::vector<int> order(n);
    for (int k = 0; k < n; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return blocks[a] > blocks[b]; });
    std::vector<std::vector<int>> groups;
    std::vector<int64_t> load;
    for (int k : order) {
        size_t gi = 0;
        while (gi < groups.size() &&
               (groups[gi].size() >= (size_t)kMultiMax || (cap > 0 && load[gi] + blocks[k] > cap)))
            gi++;
        if (gi == groups.size()) {
            groups.emplace_back();
            load.push_back(0);
        }
        groups[gi].push_back(k);
        load[gi] += blocks[k];
    }
    struct OwnStream {
        hipStream_t s = nullptr;
        ~OwnStream()
        {
            if (s) (void)hipStreamDestroy(s);
        }
    } own;
    hipStream_t s = S(stream);
    if (!s) {
        HIP_TRY(hipStreamCreateWithFlags(&own.s, hipStreamNonBlocking));
        s = own.s;
    }
    for (const auto &grp : groups) {
        std::vector<int> idx(grp.begin(), grp.end());
        std::sort(idx.begin(), idx.end());  // problem order within a launch
        const gf_status st = run_sa_multi(idx.data(), (int)idx.size(), lats, U, gauge, tmax, tmin, steps, micro, mode,
                                          graph, seeds, counters, temperatures, s);
        if (st != GF_OK) {
            const std::string msg = gf_last_error();
            return fail(st, "batch launch group of problem %d: %s", idx[0], msg.c_str());
        }
    }
    return GF_OK;
}

// ------------------------------------------------- fused-check drivers
namespace {

// bytes of a fused run's device scratch for `blocks` blocks per parity: the
// DevState (in 256 bytes), then (theta, gff) of every block of each parity
constexpr size_t fused_scratch_bytes(int64_t blocks) { return 256 + sizeof(double) * 4 * (size_t)blocks; }

struct FusedCtx {
    Scratch dev;                // DevState, then 2 * nblocks doubles for each parity
    DevState *st = nullptr;     // device
    DevState *host = nullptr;   // pinned, or own_host
    DevState own_host;          // the allocation-free driver's (pageable) copy
    double *partial = nullptr;
    double *even() { return partial; }
    double *odd() { return partial + 2 * nblocks; }
    int64_t nblocks = 0;
    gf_status init(int64_t blocks, hipStream_t s)  // partial sums for up to `blocks` blocks per parity
    {
        nblocks = blocks;
        HIP_TRY(dev.alloc(fused_scratch_bytes(nblocks), s));  // state, even, odd
        st = dev.as<DevState>();
        partial = (double *)((char *)dev.p + 256);
        HIP_TRY(pinned_state(&host));
        return GF_OK;
    }
    // the same layout in caller-provided device memory (fused_scratch_bytes(blocks)
    // bytes); the host copy of the state is this object's own
    void init_at(void *mem, int64_t blocks)
    {
        nblocks = blocks;
        st = (DevState *)mem;
        partial = (double *)((char *)mem + 256);
        host = &own_host;
    }
    gf_status read(hipStream_t s)
    {
        HIP_TRY(hipMemcpyAsync(host, st, sizeof(DevState), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return GF_OK;
    }
};

constexpr int kChunk = 32;  // sweeps enqueued between two host reads of the device state

double gff_prefactor(int gauge) { return gauge == GF_COULOMB ? 1. / 9. : 1. / 12.; }

// theta below which a single-precision phase is tested for stagnation: the
// rounding floor of FP32 sweeps lies near 1e-12 for the squared theta of
// Landau, Coulomb and U1xU1, and higher for MAG's theta, a sum of |X|
double stall_floor(int gauge) { return gauge == GF_MAG ? 1e-5 : 1e-9; }
// how stagnation is judged (k_fused_check): Landau, Coulomb and U1xU1 switch
// near the floor itself, where a phase that gains less than 10% in a few
// checks has arrived; MAG switches far above its floor while converging by a
// few percent per hundred sweeps at large volume, so its phase ends only when
// the rate of descent itself collapses
int stall_rule(int gauge) { return gauge == GF_MAG ? 1 : 0; }

// One OR phase with the check fused into the sweeps.  `launch(parity,
// theta, partial, it)` enqueues one parity kernel of sweep `it`;
// `after_sweep(it)` runs after sweep `it` (reprojection, refresh).  Sweeps
// after convergence exit at once.  On return r->theta / r->gff describe the
// state after the first sweep that met `tol`, or the final state if maxiter
// was reached.
// The end state of sweep `it` is checked when (it + 1) % check_every == 0:
// its odd kernel forms the odd-site theta, the next even kernel the even-site
// theta before updating, and a check kernel completes the sum.  The phase ends
// once theta < tol or, with stall_limit > 0, on stagnation (r->converged = 2):
// below the gauge's stall floor, stall_limit checks in a row that do not lower
// the best theta by 10%, or for MAG a window of stall_limit checks whose mean
// log(theta) drops by less than log(1.1) and by far less than the window
// before it (k_fused_check).
// exact_end: the buffer must hold the state that met `tol` (the reference
// breaks right after that check, LandauGaugeFixingSU3_4D.cu:254-271), not that
// state plus the even half-sweep whose kernel formed the even sites' theta.
// Once theta falls below kNearFactor * tol the device sets `near`; from then
// on the even sites' theta comes from a read-only pass (launch theta = 3),
// checked before the even update, which exits if the check converged.  While
// the host has not seen theta within kHostNearFactor of tol the passes are not
// enqueued at all (a theta that falls from above kNearFactor * tol below tol
// within such a stretch costs one more sweep, whose end state is then the
// one checked and returned).
// Coulomb on a 4D lattice: each timeslice is its own problem and must meet
// the precision on its own.  The per-slice check needs the 64-site blocks to
// tile the slices; otherwise the lattice average is tested.
int check_segments(const int dims[4], int gauge) { return gauge == GF_COULOMB && dims[0] > 1 ? dims[0] : 1; }

// coarse_every > check_every: while the last theta the host has read is above
// kCoarseFactor * tol, check only every coarse_every sweeps.  Such a theta
// cannot reach tol within the next chunk, and stagnation is tested far below
// it, so only the cost of the checks changes.  0 checks every check_every
// sweeps throughout.  coarse_start: the phase also checks coarsely before its
// first read (a phase that starts far from tol) and turns fine for good at the
// first theta near tol; otherwise it starts fine and follows the theta read
// after each chunk, and also by its rate: once two reads at least kRateSpan
// sweeps apart show theta falling, the phase stays coarse while theta falling
// at that rate would take more than kFarSweeps sweeps (five chunks) to reach
// tol, however close to tol it is.  A phase that converges slowly (MAG's FP64
// phase at large volume gains a factor of 10 in some 10^4 sweeps) thus checks
// every sweep only near its end; theta would have to fall five times faster
// than over the last kRateSpan sweeps to cross tol before the next read.
constexpr double kCoarseFactor = 1e3;
constexpr double kNearFactor = 2.0;
constexpr double kHostNearFactor = 1e3;
constexpr int kRateSpan = 2 * kChunk;
constexpr int kFarSweeps = 5 * kChunk;

// plain(it): enqueue all of sweep `it` (no check involved) at once and return
// hipSuccess, or hipErrorNotSupported to have its two parity kernels launched
template <typename Launch, typename After, typename Plain>
gf_status fused_loop(const int dims[4], int64_t V, int gauge, double tol, int maxiter, int check_every,
                     int coarse_every, int stall_limit, bool exact_end, bool near, int64_t nblocks, FusedCtx &fc,
                     hipStream_t s, Launch launch, After after_sweep, Plain plain, int *sweeps_done,
                     gf_run_result *r, bool coarse_start = true)
{
    const int nseg = check_segments(dims, gauge);
    if (check_every < 1) check_every = 1;
    if (nblocks > fc.nblocks) return fail(GF_ERR_BAD_ARG, "internal: partial buffer too small");
    hipError_t e = launch_state_init(fc.st, tol, 1. / (3. * V), gff_prefactor(gauge) / V, stall_limit,
                                     stall_floor(gauge), stall_rule(gauge), exact_end ? 1 : 0, kNearFactor * tol,
                                     near ? 1 : 0, s);
    if (e != hipSuccess) return hip_fail(e, "state init");
    int it = 0;
    bool done = false;
    bool pending = false;  // a check of the state at the end of sweep it-1 is due
    bool coarse = coarse_every > check_every && coarse_start;
    bool host_near = exact_end;  // until the first read of this phase's state
    int rate_sweep = -1;         // the rate rule: sweep and log theta of an earlier read check,
    double rate_log = 0, rate = 0;  // and the fall of log theta per sweep since the one before
    auto check = [&](int mode, int sweep) -> gf_status {
        const hipError_t ec = launch_fused_check(fc.st, fc.even(), fc.odd(), nblocks, nseg, sweep, mode, s);
        if (ec != hipSuccess) return hip_fail(ec, "check kernel launch");
        return GF_OK;
    };
    while (it < maxiter && !done) {
        const int every = coarse ? coarse_every : check_every;
        const int n = std::min(kChunk, maxiter - it);
        for (int k = 0; k < n; k++, it++) {
            const bool check_prev = pending;                    // end of sweep it-1
            const bool check_this = (it + 1) % every == 0;      // end of sweep it
            const bool pre = check_prev && host_near;
            if (!check_prev && !check_this) {
                e = plain(it);
                if (e == hipSuccess) {
                    pending = false;
                    GF_TRY(after_sweep(it));
                    continue;
                }
                if (e != hipErrorNotSupported) return hip_fail(e, "sweep launch");
            }
            if (pre) {
                // read-only theta of the even sites (a no-op until `near`), its
                // check, then the even update, which exits if that converged
                e = launch(0, 3, fc.even(), it);
                if (e != hipSuccess) return hip_fail(e, "theta pass launch");
                GF_TRY(check(CHECK_PRE, it));
            }
            e = launch(0, check_prev ? 1 : 0, fc.even(), it);
            if (e != hipSuccess) return hip_fail(e, "fused sweep launch");
            if (check_prev) GF_TRY(check(pre ? CHECK_FUSED_IF_FAR : CHECK_FUSED, it));
            e = launch(1, check_this ? 2 : 0, fc.odd(), it);
            if (e != hipSuccess) return hip_fail(e, "fused sweep launch");
            pending = check_this;
            GF_TRY(after_sweep(it));
        }
        GF_TRY(fc.read(s));
        done = fc.host->done != 0;
        if (coarse_every > check_every && fc.host->checks > 0) {
            const double th = fc.host->theta;
            bool far = th > kCoarseFactor * tol;
            if (!coarse_start) {
                const int cs = fc.host->checked_sweep;
                if (th > 0 && (rate_sweep < 0 || cs - rate_sweep >= kRateSpan)) {
                    if (rate_sweep >= 0) rate = (rate_log - std::log(th)) / (cs - rate_sweep);
                    rate_sweep = cs;
                    rate_log = std::log(th);
                }
                if (rate > 0 && th > tol) far = std::log(th / tol) > rate * kFarSweeps;
                coarse = far;  // coarse only while a theta the host has read is far
            } else if (!far) {
                coarse = false;  // coarse until the first theta near tol
            }
        }
        if (exact_end && fc.host->checks > 0) host_near = fc.host->near || !(fc.host->theta > kHostNearFactor * tol);
    }
    // the host copy of the state is this loop's only after a chunk has run
    // (a phase given no sweeps must not count the previous phase's checks again)
    if (it > 0) r->checks += fc.host->checks;
    if (done) {
        *sweeps_done = fc.host->converged_sweep + 1;
        r->theta = fc.host->theta;
        r->gff = fc.host->gff;
        r->converged = fc.host->done;
    } else {
        *sweeps_done = it;
        r->converged = 0;
    }
    return GF_OK;
}

// the fused loop over the link sweep kernels (FP64, or FP32 with g accumulation)
template <typename R, typename After>
gf_status fused_phase(const gf_lattice *lat, R *U, float *gacc, int gauge, float omega, double tol, int maxiter,
                      int check_every, int coarse_every, int stall_limit, bool exact_end, bool near, FusedCtx &fc,
                      hipStream_t s, After after_sweep, int *sweeps_done, gf_run_result *r, bool coarse_start = true)
{
    int64_t V = 1;
    for (int i = 0; i < 4; i++) V *= lat->dims[i];
    SweepArgs<R> a{};
    a.U = U;
    a.Udw = U;
    a.gacc = gacc;
    a.state = fc.st;
    a.g = per_slice(make_geom(lat->dims, V), gauge);
    a.param = omega;
    a.nsb = 32;
    auto launch = [&](int parity, int theta, double *partial, int it) {
        a.sweep = it;
        a.parity = parity;
        a.partial = partial;
        return fast::launch_sweep<R>(lat->layout, gauge, GF_OR, theta, gacc != nullptr, a, s);
    };
    auto two_kernels = [](int) { return hipErrorNotSupported; };
    return fused_loop(lat->dims, V, gauge, tol, maxiter, check_every, coarse_every, stall_limit, exact_end, near,
                      sweep_grid(a.g), fc, s, launch, after_sweep, two_kernels, sweeps_done, r, coarse_start);
}

// projectSU3 enqueued by a driver after a sweep: skipped on the device once
// the run has converged, so that the checked state is what the caller gets.
// lat may have the drivers' internal layout kTiled12.
gf_status project_unless_done(const gf_lattice *lat, void *U, const DevState *st, hipStream_t s)
{
    int64_t V = 1;
    for (int i = 0; i < 4; i++) V *= lat->dims[i];
    hipError_t e = lat->dtype == GF_F64 ? exact::launch_project<double>(lat->layout, 4, (double *)U, V, s, st)
                                        : exact::launch_project<float>(lat->layout, 4, (float *)U, V, s, st);
    if (e != hipSuccess) return hip_fail(e, "project kernel launch");
    return GF_OK;
}

// The FP64 phases run on the drivers' own copy of the links in 64-site tiles
// (kTiled12): a block's 64 sites then read and write one contiguous 6 KB
// tile per direction instead of twelve 512-byte pieces V * 8 bytes apart.
struct Tiles64 {
    Scratch buf;
    gf_lattice lat{};
    double *p = nullptr;
    gf_status init(const gf_lattice *caller, int64_t V, hipStream_t s)
    {
        lat = *caller;
        lat.layout = kTiled12;
        HIP_TRY(buf.alloc(sizeof(double) * 4 * tiled12_reals(V), s));
        p = buf.as<double>();
        return GF_OK;
    }
};

// theta and gff of the current state as the fused check defines them (for
// Coulomb on a 4D lattice the largest timeslice theta): read-only theta
// passes over both parities and one check, for a run that did not converge
template <typename R>
gf_status final_quality_t(const gf_lattice *lat, R *U, int gauge, FusedCtx &fc, gf_run_result *r, hipStream_t s)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    SweepArgs<R> a{};
    a.U = U;
    a.Udw = U;
    a.state = fc.st;
    a.g = per_slice(make_geom(lat->dims, V), gauge);
    a.param = 1.7f;
    a.nsb = 32;
    const int64_t nb = sweep_grid(a.g);
    if (nb > fc.nblocks) return fail(GF_ERR_BAD_ARG, "internal: partial buffer too small");
    // tol 0 (never met), exact_end and near set: the passes run, the check reads them
    HIP_TRY(launch_state_init(fc.st, 0.0, 1. / (3. * V), gff_prefactor(gauge) / V, 0, 0.0, 0, 1, 0.0, 1, s));
    for (int par = 0; par < 2; par++) {
        a.parity = par;
        a.partial = par ? fc.odd() : fc.even();
        HIP_TRY(fast::launch_sweep<R>(lat->layout, gauge, GF_OR, 3, false, a, s));
    }
    HIP_TRY(launch_fused_check(fc.st, fc.even(), fc.odd(), nb, check_segments(lat->dims, gauge), 0, CHECK_PRE, s));
    GF_TRY(fc.read(s));
    r->theta = fc.host->theta;
    r->gff = fc.host->gff;
    return GF_OK;
}

gf_status final_quality(const gf_lattice *lat, void *U, int gauge, FusedCtx &fc, gf_run_result *r, hipStream_t s)
{
    return lat->dtype == GF_F64 ? final_quality_t<double>(lat, (double *)U, gauge, fc, r, s)
                                : final_quality_t<float>(lat, (float *)U, gauge, fc, r, s);
}

}  // namespace

extern "C" {

gf_status gf_run_fast(const gf_lattice *lat, void *U, int gauge, float omega, float precision, int maxiter,
                      int reproject, gf_run_result *out, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !out) return fail(GF_ERR_BAD_ARG, "null argument");
    if (gauge == GF_COULOMB && lat->dims[0] == 1)
        return fail(GF_ERR_BAD_ARG, "gf_run_fast fixes every timeslice of a 4D lattice at once");
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    if (lat->dtype == GF_F32 && lat->layout != GF_LAYOUT_R12)
        return fail(GF_ERR_BAD_ARG, "FP32 fused sweeps work on the compact 12-real layout");
    memset(out, 0, sizeof *out);
    hipStream_t s = S(stream);
    FusedCtx fc;
    GF_TRY(fc.init(sweep_grid(per_slice(make_geom(lat->dims, V), gauge)), s));
    Tiles64 tiles;
    const bool f64 = lat->dtype == GF_F64;
    if (f64) GF_TRY(tiles.init(lat, V, s));
    EventPair ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    HIP_TRY(hipEventRecord(ev.a, s));
    // FP64: the caller's links into the tiles, swept there, and back
    if (f64) HIP_TRY(launch_convert(GF_F64, lat->layout, U, GF_F64, kTiled12, tiles.p, V, 0, s));
    const gf_lattice *wl = f64 ? &tiles.lat : lat;
    void *W = f64 ? (void *)tiles.p : U;
    auto after = [&](int it) -> gf_status {
        if (reproject > 0 && it % reproject == 0) return project_unless_done(wl, W, fc.st, s);
        return GF_OK;
    };
    int done = 0;
    gf_status rc = f64 ? fused_phase<double>(wl, (double *)W, nullptr, gauge, omega, precision, maxiter, 1, 0, 0, true,
                                             false, fc, s, after, &done, out)
                       : fused_phase<float>(lat, (float *)U, nullptr, gauge, omega, precision, maxiter, 1, 0, 0, true,
                                            false, fc, s, after, &done, out);
    GF_TRY(rc);
    out->sweeps = done;
    // third rows (LandauGaugeFixingSU3_4D.cu:271): projectSU3, for FP64 as the
    // tiles are written back
    if (f64) HIP_TRY(launch_convert(GF_F64, kTiled12, tiles.p, GF_F64, lat->layout, U, V, 1, s));
    else GF_TRY(gf_project(lat, U, s));
    HIP_TRY(hipEventRecord(ev.b, s));
    float ms = 0;
    HIP_TRY(hipEventSynchronize(ev.b));
    HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    out->ms_total = ms;
    if (!out->converged) GF_TRY(final_quality(lat, U, gauge, fc, out, s));
    return GF_OK;
}

gf_status gf_time_sweeps(const gf_lattice *lat, void *U, int gauge, int nsweeps, int variant, float *ms_per_kernel,
                         void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !ms_per_kernel || nsweeps <= 0) return fail(GF_ERR_BAD_ARG, "bad argument");
    if (variant < 0 || variant > 7) return fail(GF_ERR_BAD_ARG, "unknown sweep variant %d", variant);
    if (gauge == GF_COULOMB && lat->dims[0] == 1) return fail(GF_ERR_BAD_ARG, "gf_time_sweeps times 4D lattices");
    if (lat->dtype == GF_F32 && lat->layout != GF_LAYOUT_R12)
        return fail(GF_ERR_BAD_ARG, "FP32 fused sweeps work on the compact 12-real layout");
    if (variant > 0 && variant != 5 && variant != 7 && lat->dtype != GF_F32)
        return fail(GF_ERR_BAD_ARG, "variants 1-4 and 6 time FP32 kernels");
    if (variant == 7 && lat->dtype != GF_F64) return fail(GF_ERR_BAD_ARG, "variant 7 times an FP64 kernel");
    if (variant == 7 && !gauge_sweep64_fits(V))
        return fail(GF_ERR_BAD_SHAPE, "the FP64 gauge-transformation sweep takes at most 22,369,600 sites");
    if (variant >= 2 && variant != 5) GF_TRY(check_gauge_sweep_volume(V));
    hipStream_t s = S(stream);
    if (variant == 5) {
        // the drop-in quality kernel (GaugeFixingStats::generateGaugeQuality's
        // per-site pass with its block partials), U unchanged
        Scratch part;
        HIP_TRY(part.alloc(sizeof(double) * 2 * exact::quality_blocks(V), s));
        EventPair ev;
        HIP_TRY(hipEventCreate(&ev.a));
        HIP_TRY(hipEventCreate(&ev.b));
        const int warm = std::max(2, nsweeps / 2);
        for (int k = 0; k < warm + nsweeps; k++) {
            if (k == warm) HIP_TRY(hipEventRecord(ev.a, s));
            GF_TRY(quality_launch(lat, gauge, U, nullptr, nullptr, part.as<double>(), 0, s));
        }
        HIP_TRY(hipEventRecord(ev.b, s));
        float ms = 0;
        HIP_TRY(hipEventSynchronize(ev.b));
        HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
        *ms_per_kernel = ms / nsweeps;
        return GF_OK;
    }
    FusedCtx fc;
    GF_TRY(fc.init(sweep_grid(make_geom(lat->dims, V)), s));
    Tiles64 tiles;  // FP64: the drivers' tiled copy, as gf_run_fast sweeps it
    if (lat->dtype == GF_F64) {
        GF_TRY(tiles.init(lat, V, s));
        HIP_TRY(launch_convert(GF_F64, lat->layout, U, GF_F64, kTiled12, tiles.p, V, 0, s));
    }
    Scratch gbuf, pbuf;
    struct {
        float *p = nullptr;
    } gb;
    if (variant == 7) {  // an FP64 g on the tiles, as the FP64 phase of gf_run_mixed's gauge modes
        HIP_TRY(gbuf.alloc(sizeof(double) * gauge_floats(V), s));
        gb.p = gbuf.as<float>();
        HIP_TRY(launch_gauge_identity64(gbuf.as<double>(), V, s));
    } else if (variant > 0) {
        HIP_TRY(gbuf.alloc(sizeof(float) * gauge_floats(V), s));
        gb.p = gbuf.as<float>();
        HIP_TRY(launch_gauge_identity(gb.p, V, s));
    }
    if (variant == 4 || variant == 6) {  // the links in the 24-bit fixed-point form
        HIP_TRY(pbuf.alloc(sizeof(uint32_t) * links24_words(V), s));
        HIP_TRY(launch_pack24(GF_F32, 12, U, pbuf.as<uint32_t>(), V, s));
    }
    // tol = 0: no state ever converges, every kernel does its full work
    HIP_TRY(launch_state_init(fc.st, 0.0, 1. / (3. * V), gff_prefactor(gauge) / V, 0, 0.0, 0, 0, 0.0, 0, s));
    EventPair ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    const Geom geo = make_geom(lat->dims, V);
    int slab = 0;
    GaugeSweepArgs wa{};  // variant 6: whole sweeps as the mixed run's FP32 phase runs them
    PipeBuf pipe;
    if (variant == 6) {
        wa.U0 = pbuf.as<const float>(); wa.g = gb.p; wa.geo = geo; wa.param = 1.7f; wa.remap = 1; wa.links24 = 1;
        HIP_TRY(pipe.init(geo, gauge, false, s));
        wa.pipe = pipe.get();
        wa.pipe_count = &pipe.count;
        slab = wavefront_slab(wa, gauge, false, s);
    }
    int tpass = 0;  // the gauge-transformation sweeps' block order, as the drivers use it
    if ((variant >= 2 && variant <= 4) || variant == 7) {
        GaugeSweepArgs t{};
        t.U0 = variant == 7 ? (const float *)tiles.p : variant == 4 ? pbuf.as<const float>() : (const float *)U;
        t.g = gb.p; t.geo = geo; t.param = 1.7f; t.remap = 1; t.links24 = variant == 4; t.links64 = variant == 7;
        (void)wavefront_slab(t, gauge, variant == 3 || variant == 7, s);
        tpass = t.tpass;
    }
    auto run = [&](int n, int first) -> gf_status {
        for (int it = 0; it < n; it++) {
            if (variant == 6) {
                hipError_t e = slab > 0 ? launch_gsweep_wavefront(gauge, GF_OR, false, wa, slab, s) : hipErrorNotSupported;
                if (e == hipErrorNotSupported) {  // the two kernels
                    e = hipSuccess;
                    for (int par = 0; par < 2 && e == hipSuccess; par++) {
                        GaugeSweepArgs b = wa;
                        b.parity = par;
                        e = launch_gsweep(gauge, GF_OR, 0, false, b, s);
                    }
                }
                HIP_TRY(e);
                continue;
            }
            for (int par = 0; par < 2; par++) {
                double *partial = par ? fc.odd() : fc.even();
                if (variant == 7) {
                    GaugeSweepArgs a{};
                    a.U0 = (const float *)tiles.p; a.g = gb.p; a.links64 = 1; a.state = fc.st; a.partial = partial;
                    a.geo = geo; a.parity = par; a.param = 1.7f; a.remap = 1; a.tpass = tpass;
                    HIP_TRY(launch_gsweep(gauge, GF_OR, par + 1, true, a, s));
                } else if (lat->dtype == GF_F64) {
                    SweepArgs<double> a{};
                    a.U = tiles.p; a.Udw = a.U; a.state = fc.st; a.g = geo; a.param = 1.7f; a.nsb = 32;
                    a.sweep = first + it; a.parity = par; a.partial = partial;
                    HIP_TRY(fast::launch_sweep<double>(kTiled12, gauge, GF_OR, par + 1, false, a, s));
                } else if (variant <= 1) {
                    SweepArgs<float> a{};
                    a.U = (float *)U; a.Udw = a.U; a.gacc = gb.p; a.state = fc.st; a.g = geo; a.param = 1.7f;
                    a.nsb = 32; a.sweep = first + it; a.parity = par; a.partial = partial;
                    HIP_TRY(fast::launch_sweep<float>(12, gauge, GF_OR, variant == 0 ? par + 1 : 0, variant == 1, a,
                                                      s));
                } else {
                    GaugeSweepArgs a{};
                    const bool q24 = variant == 4;
                    a.U0 = q24 ? pbuf.as<const float>() : (const float *)U;
                    a.g = gb.p; a.state = fc.st; a.partial = partial; a.geo = geo;
                    a.parity = par; a.param = 1.7f; a.remap = 1; a.links24 = q24; a.tpass = tpass;
                    HIP_TRY(launch_gsweep(gauge, GF_OR, 0, variant == 3, a, s));
                }
            }
        }
        return GF_OK;
    };
    // warm-up: long enough for the clocks and caches to reach the steady state
    // of a sustained run (a few sweeps alone read up to 20% slow)
    const int warm = std::max(2, nsweeps / 2);
    GF_TRY(run(warm, 0));
    HIP_TRY(hipEventRecord(ev.a, s));
    GF_TRY(run(nsweeps, warm));
    HIP_TRY(hipEventRecord(ev.b, s));
    float ms = 0;
    HIP_TRY(hipEventSynchronize(ev.b));
    HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    *ms_per_kernel = ms / (2.0f * nsweeps);
    GF_TRY(pipe.check(s));
    return GF_OK;
}

gf_status gf_run_mixed(const gf_lattice *lat, void *U, int gauge, float omega, float switch_precision, float precision,
                       int maxiter, int mode, int refresh, int check32, int reproject, gf_run_result *out,
                       void *stream)
{
    // FP32 checks per stagnation window (below stall_floor); MAG converges
    // more slowly and gets a longer window
    const int kStall32 = gauge == GF_MAG ? 50 : 8;
    // far above the switch threshold the FP32 phase checks 5x less often
    const int kCoarse = 5;
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !out) return fail(GF_ERR_BAD_ARG, "null argument");
    if (lat->dtype != GF_F64) return fail(GF_ERR_BAD_ARG, "the mixed driver takes an FP64 configuration");
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    if (mode < GF_MIXED_LINKS || mode > GF_MIXED_FP64) return fail(GF_ERR_BAD_ARG, "unknown mode %d", mode);
    // GF_MIXED_FP64: no FP32 phase, the FP64 phase from the start
    const bool fp32_phase = mode != GF_MIXED_FP64;
    if (mode != GF_MIXED_LINKS && fp32_phase) GF_TRY(check_gauge_sweep_volume(V));
    if (gauge == GF_COULOMB && lat->dims[0] == 1)
        return fail(GF_ERR_BAD_ARG, "gf_run_mixed fixes every timeslice of a 4D lattice at once");
    if (gauge == GF_COULOMB && mode == GF_MIXED_LINKS)
        return fail(GF_ERR_BAD_ARG, "Coulomb runs with the gauge-transformation modes");
    memset(out, 0, sizeof *out);
    hipStream_t s = S(stream);
    gf_lattice l32 = *lat;
    l32.dtype = GF_F32;
    l32.layout = GF_LAYOUT_R12;
    const Geom geo = per_slice(make_geom(lat->dims, V), gauge);
    Scratch sU32, sg, sg64;
    if (fp32_phase) {
        HIP_TRY(sU32.alloc(mode == GF_MIXED_GAUGE24 ? sizeof(uint32_t) * links24_words(V) : gf_lattice_bytes(&l32),
                           s));
        HIP_TRY(sg.alloc(sizeof(float) * gauge_floats(V), s));
    }
    // the gauge modes' FP64 phase sweeps an FP64 g on the FP64 links (the
    // link mode, and lattices beyond the FP64 gauge sweep's offsets, sweep
    // the links themselves)
    const bool gauge64 = mode != GF_MIXED_LINKS && gauge_sweep64_fits(V);
    if (gauge64) HIP_TRY(sg64.alloc(sizeof(double) * gauge_floats(V), s));
    struct {
        void *U32, *g;
    } b{sU32.p, sg.p};
    FusedCtx fc;
    GF_TRY(fc.init(sweep_grid(geo), s));
    Tiles64 tiles;  // the FP64 phase's links
    GF_TRY(tiles.init(lat, V, s));
    EventPair ev;
    hipEvent_t mid = nullptr;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    HIP_TRY(hipEventCreate(&mid));
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void)hipEventDestroy(e); }
    } mg{mid};
    float *g = (float *)b.g;
    HIP_TRY(hipEventRecord(ev.a, s));
    if (fp32_phase) {
        HIP_TRY(launch_gauge_identity(g, V, s));
        if (mode == GF_MIXED_GAUGE24) HIP_TRY(launch_pack24(GF_F64, lat->layout, U, (uint32_t *)b.U32, V, s));
        else HIP_TRY(launch_convert(lat->dtype, lat->layout, U, GF_F32, 12, b.U32, V, 0, s));
    }
    int done32 = 0;
    if (fp32_phase && mode == GF_MIXED_LINKS) {
        // FP32 links and gauge transformation; every `refresh` sweeps the
        // working links are rebuilt from the FP64 configuration and the
        // (reprojected) transformation, which removes the FP32 drift of the
        // gauge-invariant content.
        auto after32 = [&](int it) -> gf_status {
            const int k = it + 1;
            if (refresh > 0 && k % refresh == 0) {
                HIP_TRY(launch_project_gauge(g, V, s));
                HIP_TRY(launch_gauge_apply((const double *)U, lat->layout, g, b.U32, GF_F32, 12, geo, s));
            } else if (reproject > 0 && it % reproject == 0) {
                HIP_TRY(exact::launch_project<float>(12, 4, (float *)b.U32, V, s));
            }
            return GF_OK;
        };
        GF_TRY(fused_phase<float>(&l32, (float *)b.U32, g, gauge, omega, switch_precision, maxiter, check32,
                                  kCoarse * check32, kStall32, false, false, fc, s, after32, &done32, out));
    } else if (fp32_phase) {
        // the links stay U0 (FP32 copy, read only); only g is swept, and the
        // kernel reprojects it in every update
        GaugeSweepArgs ga{};
        ga.U0 = (const float *)b.U32;
        ga.g = g;
        ga.state = fc.st;
        ga.geo = geo;
        ga.param = omega;
        ga.remap = 1;
        ga.links24 = mode == GF_MIXED_GAUGE24;
        PipeBuf pipe;
        HIP_TRY(pipe.init(geo, gauge, false, s));
        ga.pipe = pipe.get();
        ga.pipe_count = &pipe.count;
        auto launch = [&](int parity, int theta, double *partial, int) {
            ga.parity = parity;
            ga.partial = partial;
            return launch_gsweep(gauge, GF_OR, theta, mode == GF_MIXED_GAUGE64, ga, s);
        };
        auto none = [](int) -> gf_status { return GF_OK; };
        // sweeps without a check: as a wavefront over timeslices where that
        // pays (lattices beyond the Infinity Cache)
        const int slab = wavefront_slab(ga, gauge, mode == GF_MIXED_GAUGE64, s);
        if (mode != GF_MIXED_GAUGE64) {
            // stage B's sweeps (below), tuned with the FP32 phase's so that a
            // later run of this kind of lattice times no candidates
            GaugeSweepArgs gb = ga;
            (void)wavefront_slab(gb, gauge, true, s);
        }
        auto plain = [&](int) {
            if (slab <= 0) return hipErrorNotSupported;
            return launch_gsweep_wavefront(gauge, GF_OR, mode == GF_MIXED_GAUGE64, ga, slab, s);
        };
        GF_TRY(fused_loop(lat->dims, V, gauge, switch_precision, maxiter, check32, kCoarse * check32, kStall32,
                          false, false, sweep_grid(geo), fc, s, launch, none, plain, &done32, out));
        if (out->converged == 2 && mode != GF_MIXED_GAUGE64 && done32 < maxiter) {
            // FP32 arithmetic stagnated above switch_precision (its rounding
            // floor, near theta ~ 2e-12 for Landau): the same g and links
            // with FP64 arithmetic, whose floor is the FP32 storage's (~4e-13)
            // -- a third of the cost of an FP64 link sweep
            auto launch64 = [&](int parity, int theta, double *partial, int) {
                ga.parity = parity;
                ga.partial = partial;
                return launch_gsweep(gauge, GF_OR, theta, true, ga, s);
            };
            const int slab64 = wavefront_slab(ga, gauge, true, s);
            auto plain64 = [&](int) {
                if (slab64 <= 0) return hipErrorNotSupported;
                return launch_gsweep_wavefront(gauge, GF_OR, true, ga, slab64, s);
            };
            int done_b = 0;
            GF_TRY(fused_loop(lat->dims, V, gauge, switch_precision, maxiter - done32, check32, kCoarse * check32,
                              kStall32, false, false, sweep_grid(geo), fc, s, launch64, none, plain64, &done_b, out));
            done32 += done_b;
        }
        GF_TRY(pipe.check(s));
    }
    out->sweeps_fp32 = done32;
    const bool reached_switch = fp32_phase && out->converged == 1;
    if (fp32_phase) {
        // switch: apply the accumulated transformation to the FP64
        // configuration, into the FP64 phase's tiles
        HIP_TRY(launch_project_gauge(g, V, s));
        HIP_TRY(launch_gauge_apply((const double *)U, lat->layout, g, tiles.p, GF_F64, kTiled12, geo, s));
    } else {
        HIP_TRY(launch_convert(GF_F64, lat->layout, U, GF_F64, kTiled12, tiles.p, V, 0, s));
    }
    HIP_TRY(hipEventRecord(mid, s));
    auto after64 = [&](int it) -> gf_status {
        if (reproject > 0 && it % reproject == 0) return project_unless_done(&tiles.lat, tiles.p, fc.st, s);
        return GF_OK;
    };
    int done64 = 0;
    out->converged = 0;
    const int left = std::max(0, maxiter - done32);
    // an FP32 phase that went down to (near) the precision hands over a state
    // that may meet it after a sweep or two: check it before the even update
    // from the start (not after a stagnated phase, whose theta is further off:
    // the device sets `near` once a check comes close)
    const bool near = reached_switch && switch_precision < kNearFactor * precision;
    // a check costs the fused theta in the sweep kernels and a check kernel
    // (MAG's a third of a sweep in this phase: its theta is formed from the
    // links in every wave), so the phase checks every 16 sweeps while a theta
    // the host has read is far from the precision (fused_loop's rules: MAG
    // far from the target converges by a few percent per hundred sweeps)
    const int coarse64 = 16;
    if (gauge64) {
        // FP64 sweeps of a second gauge transformation g64 (from 1) on the
        // links G U0 G^dagger, read only: the same OR updates as sweeping the
        // links, with 1056 instead of 1536 bytes per site-update; g64 is
        // re-orthonormalised in every update, so nothing is reprojected
        double *g64 = sg64.as<double>();
        HIP_TRY(launch_gauge_identity64(g64, V, s));
        GaugeSweepArgs gd{};
        gd.U0 = (const float *)tiles.p;
        gd.g = (float *)g64;
        gd.state = fc.st;
        gd.geo = geo;
        gd.param = omega;
        gd.remap = 1;
        gd.links64 = 1;
        auto launch64 = [&](int parity, int theta, double *partial, int) {
            gd.parity = parity;
            gd.partial = partial;
            return launch_gsweep(gauge, GF_OR, theta, true, gd, s);
        };
        const int slab64 = wavefront_slab(gd, gauge, true, s);
        auto plain64 = [&](int) {
            if (slab64 <= 0) return hipErrorNotSupported;
            return launch_gsweep_wavefront(gauge, GF_OR, true, gd, slab64, s);
        };
        auto none = [](int) -> gf_status { return GF_OK; };
        GF_TRY(fused_loop(lat->dims, V, gauge, precision, left, 1, coarse64, 0, true, near, sweep_grid(geo), fc, s,
                          launch64, none, plain64, &done64, out, false));
        out->sweeps = done32 + done64;
        // U = G64 (G U0 G^dagger) G64^dagger into the caller's links, projected
        // (third rows, LandauGaugeFixingSU3_4D.cu:271)
        HIP_TRY(launch_gauge_apply64(tiles.p, g64, (double *)U, lat->layout, geo, 1, s));
    } else {
        GF_TRY(fused_phase<double>(&tiles.lat, tiles.p, nullptr, gauge, omega, precision, left, 1, coarse64, 0, true,
                                   near, fc, s, after64, &done64, out, false));
        out->sweeps = done32 + done64;
        // back to the caller's links, projected (third rows, LandauGaugeFixingSU3_4D.cu:271)
        HIP_TRY(launch_convert(GF_F64, kTiled12, tiles.p, GF_F64, lat->layout, U, V, 1, s));
    }
    HIP_TRY(hipEventRecord(ev.b, s));
    float ms = 0, ms1 = 0;
    HIP_TRY(hipEventSynchronize(ev.b));
    HIP_TRY(hipEventElapsedTime(&ms, ev.a, ev.b));
    HIP_TRY(hipEventElapsedTime(&ms1, ev.a, mid));
    out->ms_total = ms;
    out->ms_fp32 = ms1;
    if (!out->converged) GF_TRY(final_quality(lat, U, gauge, fc, out, s));
    return GF_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free OR driver
// gf_run_fast for callers that cannot allocate: the run's device state and
// block partials live in `scratch`, a file-scope device array that each GPU
// holds once for the life of the process.  Host code never passes the
// array's name to a kernel (there it is a host-side stand-in): the device
// address of each GPU's copy comes from hipGetSymbolAddress, taken once per
// device on the first call with that device current.
namespace {
constexpr int64_t kNoAllocBlocks = 131072;  // 64-site blocks per half-sweep: up to 16,777,216 sites (64^4)
constexpr size_t kNoAllocScratchBytes = fused_scratch_bytes(kNoAllocBlocks);
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
gf_status noalloc_scratch(void **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GF_ERR_BAD_ARG, "device %d beyond the %d the library tracks", dev,
                                                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GF_ERR_HIP, "hipGetSymbolAddress returned a null address for the scratch array");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return GF_OK;
}
}  // namespace

extern "C" gf_status gf_run_fast_noalloc(const gf_lattice *lat, void *U, int gauge, float omega, float precision,
                                         int maxiter, int reproject, gf_run_result *out, void *stream)
{
    int64_t V;
    GF_TRY(check_lattice(lat, &V));
    if (!U || !out) return fail(GF_ERR_BAD_ARG, "null argument");
    if (gauge == GF_COULOMB && lat->dims[0] == 1)
        return fail(GF_ERR_BAD_ARG, "gf_run_fast_noalloc fixes every timeslice of a 4D lattice at once");
    if (gauge < GF_LANDAU || gauge > GF_U1XU1) return fail(GF_ERR_BAD_ARG, "unknown gauge %d", gauge);
    if (lat->dtype == GF_F32 && lat->layout != GF_LAYOUT_R12)
        return fail(GF_ERR_BAD_ARG, "FP32 fused sweeps work on the compact 12-real layout");
    const int64_t nblocks = sweep_grid(per_slice(make_geom(lat->dims, V), gauge));
    if (nblocks > kNoAllocBlocks)
        return fail(GF_ERR_BAD_SHAPE,
                    "gf_run_fast_noalloc: %lld blocks of 64 sites per half-sweep exceed its fixed scratch of %lld "
                    "(%zu bytes; at most 16777216 sites, 64^4); use gf_run_fast",
                    (long long)nblocks, (long long)kNoAllocBlocks, kNoAllocScratchBytes);
    memset(out, 0, sizeof *out);
    hipStream_t s = S(stream);
    void *mem = nullptr;
    GF_TRY(noalloc_scratch(&mem));
    FusedCtx fc;
    fc.init_at(mem, nblocks);
    const auto t0 = std::chrono::steady_clock::now();
    auto after = [&](int it) -> gf_status {
        if (reproject > 0 && it % reproject == 0) return project_unless_done(lat, U, fc.st, s);
        return GF_OK;
    };
    int done = 0;
    gf_status rc = lat->dtype == GF_F64
                       ? fused_phase<double>(lat, (double *)U, nullptr, gauge, omega, precision, maxiter, 1, 0, 0, true,
                                             false, fc, s, after, &done, out)
                       : fused_phase<float>(lat, (float *)U, nullptr, gauge, omega, precision, maxiter, 1, 0, 0, true,
                                            false, fc, s, after, &done, out);
    GF_TRY(rc);
    out->sweeps = done;
    // third rows (LandauGaugeFixingSU3_4D.cu:271), projected in place as
    // gf_run_fast projects its copy on the way back
    if (lat->dtype == GF_F64) HIP_TRY(launch_convert(GF_F64, lat->layout, U, GF_F64, lat->layout, U, V, 1, s));
    else GF_TRY(gf_project(lat, U, s));
    HIP_TRY(hipStreamSynchronize(s));
    out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!out->converged) GF_TRY(final_quality(lat, U, gauge, fc, out, s));
    return GF_OK;
}
