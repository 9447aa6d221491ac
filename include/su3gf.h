/*
 * su3gf.h -- C ABI of the MI355X (gfx950) SU(3) gauge-fixing library.
 *
 * Plain C: raw device pointers, sizes and status codes; no C++ or torch types.
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference's src/ directory).  The C++ headers in include/culgt/ keep
 * the reference's class names and forward to these functions.
 *
 * Link storage: component `comp` of link (site idx, direction mu) lives at
 *   U[idx + V*(comp + stride*mu)],  comp = c + 2*(j + 3*i)
 * with idx the parity-split site index (even sites first), i/j row/column,
 * c = 0 real / 1 imaginary.  stride = 18 is the reference's GpuPattern layout
 * (lattice/access_pattern/GpuPattern.hxx:53-55); stride = 12 stores rows 0,1
 * only.  dims = {Nt, Nx, Ny, Nz}, mu = 0 is time, every extent even.
 * A Coulomb timeslice buffer is a lattice with dims {1, Nx, Ny, Nz}.
 *
 * Every entry returns a gf_status; gf_last_error() describes the last failure
 * on the calling thread.  `stream` is a hipStream_t (NULL = default stream).
 */
#ifndef SU3GF_H_
#define SU3GF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GF_OK = 0,
    GF_ERR_BAD_SHAPE = 1,     /* odd or non-positive extent, size mismatch */
    GF_ERR_BAD_ARG = 2,       /* null pointer, unknown enum, bad parameter */
    GF_ERR_HIP = 3,           /* a HIP runtime call failed */
    GF_ERR_NOT_CONVERGED = 4, /* a driver hit its iteration limit */
    GF_ERR_NO_DEVICE = 5,     /* no usable gfx950 device */
    GF_ERR_IO = 6             /* a configuration file cannot be read or written (su3gf_io.h) */
} gf_status;

/* lattice/datatype/lattice_typedefs.h:39 (enum GaugeType) */
typedef enum { GF_LANDAU = 0, GF_COULOMB = 1, GF_MAG = 2, GF_U1XU1 = 3 } gf_gauge;

/* gaugefixing/algorithms/: OrUpdate, MicroUpdate, SaUpdate, SrUpdate, RandomUpdate */
typedef enum { GF_OR = 0, GF_MICRO = 1, GF_SA = 2, GF_SR = 3, GF_RANDOM = 4 } gf_algo;

/* OR'ed into gf_sweep's algo (GF_OR, GF_MICRO or GF_SA) on FP32 links: the
 * SU(2) update rule is computed in double from the FP32 subgroup sums and
 * rounded back, as in the reference's single-precision build with
 * DPUPDATES=true (-DUSE_DP_ORUPDATE / USE_DP_MICROUPDATE / USE_DP_SAUPDATE,
 * gaugefixing/apps/Makefile:46-48; algorithms/OrUpdate.hxx:54-64,
 * MicroUpdate.hxx:45-54, SaUpdate.hxx:54-98).  A no-op on FP64 links. */
#define GF_UPDATE_DP 0x100

typedef enum { GF_F32 = 0, GF_F64 = 1 } gf_dtype;

typedef enum { GF_LAYOUT_R12 = 12, GF_LAYOUT_PATTERN18 = 18 } gf_layout;

/* lattice/datatype/lattice_typedefs.h:42 (enum StoppingCrit) */
typedef enum { GF_CRIT_MAX = 0, GF_CRIT_AVERAGE = 1 } gf_crit;

/* single-precision phase of gf_run_mixed (no reference counterpart) */
typedef enum {
    GF_MIXED_LINKS = 0,   /* FP32 links updated together with an FP32 g(x) (any gauge) */
    GF_MIXED_GAUGE32 = 1, /* g(x) swept on read-only FP32 links, FP32 arithmetic */
    GF_MIXED_GAUGE64 = 2, /* g(x) swept on read-only FP32 links, FP64 arithmetic */
    GF_MIXED_GAUGE24 = 3, /* g(x) swept on read-only links in 24-bit fixed point, FP32 arithmetic */
    GF_MIXED_FP64 = 4     /* no FP32 phase: the FP64 phase from the start */
} gf_mixed_mode;

typedef struct {
    int dims[4];    /* {Nt, Nx, Ny, Nz} */
    int dtype;      /* gf_dtype */
    int layout;     /* gf_layout */
} gf_lattice;

/* ---------------------------------------------------------------- status */
const char *gf_last_error(void);
const char *gf_version(void);

/* ------------------------------------------------- device and memory
 * Thin wrappers over the HIP runtime so that callers without a HIP binding
 * (ctypes, cgo, JNI) can own device buffers.  The blocking copies and
 * gf_memset return after the device has finished them.  These replace the apps'
 * cudaMalloc / cudaMemcpy / cudaFree calls
 * (gaugefixing/apps/LandauGaugeFixingSU3_4D.cu:98-116,189,280). */
gf_status gf_device_info(int *count, char *name, size_t name_len);
gf_status gf_malloc(void **ptr, size_t bytes);
gf_status gf_free(void *ptr);
gf_status gf_memcpy_h2d(void *dst, const void *src, size_t bytes);
gf_status gf_memcpy_d2h(void *dst, const void *src, size_t bytes);
gf_status gf_memcpy_d2d(void *dst, const void *src, size_t bytes);   /* returns when done */
/* stream-ordered copy of any direction (host buffers should be pinned) */
gf_status gf_memcpy_async(void *dst, const void *src, size_t bytes, void *stream);
gf_status gf_memset(void *ptr, int value, size_t bytes);
gf_status gf_synchronize(void);
gf_status gf_stream_create(void **stream);
gf_status gf_stream_destroy(void *stream);
gf_status gf_stream_synchronize(void *stream);
gf_status gf_event_create(void **event);
gf_status gf_event_destroy(void *event);
gf_status gf_event_record(void *event, void *stream);
gf_status gf_event_elapsed_ms(void *start, void *stop, float *ms);

/* Bytes of one link array for a lattice (V * 4 * layout * sizeof(real)). */
size_t gf_lattice_bytes(const gf_lattice *lat);

/* ---------------------------------------------------- operator API
 * One parity half-sweep with update rule `algo`.
 * Replaces LandauKernelsSU3::{orStep,microStep,saStep,randomTrafo}
 *   (gaugefixing/LandauKernelsSU3.hxx:55-105, kernels :216-286),
 * MAGKernelsSU3::{orStep,microStep,saStep,srStep,randomTrafo}
 *   (gaugefixing/MAGKernelsSU3.hxx:57-110, kernels :248-327),
 * U1xU1KernelsSU3::{orStep,microStep,saStep,randomTrafo}
 *   (gaugefixing/U1xU1KernelsSU3.hxx:55-105, kernels :216-286), and for
 * gauge = GF_COULOMB, CoulombKernelsSU3::{orStep,microStep,saStep,randomTrafo}
 *   (gaugefixing/CoulombKernelsSU3.hxx:53-101, kernels :208-279), where U is
 * the timeslice t buffer (UtUp), Udw the slice t-1 buffer (UtDw) and dims
 * {1, Nx, Ny, Nz}.  Udw is ignored for the other gauges.
 * param: OR/SR parameter or SA temperature (a float, as in the reference).
 * seed, counter: the Philox key word and launch counter of the RNG updates
 * (SA, SR, RANDOM); nsb: the reference's sites-per-block, which enters the RNG
 * key (LandauKernelsSU3.hxx:272); 0 selects the reference default 32.
 * The neighbour table the reference passes is computed on the fly. */
gf_status gf_sweep(const gf_lattice *lat, void *U, void *Udw, int parity, int gauge, int algo,
                   float param, int seed, uint32_t counter, int nsb, void *stream);

/* Reunitarize every link and, for the 18-real layout, write its third row.
 * Replaces CommonKernelsSU3::projectSU3 (gaugefixing/CommonKernelsSU3.hxx:40-62,71-90). */
gf_status gf_project(const gf_lattice *lat, void *U, void *stream);

/* Random SU(3) links (product of three random SU(2) subgroup elements).
 * Replaces CommonKernelsSU3::setHot (gaugefixing/CommonKernelsSU3.hxx:92-130). */
gf_status gf_set_hot(const gf_lattice *lat, void *U, int seed, uint32_t counter, void *stream);

/* Copy links between layouts / precisions (rows 0,1; the third row is
 * reconstructed for an 18-real destination).  reunitarize != 0 projects each
 * link in the destination precision.  No reference counterpart: it moves a
 * configuration between the reference layout and the compact one. */
gf_status gf_convert(const gf_lattice *src_lat, const void *src, const gf_lattice *dst_lat, void *dst,
                     int reunitarize, void *stream);

/* The read-only link copy of GF_MIXED_GAUGE24 and gf_run_sa mode 2: rows 0,1
 * of every link in 24-bit fixed point.  Each real v becomes
 * q = round(v (2^23 - 1)) (v clamped to [-1, 1], ties to even), a 24-bit
 * two's-complement integer; four of them, a b c d in the order
 * Re u00, Im u00, Re u01, ... Im u12, share three words:
 *   w0 = a | b << 24,   w1 = b >> 8 | c << 16,   w2 = c >> 16 | d << 8.
 * The words lie in tiles of 64 sites per direction: with T = ceil(V / 64),
 * word w (0..8) of link (i, mu) is at
 *   P[576 T mu + 9 (i - i % 64) + 64 w + i % 64],
 * and P holds 2304 T words (gf_links24_words).
 * gf_pack_links24 writes P from U (precision and layout of lat);
 * gf_unpack_links24 decodes P the way the sweep kernels do, as
 * (float)q * (float)(1 / (2^23 - 1)), into FP32 rows 0,1 in the compact layout
 * (U32 holds 48 V floats).  No reference counterpart. */
size_t gf_links24_words(const int dims[4]);
gf_status gf_pack_links24(const gf_lattice *lat, const void *U, uint32_t *P, void *stream);
gf_status gf_unpack_links24(const int dims[4], const uint32_t *P, float *U32, void *stream);

/* ------------------------------------------------------ quality / stats
 * Replaces GaugeFixingStats<Ndim,Nc,GType,StoppingCrit>
 * (gaugefixing/GaugeFixingStats.hxx:55-99,186-232): owns the reduction
 * scratch for one lattice shape.  gauge selects the quality kernel
 * (Landau/U1xU1: LandauKernelsSU3.hxx:114-191; Coulomb: CoulombKernelsSU3.hxx:110-183;
 * MAG: MAGKernelsSU3.hxx:116-223). */
typedef struct gf_stats gf_stats;

gf_status gf_stats_create(const int dims[4], int gauge, int crit, gf_stats **out);
gf_status gf_stats_destroy(gf_stats *st);

/* generateGaugeQuality + getCurrentGff/getCurrentA: gff and theta with the
 * reference's prefactors (gff/(Nc*Ndim*V) -- Coulomb Nc*(Ndim-1) --, theta/Nc,
 * and /V for GF_CRIT_AVERAGE). */
gf_status gf_stats_quality(gf_stats *st, const gf_lattice *lat, const void *U, double *gff, double *theta,
                           void *stream);

/* generateGaugeQualityPerSite (e.g. LandauKernelsSU3.hxx:71-74): per-site
 * gff_x and theta_x into caller-owned device arrays of V doubles each.
 * Coulomb on a 4D lattice gives every timeslice's spatial quality (and
 * gf_stats_quality its average over the slices). */
gf_status gf_quality_per_site(const gf_lattice *lat, int gauge, const void *U, double *dGff, double *dA,
                              void *stream);

/* ---------------------------------------------------------- drivers */
typedef struct {
    int sweeps;          /* OR sweeps done (per phase summed) */
    int sweeps_fp32;     /* of which in FP32 (mixed driver) */
    int checks;          /* quality evaluations */
    double gff;          /* last evaluated functional */
    double theta;        /* last evaluated theta */
    double ms_total;     /* device time of the whole run (hipEvents) */
    double ms_fp32;      /* device time of the FP32 phase (mixed driver) */
    int converged;       /* 1: theta < precision reached (2: a phase stagnated) */
} gf_run_result;

/* OR loop with the reference's control semantics
 * (gaugefixing/apps/LandauGaugeFixingSU3_4D.cu:244-262): for i < maxiter:
 * two parity sweeps; projectSU3 when i % reproject == 0; a quality check when
 * i % checkprecision == 0 that stops once theta < precision. */
gf_status gf_run_or(const gf_lattice *lat, void *U, int gauge, float omega, float precision, int maxiter,
                    int checkprecision, int reproject, gf_run_result *out, void *stream);

/* Simulated annealing with the reference apps' temperature schedules:
 *   Landau / Coulomb: `steps` temperatures from tmax down by (tmax-tmin)/steps
 *     (gaugefixing/apps/LandauGaugeFixingSU3_4D.cu:205-240,
 *     CoulombGaugeFixingSU3_4D.cu:229-265);
 *   U1xU1: down to tmin with 15 times finer steps for 2 <= T <= 2.7
 *     (U1xU1GaugeFixingSU3_4D.cu:212-251);
 *   MAG: the same with 0.7 <= T <= 0.88 and five heatbath steps per
 *     temperature (MAGaugeFixingSU3_4D.cu:222-265).
 * A heatbath step is an SA sweep of both parities (RNG counters *counter,
 * *counter + 1, ...; *counter is advanced like PhiloxWrapper::getNextCounter)
 * followed by `micro` microcanonical sweeps; projectSU3 when i % reproject
 * == 0 for temperature i.  Udw: the slice t-1 buffer of a Coulomb timeslice
 * (projected together with U), otherwise ignored.  *temperatures (optional)
 * receives the number of temperatures visited.
 * mode 0: the reference's sweeps on U in the precision of lat (asynchronous
 * on `stream`).  mode 1 (FP64 lat; any gauge, Coulomb on a 4D lattice with
 * every timeslice at once): the gauge transformation is annealed in FP32 on a
 * read-only FP32 copy of U (same update rules and random numbers, FP32
 * arithmetic as in the reference's single-precision build), then applied to U
 * in FP64 and U reprojected; returns when done.  mode 2: as mode 1 with the
 * copy of U in 24-bit fixed point (see GF_MIXED_GAUGE24).  Modes 1 and 2
 * take at most 44,739,200 sites, as the gauge modes of gf_run_mixed.
 * mode | GF_SA_GRAPH: the whole schedule (every launch, with its temperature
 * and RNG counter) is captured into one HIP graph on `stream` (or a stream of
 * the call's own when `stream` is null) and replayed once, so the host
 * submits one graph instead of (2 + 2 micro) launches per heatbath step;
 * bit-identical results.  Returns when done.  No reference counterpart. */
#define GF_SA_GRAPH 0x10
gf_status gf_run_sa(const gf_lattice *lat, void *U, void *Udw, int gauge, float tmax, float tmin, int steps,
                    int micro, int reproject, int mode, int seed, uint32_t *counter, int *temperatures,
                    void *stream);

/* OR to theta < precision with the quality check fused into the sweep
 * kernels: theta of the odd sites is formed after the odd update, theta of
 * the even sites before the next even update; both write per-block partials
 * that a one-block check kernel reduces in a fixed order.  It sets a device
 * flag on convergence that makes every later sweep of the run exit at once.
 * The host reads the flag once per 32 sweeps.  A 4D lattice with
 * gauge = GF_COULOMB fixes every timeslice at once (spatial links enter the
 * functional; the time links are transformed from both ends) until the
 * largest timeslice theta meets the precision.  The 4D checkerboard takes
 * the parity of t+x+y+z, so an odd timeslice has its odd sites updated
 * first, where the reference's per-slice loop sweeps every slice even-first:
 * the same maximum, reached along a slightly different path.  Every sweep is
 * checked (the reference's checkprecision = 1); out->sweeps is the index + 1
 * of the sweep whose end state met the precision, out->gff / out->theta are
 * that state's, and the array holds that state, projected (the reference
 * breaks after the check and projects, LandauGaugeFixingSU3_4D.cu:254-271).
 * A check fused into the even kernel sees the state only after that kernel
 * has updated it, so once theta < 2 * precision each check takes the even
 * sites' theta from a read-only pass before the even update, which does not
 * run if the check converged.  Should theta fall from above 2 * precision to
 * below it in one sweep, the run goes on and ends on the next sweep's check
 * (one sweep more than the reference loop).  projectSU3 after sweep i when
 * i % reproject == 0, unless sweep i follows convergence.  Not converged:
 * out->theta / out->gff describe the final state, theta as the check
 * defines it (for Coulomb the largest timeslice theta).  No reference
 * counterpart (the reference
 * checks with a separate kernel and a host copy,
 * gaugefixing/GaugeFixingStats.hxx:186-232). */
gf_status gf_run_fast(const gf_lattice *lat, void *U, int gauge, float omega, float precision, int maxiter,
                      int reproject, gf_run_result *out, void *stream);

/* gf_run_fast without any allocation, for callers that cannot allocate
 * device or host memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * caller passes the links and the result as for gf_run_fast; the run's device
 * state and its per-block theta/gff partials live in a fixed array of the
 * library (4 MiB + 256 bytes on each GPU it runs on, kept until the process
 * ends), so lattices of at most 131072 64-site blocks per half-sweep
 * (16,777,216 sites, 64^4) are accepted, larger ones rejected with
 * GF_ERR_BAD_SHAPE.  It sweeps the caller's links in place (no tiled copy)
 * and returns the same sweeps, theta, gff and links as gf_run_fast.
 * out->ms_total is host wall time (no events are created).  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams
 * or threads) race and must not be made. */
gf_status gf_run_fast_noalloc(const gf_lattice *lat, void *U, int gauge, float omega, float precision, int maxiter,
                              int reproject, gf_run_result *out, void *stream);

/* Average duration of one sweep kernel as the drivers launch it, from HIP
 * events around nsweeps back-to-back OR sweeps on `stream` (after
 * max(2, nsweeps / 2) warm-up sweeps).  variant 0: the link sweep in the precision of lat with
 * the fused check (gf_run_fast and the FP64 phase check every sweep); FP32
 * only (lat = FP32, 12-real layout), without the check (the FP32 phases check
 * every check32 sweeps): 1 FP32 links with gauge accumulation, 2 / 3 the
 * gauge-transformation sweep with FP32 / FP64 arithmetic (U is then the
 * read-only link copy), 4 the same as 2 on the links packed into 24-bit
 * fixed point; 5 the drop-in quality kernel (gf_stats_quality's per-site
 * pass, nsweeps launches, any dtype and layout; the time is per launch);
 * 6 as 4, each sweep run as the mixed run's FP32 phase runs a sweep without
 * a check (a wavefront over timeslices where gf_set_wavefront's choice is
 * one), the time per half of a sweep; 7 (lat FP64) the FP64
 * gauge-transformation sweep with the fused check, on the tiled copy of the
 * links held read only (the FP64 phase of gf_run_mixed's gauge modes).
 * U is modified by variants 0 and 1.  Any gauge but Coulomb on a single
 * timeslice. */
gf_status gf_time_sweeps(const gf_lattice *lat, void *U, int gauge, int nsweeps, int variant,
                         float *ms_per_kernel, void *stream);

/* The gauge-transformation sweep (the FP32 phase of gf_run_mixed's gauge
 * modes, gf_run_sa modes 1 and 2) runs each sweep without a check as a
 * wavefront over timeslices when that is faster: launch k updates the even
 * sites of a slab of timeslices and the odd sites of the slab before it,
 * shifted back by one timeslice, so a timeslice's links are read from HBM
 * once per sweep instead of once per half-sweep (the second read comes from
 * the Infinity Cache).  Every site sees the neighbour values of the
 * two-kernel sweep: results are bit-identical either way.  The slab is
 * chosen per lattice shape, gauge and mode by timing a few sweeps with each
 * candidate on the first run of that kind in the process.
 * GF_WAVEFRONT_PIPELINED stands for the pipelined sweep, one launch per
 * sweep: each XCD takes Nt/8 timeslices, and the odd update of a block of 64
 * sites follows the even updates of its neighbours within the same launch
 * (per-block completion flags), so that the second reads of its links and
 * neighbours come from a cache instead of HBM.  Landau, Coulomb (every
 * timeslice at once) and U1xU1; Nt a multiple of 8, a timeslice's sites of a
 * parity a whole number of 64-site blocks; bit-identical results as well.
 * gf_set_wavefront: -1 choose by timing (default), 0 never, n > 0 slabs of
 * n timeslices (or GF_WAVEFRONT_PIPELINED); *previous (optional) receives the
 * old setting.
 * gf_wavefront_choice: the slab in use for gf_run_mixed of lat with `gauge`
 * and gauge-sweep `mode` (0 two kernels, -1 not chosen yet).
 * No reference counterpart. */
#define GF_WAVEFRONT_PIPELINED 1000000
gf_status gf_set_wavefront(int slab, int *previous);
gf_status gf_wavefront_choice(const gf_lattice *lat, int gauge, int mode, int *slab);
/* Pipelined sweeps (GF_WAVEFRONT_PIPELINED) this process has launched, so a
 * caller can see which path its sweeps took.  gf_set_pipelined_lag: the extra
 * even blocks each XCD updates before it starts the odd block they are ahead
 * of (-1 chosen by the timing with the slab, default; n >= 0 fixed): more
 * lets the even updates an odd block waits for finish first, less keeps that
 * block's links in the L2.  Results do not depend on it.  No reference
 * counterpart. */
gf_status gf_pipelined_launches(long long *n);
gf_status gf_set_pipelined_lag(int lag, int *previous);

/* The gauge-transformation sweep's block order across the eight XCDs (each
 * with its own L2): each XCD takes a contiguous range of timeslices and walks
 * it with t innermost, so that a block's t neighbours were read by the block
 * before it.  Walking the range in passes of n timeslices brings a block's x
 * neighbours closer in that sequence, at the cost of the t neighbours across
 * a pass boundary.  The results do not depend on the order (the sites of a
 * parity are independent); the time does.  Chosen per lattice shape, gauge
 * and mode by timing, together with the wavefront slab.
 * The drop-in quality kernel (gf_stats_quality, gf_quality_per_site) walks
 * the XCDs' timeslices the same way, and its order is tuned on its first
 * launch for a kind of lattice (timing launches read the links and write
 * nothing), under the same setting.
 * gf_set_block_order: -1 choose by timing (default), 0 one pass, n > 0 passes
 * of n timeslices (where n divides the XCD's range; otherwise one pass);
 * *previous (optional) receives the old setting.
 * gf_block_order_choice: the order in use for gf_run_mixed of lat with
 * `gauge` and gauge-sweep `mode` (-1 not chosen yet).
 * No reference counterpart. */
gf_status gf_set_block_order(int tpass, int *previous);
gf_status gf_block_order_choice(const gf_lattice *lat, int gauge, int mode, int *tpass);

/* Mixed precision.  U (FP64, layout of lat) holds the configuration U0; an
 * FP32 phase accumulates each site's gauge transformation g(x) until
 * theta < switch_precision (checked every check32 sweeps; the phase also ends
 * at the rounding floor of single precision: with theta below 1e-9, 8
 * checks in a row that do not lower its best theta by 10%; for MAG, with
 * theta below 1e-5, a window of 50 checks whose mean log(theta) drops by less
 * than log(1.1) from the window before, and by less than log(1.001) or than
 * a quarter of that window's own drop, so that slow but steady convergence
 * goes on); then
 * U1 = G(x) U0_mu(x) G(x+mu)^dagger in FP64 with G the FP64 projection of g,
 * so FP32 rounding never alters the gauge orbit, and fused-check FP64 sweeps
 * continue to theta < precision, checked every sweep once theta, falling at
 * the rate the checks show, is within 160 sweeps of the precision, and every
 * 16 sweeps before that (the same stopping sweep, fewer checks); the array then holds
 * the state that met it, as in gf_run_fast (with switch_precision <
 * 2 * precision the FP64 checks are made before the even update from the
 * first sweep on).  In the gauge-transformation modes (and GF_MIXED_FP64)
 * the FP64 phase sweeps a second gauge transformation g64 (FP64, from 1) on
 * U1 held read only, the same OR updates with 1056 instead of 1536 bytes per
 * site-update, and writes g64 U1 g64^dagger back, projected; it takes lattices
 * up to 22,369,600 sites, 68^4 (the link mode, and larger lattices, sweep the FP64
 * links themselves, reprojected every `reproject` sweeps).
 *   mode GF_MIXED_LINKS: a compact FP32 copy of the links is swept and g
 *     updated alongside; every `refresh` sweeps the FP32 links are rebuilt
 *     as G U0 G^dagger (reprojected every `reproject` sweeps in between).
 *   mode GF_MIXED_GAUGE32 / GAUGE64: the FP32 copy of U0 is read only and
 *     the sweep updates g alone (links formed as g U0 g^dagger on the fly),
 *     in FP32 or FP64 arithmetic; every gauge (Coulomb on a 4D lattice:
 *     every timeslice at once; MAG keeps each site's eight links in registers
 *     and applies every SU(2) step to them).  refresh and reproject are not
 *     used in this phase.
 *   mode GF_MIXED_GAUGE24: as GAUGE32 with U0 held in 24-bit fixed point
 *     (each real as round(v (2^23 - 1)), four to three 32-bit words): 36
 *     instead of 48 bytes per link, a quantisation error of at most 2^-24
 *     per real, the size of an FP32 rounding error of an entry near 1.
 *   In GAUGE32 and GAUGE24, an FP32 phase that stagnates above
 *   switch_precision continues on the same g and links with FP64
 *   arithmetic (the GAUGE64 kernel, whose floor is lower) until
 *   switch_precision or stagnation again; out->sweeps_fp32 counts both.
 *   mode GF_MIXED_FP64: no FP32 phase (switch_precision, refresh and check32
 *     unused): the FP64 phase alone, on the caller's links -- an all-FP64
 *     run, on the gauge-transformation sweep where the volume allows.
 * The gauge-transformation modes take at most 44,739,200 sites (80^4 fits;
 * GF_ERR_BAD_SHAPE beyond), GF_MIXED_LINKS any volume below 2^31.
 * No reference counterpart (the reference's DPUPDATES is a compile-time
 * FP32/FP64 mix, gaugefixing/apps/Makefile:46-48). */
gf_status gf_run_mixed(const gf_lattice *lat, void *U, int gauge, float omega, float switch_precision,
                       float precision, int maxiter, int mode, int refresh, int check32, int reproject,
                       gf_run_result *out, void *stream);

/* Batches of independent lattices: gauge copies of one configuration,
 * lattices of several sizes (the reference's apps fix copies one after
 * another, LandauGaugeFixingSU3_4D.cu:184-286; there is no batch call in the
 * reference).  Problem k is the 4D lattice lats[k] with its configuration
 * U[k] (a device buffer of its own); its result is that of the single-lattice
 * call with the same arguments.  Lattices whose half-sweep has fewer than
 * 4096 blocks of 64 sites run concurrently, up to max_concurrent at a time,
 * each on a host thread and a stream of its own; larger ones one after
 * another.  Every lattice is checked before any runs; a failure is returned
 * as "batch problem k: <message>" for the first failing k.  The call returns
 * when every problem is done.  streams: null, or max_concurrent streams of
 * the caller's (host thread t runs its problems on streams[t], the large
 * problems run on streams[0]); with null each problem gets a stream created
 * and destroyed by the call.  Streams created early in a process -- before
 * its first device allocation -- have run concurrent copies up to 35% faster
 * than streams created later, with the runtime's 4 hardware queues per
 * process: a caller that batches repeatedly keeps a set of its own. */
gf_status gf_batch_run_mixed(int n, const gf_lattice *lats, void *const *U, int gauge, float omega,
                             float switch_precision, float precision, int maxiter, int mode, int refresh,
                             int check32, int reproject, int max_concurrent, void *const *streams,
                             gf_run_result *out);
/* gf_run_sa for each problem, with seeds[k] and counters[k] (in/out: the
 * launch counter after the schedule); temperatures (optional) receives each
 * problem's number of temperatures; streams as for gf_batch_run_mixed.
 * mode | GF_SA_ONE_LAUNCH (modes 1 and 2; Landau, Coulomb on 4D lattices
 * and U1xU1): instead of a launch per lattice, each half-sweep of several
 * lattices is one launch whose blocks are the lattices' blocks one after
 * another (up to 16 lattices of any sizes).  The lattices are packed into
 * such groups largest first, each into the first group with room under the
 * limit of gf_set_batch_launch_blocks; the groups anneal one after another on
 * streams[0] (or a stream of the call's own), each group's schedule captured
 * into one graph when GF_SA_GRAPH is set too.  Every problem's result and
 * counter are those of gf_run_sa, bit for bit.  max_concurrent is not used.
 * No reference counterpart. */
#define GF_SA_ONE_LAUNCH 0x20
gf_status gf_batch_run_sa(int n, const gf_lattice *lats, void *const *U, int gauge, float tmax, float tmin,
                          int steps, int micro, int reproject, int mode, const int *seeds, uint32_t *counters,
                          int *temperatures, int max_concurrent, void *const *streams);

/* The most blocks of 64 sites of one parity that a GF_SA_ONE_LAUNCH launch
 * gathers from several lattices (default 8192, the blocks of a 32^4 lattice,
 * whose working set the 256 MB Infinity Cache holds in the annealing's
 * 24-bit form; a lattice with more blocks forms a launch of its own); 0: no
 * limit but the 16 lattices.  *previous (optional) receives the old limit.
 * No reference counterpart. */
gf_status gf_set_batch_launch_blocks(int blocks, int *previous);

/* --------------------------------------------------- synthetic inputs
 * Host-side generator of the benchmark configurations (not a reference
 * interface): links U = exp(i sum_a c_a lambda_a / 2), c_a ~ N(0, sigma^2)
 * (Philox-keyed by seed and link), then, when haar_seed != 0, a Haar-random
 * SU(3) gauge transformation g(x) (Philox-keyed by haar_seed and site):
 * U_mu(x) <- g(x) U_mu(x) g(x+mu)^dagger.  Writes all 18 reals in the
 * GpuPattern layout into host memory U (V*72 doubles).  sigma = 0 gives unit
 * links (a pure gauge after the transformation). */
gf_status gf_make_config(const int dims[4], double sigma, uint32_t seed, uint32_t haar_seed, double *U);

#ifdef __cplusplus
}
#endif

#endif /* SU3GF_H_ */
