// su3gf_kernels.h -- kernel argument blocks and launchers shared by the
// kernel translation units and the C ABI implementation.
#pragma once
#include "su3gf_device.h"
#include "su3gf_pipe.h"
#include "../../include/su3gf.h"

namespace su3gf {

constexpr int kBlock = 256;  // 4 waves of 64 lanes: one wave per direction, 64 sites per block

// Device-resident state of a fused-theta driver run.  Written by the check
// kernel that follows each even sweep kernel; read by the host after a chunk.
struct DevState {
    int done;             // 1: theta < tol reached; 2: theta stagnated
    int converged_sweep;  // index of the sweep whose end state met tol (or stagnated)
    int checked_sweep;    // index of the sweep whose end state `theta` describes
    int checks;           // number of completed checks
    int stall;            // rule 0: checks in a row without a 10% gain; 1: checks in the current window
    int stall_limit;      // that many checks stagnate (rule 0) or make a window (rule 1); 0: no stagnation exit
    double stall_below;   // stagnation is only tested for theta below this
    int stall_rule;       // 0: the best theta's gain; 1: windows' mean log(theta), for slow convergence
    double best;          // rule 0: the best theta, lowered in steps of 10%
    int windows;          // rule 1: windows completed below stall_below
    double win_logsum;    // sum of log(theta) over the current window
    double win_mean;      // mean of log(theta) over the previous window
    double win_drop;      // the previous window's decrease of that mean
    double tol;           // theta threshold
    double theta_scale;   // 1 / (Nc * V)
    double gff_scale;     // 1 / (Nc * Ndim * V)
    double theta;         // theta of sweep `checked_sweep` (prefactors applied)
    double gff;           // gff of the same state
    // A phase that must end on the state its last check measured (the FP64
    // phases, exact_end = 1) cannot stop on a theta formed by an even kernel
    // that has already updated.  Once theta < near_below (a small multiple of
    // tol), `near` is set: each check is then made by a read-only theta pass
    // before the even update, which the check can still cancel.
    int exact_end;
    int near;
    double near_below;
};

// k_fused_check modes: which theta the check reads
enum CheckMode {
    CHECK_FUSED = 0,    // the even sweep kernel's (already updated); acts unless done
    CHECK_PRE = 1,      // the read-only pass before the even update; acts only if `near`
    CHECK_FUSED_IF_FAR = 2,  // as CHECK_FUSED, but only while `near` is not set
};

template <typename R> struct SweepArgs {
    R *U;          // links (timeslice t for Coulomb)
    R *Udw;        // Coulomb: timeslice t-1 buffer; otherwise == U
    float *gacc;   // gauge transformation accumulated by the FP32 phase (or null)
    DevState *state;   // fused-theta state (or null)
    double *partial;   // per-block (theta, gff) partials of this kernel (fused check)
    Geom g;
    int parity;
    int sweep;     // sweep index (fused check)
    float param;   // OR/SR parameter or SA temperature
    uint32_t seed, counter;
    int64_t nsb;   // sites per block of the reference launch (RNG key)
};

template <typename R> struct QualityArgs {
    const R *U;
    Geom g;
    double *dGff, *dA;  // optional per-site outputs (V doubles each)
    double *partial;    // optional per-block (theta, gff) partials
    int crit_max;
    // block order, set by the launcher: a block takes the sites of both
    // parities with parity-split index in [128 L, 128 L + 128) (256
    // consecutive sites of the lexicographic order) for its logical block L
    int remap;              // 0 plain, 1 one contiguous eighth per XCD, 2 the same with t innermost
    int64_t nblocks;        // logical blocks (quality_blocks)
    int64_t slice_blocks;   // remap 2: logical blocks per timeslice
    int slices_per_xcd;     // remap 2
    int tpass;              // remap 2: each XCD's timeslices walked in passes of tpass (0: one pass)
    FastDiv tpass_div, pass_div;  // division by tpass and by tpass * slice_blocks
};


// Single-precision sweep on the gauge transformation (su3gf_gauge.hip): U0 is
// read only, g(x) is updated; the links are g(x) U0_mu(x) g(x+mu)^dagger.
struct GaugeSweepArgs {
    const float *U0;   // compact FP32 links (rows 0,1, stride 12), or the 24-bit words (links24)
    float *g;          // gauge transformation, rows 0,1 (stride 12, one field)
    DevState *state;   // fused-theta state (or null)
    double *partial;   // per-block (theta, gff) partials (fused check)
    Geom geo;
    int parity;
    float param;       // OR parameter or SA temperature
    int remap;         // block order: 0 plain, else chosen by the launcher (XCD-aware)
    int slices_per_xcd;  // set by the launcher for remap = 2
    int64_t slice_blocks;
    int tpass;           // remap = 2: each XCD walks its timeslices in passes of tpass (a divisor of
                         // slices_per_xcd; 0 = all of them in one pass), t innermost within a pass
    FastDiv tpass_div, pass_div;  // division by tpass and by tpass * slice_blocks
    uint32_t seed, counter;  // SA / random transformation (Philox key and counter)
    int64_t nsb;       // sites per block of the reference launch (RNG key)
    int links24;       // U0 in the 24-bit fixed-point form of launch_pack24 (OR, SA, microcanonical)
    int links64;       // U0 and g in FP64 tiles (kTiled12, 12 doubles per site): the mixed driver's
                       // FP64 phase (OR with FP64 arithmetic; theta 0-3, 3 the read-only pass)
    // Wavefront launch (wf = 1, set by launch_gsweep_wavefront): one launch
    // sweeps the even sites of logical blocks [wf_even0, wf_even0 + wf_even_n)
    // with its first wf_even_grid blocks and the odd sites of
    // [wf_odd0, wf_odd0 + wf_odd_n), then of [wf_odd1, wf_odd1 + wf_odd1_n),
    // with the rest; each range is dealt to the XCDs in contiguous eighths
    // (grids padded to multiples of 8, padding blocks idle).  Odd sites draw
    // their random numbers with counter_odd.
    int wf;
    int64_t wf_even0, wf_even_n, wf_even_grid;
    int64_t wf_odd0, wf_odd_n, wf_odd1, wf_odd1_n;
    uint32_t counter_odd;
    // Pipelined sweep (launch_gsweep_pipe): the lattice's synchronisation
    // words (allocated and zeroed by the caller, pipe_sync_words) and the
    // host's count of launches on them; the plan and its division by T, set
    // by the launcher
    PipeSync *pipe;
    uint32_t *pipe_count;  // host memory: the launches made on `pipe` so far
    uint32_t pipe_launch;  // set by the launcher: this launch's number
    int pipe_slack;        // extra even items before the first odd one (pipe_plan's slack)
    PipePlan plan;
    FastDiv pipe_tdiv;
};
// algo: GF_OR, GF_MICRO, GF_SA or GF_RANDOM; theta (OR only): 0 none,
// 1 pre-update (even kernel), 2 post-update (odd kernel), 3 the read-only
// theta pass (links64 only, while state->near is set); fp64_arith: compute
// in FP64 on the FP32 data (required with links64)
hipError_t launch_gsweep(int gauge, int algo, int theta, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s);
// One whole sweep (even then odd sites, no fused check) of a 4D lattice as a
// wavefront over timeslices: launch k updates the even sites of timeslices
// [kS, (k+1)S) and the odd sites of [(k-1)S - 1, kS - 1) (slab = S; slice 0's
// and the last slices' odd sites in a final launch), so every site sees the
// neighbour values of the two-kernel sweep -- the results are bit-identical
// -- while a timeslice's links, read by its even update, are read again by
// its odd update one launch later, from the Infinity Cache instead of HBM.
// The odd sites use counter a.counter + 1 (the two-kernel sweep's odd
// launch).  Returns hipErrorNotSupported where the wavefront does not apply
// (fewer than 4 timeslices, or a timeslice's sites of a parity not whole
// 64-site blocks); the caller then launches the two kernels.
hipError_t launch_gsweep_wavefront(int gauge, int algo, bool fp64_arith, const GaugeSweepArgs &a, int slab,
                                   hipStream_t s);
// The wavefront slab that stands for the pipelined sweep in the tuner's and
// the drivers' choice (launch_gsweep_wavefront passes it on to
// launch_gsweep_pipe).
constexpr int kPipelinedSlab = GF_WAVEFRONT_PIPELINED;
// One whole sweep (even then odd sites, no fused check) in one launch: the
// pipelined schedule of su3gf_pipe.h.  Bit-identical to the two-kernel sweep
// (odd sites draw their random numbers with counter a.counter + 1).  Landau
// and U1xU1 on FP32 or 24-bit links, OR, microcanonical and heatbath; needs
// a.pipe.  Returns hipErrorNotSupported where it does not apply (the caller
// then launches the two kernels).
hipError_t launch_gsweep_pipe(int gauge, int algo, bool fp64_arith, const GaugeSweepArgs &a, hipStream_t s);
// Up to kMultiMax lattices swept in one launch (launch_gsweep_multi): each
// lattice's read-only links U0 (FP32 rows 0,1 or the 24-bit words) and its g,
// with the geometry and the RNG key of its single-lattice sweep.
constexpr int kMultiMax = 16;
struct MultiItem {
    const float *U0;
    float *g;
    Geom geo;
    uint32_t seed, counter;  // Philox key and the counter before counter_add
};
struct MultiSweepArgs {
    MultiItem item[kMultiMax];
    int32_t first[kMultiMax + 1];  // first block of each lattice in the concatenation; first[n] = total
    int n;
    int parity;
    float param;           // SA temperature (shared by the lattices)
    uint32_t counter_add;  // added to every item's counter (then + parity)
    int32_t per_xcd;       // set by the launcher: blocks per XCD, ceil(total / 8)
};
// One parity of every lattice of m in one launch, each block running the
// two-kernel sweep's code on the block of its lattice: the results are those
// of launch_gsweep on each lattice with counter item.counter + counter_add +
// parity.  Landau, Coulomb (all timeslices at once) and U1xU1; SA and
// microcanonical; FP32 arithmetic on FP32 (links24 = 0) or 24-bit links.
hipError_t launch_gsweep_multi(int gauge, int algo, int links24, MultiSweepArgs &m, hipStream_t s);
// whether the pipelined sweep applies to a lattice and gauge (the launcher's
// shape conditions), so that callers allocate its words only then
bool gsweep_pipe_applies(const Geom &geo, int gauge, bool links64);
// pipelined sweeps launched by this process
long long pipe_launches();
// words of the 24-bit link copy: 9 per link, in whole tiles of 64 sites
__host__ __device__ inline int64_t links24_words(int64_t V) { return 4 * 9 * 64 * ((V + 63) / 64); }
// links of dtype ti and layout si (FP64 18 or 12, FP32 12) -> the 24-bit
// fixed-point rows 0,1 (9 words per link)
hipError_t launch_pack24(int ti, int si, const void *U, uint32_t *P, int64_t V, hipStream_t s);
// the 24-bit rows 0,1 -> FP32 rows 0,1 (stride 12), decoded as the sweep kernels decode them
hipError_t launch_unpack24(const uint32_t *P, float *U32, int64_t V, hipStream_t s);

// added to an update rule's number: the rule computed in double on FP32
// links (GF_UPDATE_DP); template argument of the exact sweep kernels
constexpr int kDpUpdate = 8;

// exact translation unit (reference order, no FMA contraction)
namespace exact {
// algo: a gf_algo, or GF_OR / GF_MICRO / GF_SA + kDpUpdate (R = float)
template <typename R> hipError_t launch_sweep(int stride, int gauge, int algo, const SweepArgs<R> &a, hipStream_t s);
// st: if given, the kernel does nothing once st->done is set (a driver's
// reprojection enqueued after the sweep that converged)
template <typename R>
hipError_t launch_project(int stride, int nfields, R *U, int64_t V, hipStream_t s, const DevState *st = nullptr);
template <typename R> hipError_t launch_set_hot(R *U, int64_t V, uint32_t seed, uint32_t counter, hipStream_t s);
template <typename R> hipError_t launch_quality(int stride, int gauge, const QualityArgs<R> &a, hipStream_t s);
int64_t quality_blocks(int64_t V);
}  // namespace exact

// fast translation unit (contracted, fused theta, gauge accumulation)
namespace fast {
// theta: 0 none, 1 pre-update (even kernel), 2 post-update (odd kernel),
// 3 read-only theta pass (OR only; runs only while state->near is set)
template <typename R>
hipError_t launch_sweep(int stride, int gauge, int algo, int theta, bool gacc, const SweepArgs<R> &a, hipStream_t s);
}  // namespace fast

// utility kernels (su3gf_kernels.hip)
hipError_t launch_convert(int ti, int si, const void *src, int to, int so, void *dst, int64_t V, int reun,
                          hipStream_t s);
hipError_t launch_reduce_final(const double *partial, int64_t n, int crit_max, double *out, hipStream_t s);
hipError_t launch_gauge_identity(float *g, int64_t V, hipStream_t s);
hipError_t launch_project_gauge(float *g, int64_t V, hipStream_t s);
// out_mu(x) = G(x) U0_mu(x) G(x+mu)^dagger with G = FP64 projection of g(x);
// U0 FP64 in layout s0; out of type `to` in layout so (may alias U0 if same type/layout)
hipError_t launch_gauge_apply(const double *U0, int s0, const float *g, void *out, int to, int so, const Geom &geo,
                              hipStream_t s);
// the FP64 phase's g (12 doubles per site in 64-site tiles) set to 1, and
// out = G U1 G^dagger with U1 in kTiled12 (FP64), out FP64 in layout so
// (18 or 12), reprojected when reun
hipError_t launch_gauge_identity64(double *g, int64_t V, hipStream_t s);
hipError_t launch_gauge_apply64(const double *U1, const double *g, double *out, int so, const Geom &geo, int reun,
                                hipStream_t s);
hipError_t launch_state_init(DevState *st, double tol, double theta_scale, double gff_scale, int stall_limit,
                             double stall_below, int stall_rule, int exact_end, double near_below, int near,
                             hipStream_t s);
// theta of the state after sweep `sweep - 1` from the even partials of sweep
// `sweep` (pre-update) and the odd partials of sweep `sweep - 1` (post-update);
// nseg > 1: the blocks form nseg independent problems (timeslices), theta is
// the largest of theirs; mode: a CheckMode
hipError_t launch_fused_check(DevState *st, const double *even, const double *odd, int64_t nblocks, int nseg,
                              int sweep, int mode, hipStream_t s);

}  // namespace su3gf
