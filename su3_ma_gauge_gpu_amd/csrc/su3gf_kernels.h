// su3gf_kernels.h -- kernel argument blocks and launchers shared by the
// kernel translation unit and the C ABI implementation.
#pragma once
#include "su3gf_device.h"
#include "../../include/su3gf.h"

namespace su3gf {

constexpr int kBlock = 256;  // 4 waves of 64 lanes = 64 sites per block

template <typename R> struct SweepArgs {
    R *U;          // links (timeslice t for Coulomb)
    R *Udw;        // Coulomb: timeslice t-1 buffer; otherwise == U
    Geom g;
    int parity;
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
};

template <typename R> hipError_t launch_sweep(int stride, int gauge, int algo, const SweepArgs<R> &a, hipStream_t s);
template <typename R> hipError_t launch_project(int stride, R *U, int64_t V, hipStream_t s);
template <typename R> hipError_t launch_set_hot(R *U, int64_t V, uint32_t seed, uint32_t counter, hipStream_t s);
template <typename R> hipError_t launch_quality(int stride, int gauge, const QualityArgs<R> &a, hipStream_t s);
hipError_t launch_convert(int ti, int si, const void *src, int to, int so, void *dst, int64_t V, int reun,
                          hipStream_t s);
hipError_t launch_reduce_final(const double *partial, int64_t n, int crit_max, double *out, hipStream_t s);
int64_t quality_blocks(int64_t V);

}  // namespace su3gf
