// su3gf_fast.hip -- sweep kernels of the fused-check drivers (FMA
// contraction, hardware rsqrt): theta/gff of the lattice are accumulated inside
// the sweep (odd sites after the odd update, even sites before the next even
// update, finished by the last block of each kernel), and the FP32 phase of the
// mixed-precision driver also accumulates each site's gauge transformation.
#include "su3gf_sweep.h"

namespace su3gf {
namespace fast {

template <typename R, int G, int A, int S, int T, bool GA>
static hipError_t launch_t(const SweepArgs<R> &a, hipStream_t s)
{
    hipLaunchKernelGGL((k_sweep<R, G, A, S, false, T, GA>), dim3((unsigned)sweep_grid(a.g)), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

template <typename R, int G, int A, int S, bool GA>
static hipError_t launch_theta(int theta, const SweepArgs<R> &a, hipStream_t s)
{
    if constexpr (A == GF_OR || A == GF_SR) {
        switch (theta) {
        case 1: return launch_t<R, G, A, S, 1, GA>(a, s);
        case 2: return launch_t<R, G, A, S, 2, GA>(a, s);
        case 3:
            // the read-only pass: one instantiation per gauge and layout suffices
            if constexpr (A == GF_OR && !GA) {
                if (!a.state) return hipErrorInvalidValue;
                return launch_t<R, G, A, S, 3, false>(a, s);
            } else {
                return hipErrorInvalidValue;
            }
        default: return launch_t<R, G, A, S, 0, GA>(a, s);
        }
    }
    else {
        if (theta != 0) return hipErrorInvalidValue;
        return launch_t<R, G, A, S, 0, GA>(a, s);
    }
}

template <typename R, int G, int S, bool GA>
static hipError_t launch_a(int algo, int theta, const SweepArgs<R> &a, hipStream_t s)
{
    switch (algo) {
    case GF_OR: return launch_theta<R, G, GF_OR, S, GA>(theta, a, s);
    case GF_MICRO: return launch_theta<R, G, GF_MICRO, S, GA>(theta, a, s);
    case GF_SA: return launch_theta<R, G, GF_SA, S, GA>(theta, a, s);
    case GF_SR:
        if constexpr (G == GF_MAG) return launch_theta<R, G, GF_SR, S, GA>(theta, a, s);
        else return hipErrorInvalidValue;
    default: return launch_theta<R, G, GF_RANDOM, S, GA>(theta, a, s);
    }
}

template <typename R, int S, bool GA>
static hipError_t launch_g(int gauge, int algo, int theta, const SweepArgs<R> &a, hipStream_t s)
{
    switch (gauge) {
    case GF_LANDAU: return launch_a<R, GF_LANDAU, S, GA>(algo, theta, a, s);
    case GF_COULOMB: return launch_a<R, GF_COULOMB, S, GA>(algo, theta, a, s);
    case GF_MAG: return launch_a<R, GF_MAG, S, GA>(algo, theta, a, s);
    default: return launch_a<R, GF_U1XU1, S, GA>(algo, theta, a, s);
    }
}

template <>
hipError_t launch_sweep<double>(int stride, int gauge, int algo, int theta, bool gacc, const SweepArgs<double> &a,
                                hipStream_t s)
{
    if (gacc) return hipErrorInvalidValue;  // the FP64 phase works on the transformed links directly
    if (stride == kTiled12) return launch_g<double, kTiled12, false>(gauge, algo, theta, a, s);
    if (stride == 18) return launch_g<double, 18, false>(gauge, algo, theta, a, s);
    return launch_g<double, 12, false>(gauge, algo, theta, a, s);
}

template <>
hipError_t launch_sweep<float>(int stride, int gauge, int algo, int theta, bool gacc, const SweepArgs<float> &a,
                               hipStream_t s)
{
    if (stride != 12) return hipErrorInvalidValue;  // FP32 works on the compact layout
    if (gacc) return launch_g<float, 12, true>(gauge, algo, theta, a, s);
    return launch_g<float, 12, false>(gauge, algo, theta, a, s);
}

}  // namespace fast
}  // namespace su3gf
