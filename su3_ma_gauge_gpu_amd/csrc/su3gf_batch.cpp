// su3gf_batch.cpp -- independent lattices fixed in one call: gauge copies of
// one configuration, lattices of several sizes.  Host code on top of the
// single-lattice drivers of include/su3gf.h.
//
// A lattice whose parity sweep has fewer than kFillBlocks blocks of 64 sites
// does not fill the GPU by itself: such problems run concurrently, each on a
// host thread and a stream of its own (up to max_concurrent at a time), so
// that their launches overlap.  Larger ones run one after another, where
// concurrent problems would only evict each other's data from the caches.
// Every problem's result is that of the single-lattice call with the same
// arguments: the problems share nothing.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "su3gf.h"
#include "su3gf_error.h"

namespace {

constexpr long long kFillBlocks = 4096;

struct Outcome {
    gf_status st = GF_OK;
    std::string msg;
};

gf_status check_problems(int n, const gf_lattice *lats, void *const *U, int max_concurrent, void *const *streams)
{
    if (n < 0) return su3gf::set_error(GF_ERR_BAD_ARG, "negative batch size");
    if (streams)
        for (int t = 0; t < std::max(1, max_concurrent); t++)
            if (!streams[t]) return su3gf::set_error(GF_ERR_BAD_ARG, "null stream in the caller's streams");
    if (n > 0 && (!lats || !U)) return su3gf::set_error(GF_ERR_BAD_ARG, "null argument");
    for (int k = 0; k < n; k++) {
        if (gf_lattice_bytes(&lats[k]) == 0) {
            const std::string m = "batch problem " + std::to_string(k) + ": invalid lattice shape or type";
            return su3gf::set_error(GF_ERR_BAD_SHAPE, m.c_str());
        }
        if (!U[k]) {
            const std::string m = "batch problem " + std::to_string(k) + ": null configuration";
            return su3gf::set_error(GF_ERR_BAD_ARG, m.c_str());
        }
    }
    return GF_OK;
}

// job(k, stream) for every k; the first failure (in problem order) is returned.
// streams: null, or max_concurrent caller streams (worker t runs its problems
// on streams[t], the large problems run on streams[0]); with null every
// problem gets a stream of its own, created and destroyed here.
template <typename Job>
gf_status run_batch(int n, const gf_lattice *lats, int max_concurrent, void *const *streams, Job job)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return su3gf::set_error(GF_ERR_NO_DEVICE, "no HIP device");
    std::vector<Outcome> out(n);
    std::vector<int> small, large;
    for (int k = 0; k < n; k++) {
        const gf_lattice &l = lats[k];
        const long long half = (long long)l.dims[0] * l.dims[1] * l.dims[2] * l.dims[3] / 2;
        ((half + 63) / 64 < kFillBlocks ? small : large).push_back(k);
    }
    auto one = [&](int k, int worker) {
        Outcome &o = out[k];
        if (hipSetDevice(dev) != hipSuccess) {  // the device is per host thread
            o.st = GF_ERR_HIP;
            o.msg = "cannot select the caller's device";
            return;
        }
        void *s = streams ? streams[worker] : nullptr;
        if (!streams) {
            o.st = gf_stream_create(&s);
            if (o.st != GF_OK) {
                o.msg = gf_last_error();
                return;
            }
        }
        o.st = job(k, s);
        if (o.st != GF_OK) o.msg = gf_last_error();
        // the annealing of mode 0 returns with its sweeps still queued
        const gf_status sy = gf_stream_synchronize(s);
        if (o.st == GF_OK && sy != GF_OK) {
            o.st = sy;
            o.msg = gf_last_error();
        }
        if (!streams) {
            const gf_status de = gf_stream_destroy(s);
            if (o.st == GF_OK && de != GF_OK) {
                o.st = de;
                o.msg = gf_last_error();
            }
        }
    };
    std::atomic<size_t> next{0};
    const int nthreads = std::min(std::max(1, max_concurrent), (int)small.size());
    std::vector<std::thread> pool;
    try {
        for (int t = 0; t < nthreads; t++)
            pool.emplace_back([&, t] {
                for (size_t i; (i = next++) < small.size();) one(small[i], t);
            });
    } catch (...) {
        // a thread that cannot start leaves its share to the others, or to
        // this thread if none started (no exception may cross the C ABI)
    }
    if (pool.empty())
        for (size_t i; (i = next++) < small.size();) one(small[i], 0);
    for (auto &t : pool) t.join();
    for (int k : large) one(k, 0);
    for (int k = 0; k < n; k++)
        if (out[k].st != GF_OK) {
            const std::string m = "batch problem " + std::to_string(k) + ": " + out[k].msg;
            return su3gf::set_error(out[k].st, m.c_str());
        }
    return GF_OK;
}

}  // namespace

extern "C" gf_status gf_batch_run_mixed(int n, const gf_lattice *lats, void *const *U, int gauge, float omega,
                                        float switch_precision, float precision, int maxiter, int mode, int refresh,
                                        int check32, int reproject, int max_concurrent, void *const *streams,
                                        gf_run_result *out)
{
    const gf_status st = check_problems(n, lats, U, max_concurrent, streams);
    if (st != GF_OK) return st;
    if (n == 0) return GF_OK;
    if (!out) return su3gf::set_error(GF_ERR_BAD_ARG, "null argument");
    return run_batch(n, lats, max_concurrent, streams, [&](int k, void *s) {
        return gf_run_mixed(&lats[k], U[k], gauge, omega, switch_precision, precision, maxiter, mode, refresh,
                            check32, reproject, &out[k], s);
    });
}

extern "C" gf_status gf_batch_run_sa(int n, const gf_lattice *lats, void *const *U, int gauge, float tmax,
                                     float tmin, int steps, int micro, int reproject, int mode, const int *seeds,
                                     uint32_t *counters, int *temperatures, int max_concurrent,
                                     void *const *streams)
{
    const gf_status st = check_problems(n, lats, U, max_concurrent, streams);
    if (st != GF_OK) return st;
    if (n == 0) return GF_OK;
    if (!seeds || !counters) return su3gf::set_error(GF_ERR_BAD_ARG, "null argument");
    if (mode & GF_SA_ONE_LAUNCH)
        return su3gf::batch_sa_one_launch(n, lats, U, gauge, tmax, tmin, steps, micro, mode, seeds, counters,
                                          temperatures, streams ? streams[0] : nullptr);
    return run_batch(n, lats, max_concurrent, streams, [&](int k, void *s) {
        return gf_run_sa(&lats[k], U[k], nullptr, gauge, tmax, tmin, steps, micro, reproject, mode, seeds[k],
                         &counters[k], temperatures ? &temperatures[k] : nullptr, s);
    });
}
