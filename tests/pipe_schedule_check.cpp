// Checks the pipelined sweep's schedule (su3gf_pipe.h) on the CPU, site by
// site: every queue's tickets give each of its even and odd items exactly
// once; every even block that holds a neighbour of an odd site lies within
// the kernel's wait set of that site's odd item (slices t-1..t+1, blocks
// q-R..q+R cyclically) and comes at an earlier ticket than the odd item (in
// its own queue's ticket order, the queues sharing the schedule's shape);
// for several extra lags (slack), down to none.
// Prints "ok <shape>" per lattice, or the first violation and exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../su3_ma_gauge_gpu_amd/csrc/su3gf_pipe.h"

using namespace su3gf;

static int fail(const char *what, const int N[4], long a, long b)
{
    std::printf("FAIL %s at %dx%dx%dx%d: %ld %ld\n", what, N[0], N[1], N[2], N[3], a, b);
    return 1;
}

static int check(const int N[4], int slack)
{
    PipePlan p;
    if (!pipe_plan(N, p, slack)) {
        std::printf("skip %dx%dx%dx%d\n", N[0], N[1], N[2], N[3]);
        return 0;
    }
    // tickets -> items: a bijection per queue, inverse of pipe_ticket_*
    std::vector<int> seen(2 * p.n, 0);
    for (int64_t k = 0; k < 2 * p.n; k++) {
        int odd;
        int64_t idx;
        pipe_item(p, k, odd, idx);
        if (idx < 0 || idx >= p.n) return fail("item index out of range", N, (long)k, (long)idx);
        if (seen[odd * p.n + idx]++) return fail("item given twice", N, (long)k, (long)idx);
        const int64_t back = odd ? pipe_ticket_odd(p, idx) : pipe_ticket_even(p, idx);
        if (back != k) return fail("ticket inverse", N, (long)k, (long)back);
    }
    const int64_t slice = (int64_t)N[1] * N[2] * N[3], half = slice / 2;
    auto block_of = [&](const int c[4], int64_t &t, int64_t &q) {
        const int64_t lex = ((((int64_t)c[0] * N[1] + c[1]) * N[2] + c[2]) * N[3] + c[3]);
        const int64_t pidx = lex >> 1;  // index within the parity
        t = pidx / half;
        q = (pidx - t * half) / 64;
    };
    int c[4];
    for (c[0] = 0; c[0] < N[0]; c[0]++)
        for (c[1] = 0; c[1] < N[1]; c[1]++)
            for (c[2] = 0; c[2] < N[2]; c[2]++)
                for (c[3] = 0; c[3] < N[3]; c[3]++) {
                    if (((c[0] + c[1] + c[2] + c[3]) & 1) == 0) continue;  // odd sites
                    int64_t t, q;
                    block_of(c, t, q);
                    if (t != c[0]) return fail("timeslice of a block", N, (long)t, (long)c[0]);
                    const int r = (int)(t / p.T);
                    const int64_t tk = pipe_ticket_odd(p, q * p.T + (t - (int64_t)r * p.T));
                    for (int mu = 0; mu < 4; mu++)
                        for (int s = -1; s <= 1; s += 2) {
                            int d[4] = {c[0], c[1], c[2], c[3]};
                            d[mu] = (d[mu] + s + N[mu]) % N[mu];
                            int64_t t2, q2;
                            block_of(d, t2, q2);
                            // inside the wait set: slices t-1..t+1, blocks within R cyclically
                            const int64_t dt = (t2 - t + N[0]) % N[0];
                            if (!(dt == 0 || dt == 1 || dt == N[0] - 1))
                                return fail("neighbour slice outside t-1..t+1", N, (long)t, (long)t2);
                            const int64_t dq = (q2 - q + p.Q) % p.Q;
                            if (!(dq <= p.R || p.Q - dq <= p.R))
                                return fail("neighbour block beyond R", N, (long)q, (long)q2);
                            // that even item at an earlier ticket of its queue
                            const int r2 = (int)(t2 / p.T);
                            const int64_t tk2 = pipe_ticket_even(p, q2 * p.T + (t2 - (int64_t)r2 * p.T));
                            if (!(tk2 < tk)) return fail("dependency not earlier", N, (long)tk, (long)tk2);
                        }
                }
    std::printf("ok %dx%dx%dx%d slack=%d T=%d Q=%d R=%d L=%ld\n", N[0], N[1], N[2], N[3], slack, p.T, p.Q, p.R,
                (long)p.L);
    return 0;
}

int main(int argc, char **argv)
{
    if ((argc - 1) % 4 != 0 || argc < 5) {
        std::printf("usage: %s Nt Nx Ny Nz [Nt Nx Ny Nz ...]\n", argv[0]);
        return 2;
    }
    for (int i = 1; i + 3 < argc; i += 4) {
        const int N[4] = {std::atoi(argv[i]), std::atoi(argv[i + 1]), std::atoi(argv[i + 2]), std::atoi(argv[i + 3])};
        for (int slack : {0, 1, 37, 128, 100000})
            if (check(N, slack)) return 1;
    }
    return 0;
}
