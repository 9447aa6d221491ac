// su3gf_pipe.h -- schedule of the pipelined gauge-transformation sweep.
//
// A whole checkerboard sweep (even sites, then odd sites) runs as one launch.
// The lattice's timeslices are split into 8 queues of T = Nt/8 consecutive
// slices, one queue per XCD (block b serves queue b % 8, and the blocks of a
// launch are dealt to the 8 XCDs round robin).  A queue's work items are its
// 64-site blocks of either parity, indexed idx = q T + (t - t0) (q the block's
// place in its timeslice, 0 <= q < Q; t0 the queue's first slice): blocks of
// neighbouring q are neighbours in space, so walking idx in order keeps the
// XCD on one small region of all T slices.  The queue interleaves the even
// items with the odd ones, the odd item of a block trailing its even item by
// about (R + 1) T items, where R is the reach of a block's neighbours in q.
// An odd item starts when every even item whose sites neighbour its sites has
// finished (per-block flags); the links and the neighbours' g that the odd
// item reads were read by those even items shortly before and still sit in
// the XCD's L2, so a sweep reads the read-only links from HBM about once
// instead of once per parity.
//
// Block b takes ticket b / 8 of queue b % 8.  Every odd item depends only on
// even items at earlier tickets (of its own queue or of the neighbouring
// queues, which share the schedule's shape), so on blocks of lower index, and
// even items depend on nothing: with blocks dispatched in order of their
// index, every item an odd item waits for is already running.  The waits are
// bounded all the same: one that runs out sets an error word, and every later
// wait gives up at once.
//
// This header is plain C++ as well, so that the schedule is checked on the CPU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SU3GF_PIPE_HD __host__ __device__ inline
#else
#define SU3GF_PIPE_HD inline
#endif

namespace su3gf {

constexpr int kPipeQueues = 8;

// Synchronisation words of one lattice's pipelined sweeps (device memory,
// zero when allocated).  The launches are numbered 1, 2, ... by the host; an
// even item sets its block's flag to the launch's number, which is all an odd
// item of the same launch waits for, so nothing needs clearing between
// launches.
struct PipeSync {
    uint32_t err;     // a dependency wait ran out of its spin limit
    uint32_t moved;   // blocks that ran on another XCD than their queue's
    uint32_t pad[30];
    // followed by one flag per even block (logical block t Q + q): the number
    // of the launch that last updated it, bit 31 set if a block on another
    // XCD than its queue's wrote it
};
constexpr uint32_t kPipeMovedBit = 0x80000000u;

// The plan of a lattice {Nt, Nx, Ny, Nz}.
struct PipePlan {
    int T;       // timeslices per queue
    int Q;       // 64-site blocks of one parity per timeslice
    int R;       // a block's neighbours (any direction in space) lie within R blocks of it in q, cyclically
    int64_t n;   // items of each parity per queue: Q T
    int64_t L;   // even items before the first odd one
    int64_t W;   // the odd order starts at idx = W: odd items of q < R come last
};

// R: the parity-split index within a timeslice is half the lexicographic
// index (z fastest), so the x neighbour lies Ny Nz / 2 sites away (cyclically
// within the slice) and y and z neighbours closer; a block then reaches
// blocks up to (63 + Ny Nz / 2) / 64 away.  L = (2 R + 1) T makes every even
// item an odd item depends on come at an earlier ticket (the CPU test
// enumerates the dependencies site by site); `slack` more even items let
// those finish before the odd item starts, at the price of a longer
// distance between a block's even and odd items (more data in between).
SU3GF_PIPE_HD bool pipe_plan(const int N[4], PipePlan &p, int slack = 0)
{
    if (N[0] % kPipeQueues != 0) return false;
    const int64_t half_slice = (int64_t)N[1] * N[2] * N[3] / 2;
    if (half_slice % 64 != 0) return false;
    p.T = N[0] / kPipeQueues;
    p.Q = (int)(half_slice / 64);
    p.R = (int)((63 + (int64_t)N[2] * N[3] / 2) / 64);
    if (2 * p.R + 1 > p.Q) return false;  // neighbourhood wraps onto itself: too small to pay
    p.n = (int64_t)p.Q * p.T;
    p.W = (int64_t)p.R * p.T;
    const int64_t L = (int64_t)(2 * p.R + 1) * p.T + (slack > 0 ? slack : 0);
    p.L = L < p.n ? L : p.n;
    return true;
}

// Ticket k of a queue -> odd (0 even, 1 odd) and idx.  Tickets: even items
// 0..L-1, then odd and even alternating, then the remaining odd items.  The
// odd items run in the order idx = W, W+1, ..., n-1, 0, ..., W-1.
SU3GF_PIPE_HD void pipe_item(const PipePlan &p, int64_t k, int &odd, int64_t &idx)
{
    const int64_t mid = 2 * (p.n - p.L);
    int64_t m;
    if (k < p.L) {
        odd = 0;
        idx = k;
        return;
    }
    if (k < p.L + mid) {
        const int64_t j = k - p.L;
        if (j & 1) {
            odd = 0;
            idx = p.L + (j >> 1);
            return;
        }
        m = j >> 1;
    } else {
        m = (p.n - p.L) + (k - p.L - mid);
    }
    odd = 1;
    idx = m + p.W;
    if (idx >= p.n) idx -= p.n;
}

// The ticket of even item idx / odd item idx (inverse of pipe_item).
SU3GF_PIPE_HD int64_t pipe_ticket_even(const PipePlan &p, int64_t idx)
{
    return idx < p.L ? idx : p.L + 2 * (idx - p.L) + 1;
}
SU3GF_PIPE_HD int64_t pipe_ticket_odd(const PipePlan &p, int64_t idx)
{
    int64_t m = idx - p.W;
    if (m < 0) m += p.n;
    return m < p.n - p.L ? p.L + 2 * m : p.L + 2 * (p.n - p.L) + (m - (p.n - p.L));
}

// Words of the synchronisation buffer for a lattice of `blocks` 64-site
// blocks per parity.
SU3GF_PIPE_HD int64_t pipe_sync_words(int64_t blocks)
{
    return (int64_t)(sizeof(PipeSync) / sizeof(uint32_t)) + blocks;
}

}  // namespace su3gf
