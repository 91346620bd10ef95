// The moment exchange between the blocks of a multi-block team (rows of 512
// or 1024 bins; cosenu_pair_wide, cosenu_quad_wide).
#pragma once

#include "fast_math.hpp"

namespace cosenu {
namespace fastk {

// Teams of NB > 1 blocks (rows wider than 256 bins): each block of 8
// wavefronts owns 256 bins of the team's z range, and the moments
// of a z row are summed over the team's blocks through device memory.  Every
// wavefront publishes its 6 partials of a row (the values it writes to its
// block's LDS ring) to a ring in device memory as well.  In each partner
// block one loader wavefront per ring (and partner) loads a row's words
// about a z step before its block needs the row's totals, checks their tags
// and stages the values in LDS, where every wavefront of the block reads
// them after the next barrier.
//
// A value travels as two 64-bit words, each holding 32 bits of the double
// and a 32-bit tag (the launch epoch and the z row): a word is written and
// read with a single agent-scope atomic, so a reader sees either the old or
// the new word, and a value is complete when both of its words carry the
// tag it expects.  No fence orders the words, and slots need no clearing
// between uses or launches (a stale word carries another row or epoch).
//
// Blocks form teams in the order they start (a ticket from a 64-bit atomic
// counter that never resets: launch k of a layout takes tickets k * grid ..
// (k + 1) * grid - 1), so a waiting block's partners have started already
// or are the next blocks the dispatcher starts: no block waits on one that
// cannot get a CU.  The epoch is the launch number k, so graph replays need
// no memset node; the host zeroes the counter when the layout changes.  A
// wait is bounded by the wall clock; a block that gives up sets the error
// word (read by the host at its next synchronisation) and later waits end
// at once, so a fault cannot hang the device.
struct WideX {
    unsigned long long *slots; // [rings][nteams][NB][ring rows][64 elements][2 words]
    unsigned long long *ctl;   // [0] ticket, [1] error
    long long wait_ticks;      // the wait bound in wall-clock ticks (1 s)
    unsigned tag_flip;         // 0; a test of the give-up path expects tags no block writes
    unsigned epoch;
    int team, role, nteams;
};

struct XWords {
    unsigned long long w0, w1;
};

// A block of a multi-block launch joins its team: a ticket from the launch's
// counter (ctl[0]) gives team, role and launch epoch.  Contains a barrier:
// every thread of the block calls it.
template <int NB>
__device__ inline WideX wide_join(unsigned long long *slots, unsigned long long *ctl,
                                  long long wait_ticks, unsigned tag_flip, int nteams)
{
    __shared__ unsigned long long ticket;
    if (threadIdx.x == 0)
        ticket = __hip_atomic_fetch_add(ctl, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const unsigned long long grid = gridDim.x;
    const int t = (int)(ticket % grid);
    WideX X;
    X.slots = slots;
    X.ctl = ctl;
    X.wait_ticks = wait_ticks;
    X.tag_flip = tag_flip;
    X.epoch = (unsigned)(ticket / grid);
    X.team = t / NB;
    X.role = t % NB;
    X.nteams = nteams;
    return X;
}

__device__ inline unsigned xtag(unsigned epoch, int row)
{
    return (epoch << 24) | ((unsigned)(row + 64) & 0xFFFFFFu);
}

__device__ inline void xput(unsigned long long *p, unsigned tag, double v)
{
    const unsigned long long t = (unsigned long long)tag << 32;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    __hip_atomic_store(p, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, t | (b & 0xFFFFFFFFull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline XWords xload(unsigned long long *p)
{
    return {__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
            __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
}

// The value of a slot whose words were loaded into w: re-read until both
// words carry `tag` (a wait bounded by wait_ticks).
__device__ inline double xvalue(XWords w, unsigned long long *p, unsigned tag,
                                unsigned long long *ctl, long long wait_ticks, bool &gave_up)
{
    auto ok = [&] { return (unsigned)(w.w0 >> 32) == tag && (unsigned)(w.w1 >> 32) == tag; };
    if (!ok() && !gave_up) {
        if (__hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull)
            gave_up = true; // an earlier wait timed out: results are void anyway
        const long long t0 = wall_clock64();
        while (!gave_up && !ok()) {
            __builtin_amdgcn_s_sleep(8);
            w = xload(p);
            if (!ok() && wall_clock64() - t0 > wait_ticks) {
                gave_up = true;
                __hip_atomic_store(ctl + 1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    return __longlong_as_double((long long)(((w.w0 & 0xFFFFFFFFull) << 32) | (w.w1 & 0xFFFFFFFFull)));
}

// Host: the wait bound, 1 s of the device's constant-rate wall clock.
// give_up_test: every wait looks for a tag no block writes and gives up after
// 10 us, so that the error path runs with every kernel still ending normally.
inline void wide_wait_bound(bool give_up_test, long long *ticks, unsigned *flip)
{
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;
    *ticks = give_up_test ? khz / 100 : 1000LL * khz;
    *flip = give_up_test ? 0x80000000u : 0u;
}

} // namespace fastk
} // namespace cosenu
