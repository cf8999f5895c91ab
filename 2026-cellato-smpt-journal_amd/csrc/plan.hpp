// Launch planning for one pass over one grid: which stream-kernel variant, how lanes map to
// word columns, how rows are cut into tasks. Host code.
#ifndef CELLATO_HIP_PLAN_HPP
#define CELLATO_HIP_PLAN_HPP

#include <string>

#include "cellato_hip.h"
#include "dispatch.hpp"

namespace cellato::hip {

struct pass_plan {
    const void* fn = nullptr;
    pass_args args{};
    int blocks = 0;
};

inline std::string& plan_error() {
    static thread_local std::string e;
    return e;
}

// message of the last failed C-ABI call on this thread (cellato_hip_last_error)
inline std::string& last_error() {
    static thread_local std::string e;
    return e;
}

// Rows per task when the caller passes 0. A task streams `chunk` rows plus 2T rows of
// pipeline fill and drain, so long chunks waste less; but a launch needs a few thousand
// waves to keep 1024 SIMDs busy. Pick the longest chunk that still yields about
// TARGET_TASKS tasks for `streams` row-streams of work (segments x rows over all grids of
// the launch), never shorter than max(8, 2T) and never longer than the grid.
constexpr int TARGET_TASKS = 3072;

inline int auto_chunk(long long streams, int T, int rows_per_group) {
    long long c = (streams + TARGET_TASKS - 1) / TARGET_TASKS;
    const int lo = 2 * T > 8 ? 2 * T : 8;
    if (c < lo) c = lo;
    if (c > 256) c = 256;
    if (c > rows_per_group) c = rows_per_group > 0 ? rows_per_group : 1;
    return static_cast<int>(c);
}

// width / height in cells. T = generations per launch (1 for the simple traverser).
inline int plan_pass(int rule, int layout, int width, int height, int T, bool temporal, int chunk, pass_plan& pl,
                     long long launch_streams = 0) {
    const rule_kernels& k = kernels_of_rule(rule);
    const int wx = layout == CELLATO_LAYOUT_LINEAR ? width / 64 : width / 8;
    const int h = layout == CELLATO_LAYOUT_LINEAR ? height : height / 8;
    pass_args& a = pl.args;
    if (static_cast<unsigned long long>(wx) * static_cast<unsigned long long>(h) * 8ULL >= (1ULL << 32)) {
        plan_error() = "a bit plane must be smaller than 4 GiB";
        return CELLATO_E_GEOMETRY;
    }
    a.wx = wx;
    a.h = h;
    if (wx < 64) {
        a.n = wx;
        a.groups = 64 / wx;
        a.stride = 0;
        a.nseg = 1;
        a.variant = VARIANT_GROUPS;
        pl.fn = temporal ? k.temporal[layout][TEMPORAL_GROUPS][T] : k.simple[layout][SIMPLE_GROUPS];
    } else if (wx == 64) {
        a.n = 64;
        a.groups = 1;
        a.stride = 0;
        a.nseg = 1;
        a.variant = VARIANT_RING64;
        pl.fn = temporal ? k.temporal[layout][TEMPORAL_RING64][T] : k.simple[layout][SIMPLE_RING64];
    } else if (!temporal) {
        a.n = 64;
        a.groups = 1;
        a.stride = 64;
        a.nseg = (wx + 63) / 64;
        a.variant = -1;   // halo loads: not used by the batch kernel
        pl.fn = k.simple[layout][SIMPLE_SEGMENT];
    } else {
        a.n = 64;
        a.groups = 1;
        a.stride = 62;
        a.nseg = (wx + 61) / 62;
        a.variant = VARIANT_SEGMENT;
        pl.fn = k.temporal[layout][TEMPORAL_SEGMENT][T];
    }
    if (!pl.fn) {
        plan_error() = "no kernel instantiated for this configuration";
        return CELLATO_E_TEMPORAL;
    }
    const int rows_per_group = (h + a.groups - 1) / a.groups;
    const long long streams = launch_streams > 0 ? launch_streams : static_cast<long long>(a.nseg) * rows_per_group;
    a.chunk = chunk > 0 ? chunk : auto_chunk(streams, T, rows_per_group);
    a.nchunk = (h + a.groups * a.chunk - 1) / (a.groups * a.chunk);
    const long long tasks = static_cast<long long>(a.nseg) * a.nchunk;
    if (tasks > (1LL << 30)) {
        plan_error() = "grid too large for one launch";
        return CELLATO_E_GEOMETRY;
    }
    a.ntasks = static_cast<int>(tasks);
    pl.blocks = (a.ntasks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    a.t0 = 0;
    return CELLATO_OK;
}

// segments x rows per lane group of one pass: the unit of work auto_chunk divides
inline long long pass_streams(int rule, int layout, int width, int height, int T, bool temporal) {
    pass_plan pl;
    if (plan_pass(rule, layout, width, height, T, temporal, 1, pl) != CELLATO_OK) return 0;
    return static_cast<long long>(pl.args.nseg) * ((pl.args.h + pl.args.groups - 1) / pl.args.groups);
}

} // namespace cellato::hip

#endif

