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

// Default rows per task: long enough that the 2T rows of pipeline fill and drain stay a
// small fraction of the rows streamed, short enough to give the chip a few thousand waves.
inline int default_chunk(int rows_per_group_total, int T) {
    int c = 16 > 4 * T ? 16 : 4 * T;
    // small grids: do not make a chunk longer than the grid itself
    if (c > rows_per_group_total) c = rows_per_group_total > 0 ? rows_per_group_total : 1;
    return c;
}

// width / height in cells. T = generations per launch (1 for the simple traverser).
inline int plan_pass(int rule, int layout, int width, int height, int T, bool temporal, int chunk, pass_plan& pl) {
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
    a.chunk = chunk > 0 ? chunk : default_chunk(rows_per_group, T);
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

} // namespace cellato::hip

#endif
