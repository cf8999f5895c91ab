// Launch planning for one pass over one grid: which stream-kernel variant, how lanes map to
// word columns, how rows are cut into tasks. Host code.
#ifndef CELLATO_HIP_PLAN_HPP
#define CELLATO_HIP_PLAN_HPP

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "cellato_hip.h"
#include "dispatch.hpp"

namespace cellato::hip {

struct pass_plan {
    const void* fn = nullptr;
    pass_args args{};
    int blocks = 0;
    long long streams = 0;   // rows x lane groups of work (the unit auto_chunk divides)
    bool rows = false;       // tiled planes processed as rows: args count 64-cell columns and cell rows
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

// Compute units of the current device (256 on MI355X; also the answer without a device, so
// that plans made on a host without a GPU match).
inline int device_cus() {
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0) {
            (void)hipGetLastError();
            n = 256;
        }
        return n;
    }();
    return cus;
}

// Blocks of a kernel (BLOCK_THREADS threads) one CU holds at once, from the kernel's register
// and static LDS use; cached per kernel. 8 when it cannot be asked (no device).
inline int blocks_per_cu(const void* fn) {
    static std::mutex mu;
    static std::map<const void*, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(fn);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, BLOCK_THREADS, 0) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        n = 8;
    }
    cache[fn] = n;
    return n;
}

// Bytes of scratch (private segment) a kernel needs per lane: nonzero means its registers
// spilled to memory, which multiplies its HBM traffic. Cached per kernel; 0 when it cannot
// be asked (no device: the plans made on a host without a GPU).
inline int private_bytes(const void* fn) {
    static std::mutex mu;
    static std::map<const void*, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(fn);
    if (it != cache.end()) return it->second;
    hipFuncAttributes attr{};
    int n = 0;
    if (hipFuncGetAttributes(&attr, fn) == hipSuccess) n = static_cast<int>(attr.localSizeBytes);
    else (void)hipGetLastError();
    cache[fn] = n;
    return n;
}

// words per row of a grid `width` cells wide (linear: 64 cells per word; tiled: 8)
inline int words_per_row(int layout, int width) { return layout == CELLATO_LAYOUT_LINEAR ? width / 64 : width / 8; }

// Whether a temporal pass of depth T over tiled planes of wx block words per row (and h block
// rows, if known) runs as rows (TILED_ROWS): a row of at least 64 columns of 64 cells, a rule
// with those kernels, and planes within the bound of its 16-byte out-of-range accesses.
inline bool rows_pass(const rule_kernels& k, int layout, int wx, int T, bool temporal, long long h = -1) {
    if (!temporal || layout != CELLATO_LAYOUT_TILED || T < 2 || wx % 8 != 0 || wx / 8 < 64) return false;
    if (h >= 0 && static_cast<unsigned long long>(wx) * static_cast<unsigned long long>(h) * 8ULL > OOB_WIDE) return false;
    return k.temporal_rows[wx / 8 == 64 ? 1 : 0][T] != nullptr;
}

// Kernel of one pass: the stream kernel of the lane mapping that a row of wx words takes.
inline const void* pass_kernel(const rule_kernels& k, int layout, int wx, int T, bool temporal) {
    if (!temporal) return k.simple[layout][wx < 64 ? SIMPLE_GROUPS : wx == 64 ? SIMPLE_RING64 : SIMPLE_SEGMENT];
    if (rows_pass(k, layout, wx, T, temporal)) return k.temporal_rows[wx / 8 == 64 ? 1 : 0][T];
    return k.temporal[layout][wx < 64 ? TEMPORAL_GROUPS : wx == 64 ? TEMPORAL_RING64 : TEMPORAL_SEGMENT][T];
}

// The depth a pass of up to T generations runs at: the deepest T' <= T whose kernel exists
// and needs no scratch memory (kernels that would spill are not instantiated, and any that
// spill on this device are stepped over). `batch`: the work-list kernel of batches instead.
inline int usable_depth(int rule, int layout, int wx, int T, bool batch = false) {
    const rule_kernels& k = kernels_of_rule(rule);
    for (int t = T; t > 1; --t) {
        const void* fn = batch ? k.batch[t] : pass_kernel(k, layout, wx, t, true);
        if (fn && private_bytes(fn) == 0) return t;
    }
    return 1;
}

// Rows per task when the caller passes 0. Every wave runs one task, and a CU takes its
// blocks (4 waves, one per SIMD) at once, so a launch lasts about m x (rows a task walks),
// m = blocks on the busiest CU; a task walks chunk + 2T rows rounded up to whole unrolled
// blocks of unroll_rows(T) rows. For m = 3 and m = 2 the smallest chunk (at least max(8, 2T),
// filling its last unrolled block) whose tasks fit in m blocks per CU is found, and the one
// with fewer rows per SIMD wins; ties keep m = 3 (more waves hide more latency). One block
// per CU leaves loads exposed for light rules and is chosen only when the kernel's registers
// let a CU hold fewer blocks (resident = that count) or the grid is small.
// tasks_of(chunk) = the launch's task count with that chunk.
template <class TasksOf>
inline int auto_chunk(TasksOf tasks_of, int T, int rows_per_group, int resident = 8) {
    const int U = unroll_rows(T);
    const auto align = [&](long long c) { return (c + 2 * T + U - 1) / U * U - 2 * T; };
    const long long lo = align(2 * T > 8 ? 2 * T : 8);
    const long long hi = rows_per_group > 0 ? rows_per_group : 1;
    const long long cus = device_cus();
    long long best = -1, best_cost = 0;
    const int m_hi = resident < 3 ? resident : 3, m_lo = resident < 2 ? resident : 2;
    for (int m = m_hi; m >= m_lo; --m) {
        long long c = lo;
        while (c < hi && (tasks_of(static_cast<int>(c)) + 3) / 4 > m * cus) c += U;
        if (c > hi) c = hi;
        const long long blocks = (tasks_of(static_cast<int>(c)) + 3) / 4;
        const long long per_cu = (blocks + cus - 1) / cus;
        const long long cost = per_cu * ((c + 2 * T + U - 1) / U * U);
        if (best < 0 || cost < best_cost) {
            best = c;
            best_cost = cost;
        }
    }
    return static_cast<int>(best);
}

// width / height in cells. T = generations per launch (1 for the simple traverser). chunk:
// rows of words per task (tiled: block rows), 0 = auto_chunk.
inline int plan_pass(int rule, int layout, int width, int height, int T, bool temporal, int chunk, pass_plan& pl,
                     long long launch_streams = 0) {
    const rule_kernels& k = kernels_of_rule(rule);
    // tiled planes processed as rows take the linear geometry: 64-cell columns, cell rows
    pl.rows = rows_pass(k, layout, width / 8, T, temporal, height / 8);
    const bool lin = layout == CELLATO_LAYOUT_LINEAR || pl.rows;
    const int wx = lin ? width / 64 : width / 8;
    const int h = lin ? height : height / 8;
    if (pl.rows && chunk > 0) chunk = chunk < (1 << 27) ? 8 * chunk : (1 << 30);
    pass_args& a = pl.args;
    if (static_cast<unsigned long long>(wx) * static_cast<unsigned long long>(h) * 8ULL >= (1ULL << 32)) {
        plan_error() = "a bit plane must be smaller than 4 GiB";
        return CELLATO_E_GEOMETRY;
    }
    a.wx = wx;
    a.h = h;
    a.tail_n = a.tail_groups = a.tail_first = a.first = 0;
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
        pl.fn = pl.rows ? k.temporal_rows[1][T] : temporal ? k.temporal[layout][TEMPORAL_RING64][T] : k.simple[layout][SIMPLE_RING64];
    } else if (!temporal) {
        a.n = 64;
        a.groups = 1;
        a.stride = 64;
        a.nseg = (wx + 63) / 64;
        a.variant = -1;   // halo loads: not used by the batch kernel
        pl.fn = k.simple[layout][SIMPLE_SEGMENT];
    } else {
        // full segments of 62 owned columns; the columns left over (fewer than 62) go to
        // narrow lane groups, several row bands per wave, instead of a mostly idle wave
        a.n = 64;
        a.groups = 1;
        a.stride = 62;
        a.first = -1;
        a.nseg = wx / 62;
        const int left = wx - 62 * a.nseg;
        if (left > 0) {
            a.tail_n = left + 2;
            a.tail_groups = 64 / a.tail_n;
            a.tail_first = 62 * a.nseg - 1;
        }
        a.variant = VARIANT_SEGMENT;
        pl.fn = pl.rows ? k.temporal_rows[0][T] : k.temporal[layout][TEMPORAL_SEGMENT][T];
    }
    if (!pl.fn) {
        plan_error() = "no kernel instantiated for this configuration";
        return CELLATO_E_TEMPORAL;
    }
    const int rows_per_group = (h + a.groups - 1) / a.groups;
    pl.streams = static_cast<long long>(a.nseg) * rows_per_group +
                 (a.tail_n > 0 ? (h + a.tail_groups - 1) / a.tail_groups : 0);
    // tasks with a given chunk: this pass alone, or a share of a launch of `launch_streams`
    const auto tasks_of = [&](int c) -> long long {
        if (launch_streams > 0) return (launch_streams + c - 1) / c;
        const long long nch = (h + static_cast<long long>(a.groups) * c - 1) / (static_cast<long long>(a.groups) * c);
        return a.nseg * nch + (a.tail_n > 0 ? (h + static_cast<long long>(a.tail_groups) * c - 1) /
                                                   (static_cast<long long>(a.tail_groups) * c) : 0);
    };
    // a caller's chunk is capped at the rows a lane group covers (a longer one would only make
    // every wave walk the wrapped rows again)
    a.chunk = chunk > 0 ? (chunk < rows_per_group ? chunk : rows_per_group)
                        : auto_chunk(tasks_of, T, rows_per_group, blocks_per_cu(pl.fn));
    // rows: tasks cover whole block rows (h is a multiple of 8)
    if (pl.rows) a.chunk = (a.chunk + 7) / 8 * 8 < h ? (a.chunk + 7) / 8 * 8 : h;
    const long long nchunk = (h + static_cast<long long>(a.groups) * a.chunk - 1) / (static_cast<long long>(a.groups) * a.chunk);
    const long long main_tasks = static_cast<long long>(a.nseg) * nchunk;
    const long long tasks = main_tasks + (a.tail_n > 0 ? (h + static_cast<long long>(a.tail_groups) * a.chunk - 1) /
                                                             (static_cast<long long>(a.tail_groups) * a.chunk) : 0);
    if (tasks > (1LL << 30)) {
        plan_error() = "grid too large for one launch";
        return CELLATO_E_GEOMETRY;
    }
    a.nchunk = static_cast<int>(nchunk);
    a.ntasks = static_cast<int>(tasks);
    a.main_tasks = static_cast<int>(main_tasks);
    pl.blocks = (a.ntasks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    a.t0 = 0;
    return CELLATO_OK;
}

// segments x rows per lane group of one pass: the unit of work auto_chunk divides
inline long long pass_streams(int rule, int layout, int width, int height, int T, bool temporal) {
    pass_plan pl;
    if (plan_pass(rule, layout, width, height, T, temporal, 1, pl) != CELLATO_OK) return 0;
    return pl.streams;
}

} // namespace cellato::hip

#endif

