// C ABI of include/cellato_hip.h: argument checks, the choice of kernel variant and launch
// shape for a grid, and the ping-pong launch loops of the simple and temporal traversers.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "cellato_hip.h"
#include "dispatch.hpp"
#include "plan.hpp"

namespace cellato::hip {

const rule_kernels& kernels_of_rule(int id) {
    switch (id) {
        case 0: return kernels_rule_0();
        case 1: return kernels_rule_1();
        case 2: return kernels_rule_2();
        case 3: return kernels_rule_3();
        case 4: return kernels_rule_4();
        case 5: return kernels_rule_5();
        case 6: return kernels_rule_6();
        case 7: return kernels_rule_7();
        case 8: return kernels_rule_8();
        default: return kernels_rule_9();
    }
}

namespace {

constexpr int RULES = 10;
int fail(int code, const std::string& msg) {
    last_error() = msg;
    return code;
}

int check_common(int rule, int layout, uint64_t* const* a, uint64_t* const* b, int width, int height,
                 int steps, int* result_in_b) {
    if (rule < 0 || rule >= RULES) return fail(CELLATO_E_RULE, "unknown rule id " + std::to_string(rule));
    if (layout != CELLATO_LAYOUT_LINEAR && layout != CELLATO_LAYOUT_TILED)
        return fail(CELLATO_E_LAYOUT, "unknown layout " + std::to_string(layout));
    if (!a || !b || !result_in_b) return fail(CELLATO_E_ARG, "null buffer or result pointer");
    if (steps < 0) return fail(CELLATO_E_ARG, "negative step count");
    if (width <= 0 || height <= 0) return fail(CELLATO_E_GEOMETRY, "grid size must be positive");
    if (layout == CELLATO_LAYOUT_LINEAR && width % 64 != 0)
        return fail(CELLATO_E_GEOMETRY, "`x_size` must be a multiple of `word store bits` (64), got " + std::to_string(width));
    if (layout == CELLATO_LAYOUT_TILED && (width % 8 != 0 || height % 8 != 0))
        return fail(CELLATO_E_GEOMETRY, "tiled layout needs width and height multiples of 8, got " +
                                            std::to_string(width) + "x" + std::to_string(height));
    const int P = kernels_of_rule(rule).planes;
    for (int p = 0; p < P; ++p)
        if (!a[p] || !b[p]) return fail(CELLATO_E_ARG, "null plane pointer");
    return CELLATO_OK;
}

int launch(const pass_plan& pl, hipStream_t stream) {
    pass_args args = pl.args;
    void* kargs[] = {&args};
    hipError_t e = hipLaunchKernel(pl.fn, dim3(pl.blocks), dim3(BLOCK_THREADS), kargs, 0, stream);
    if (e != hipSuccess) return fail(CELLATO_E_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return CELLATO_OK;
}

// Run `passes` launches of T generations (plus one of `rem`), ping-ponging a <-> b.
int run_passes(int rule, int layout, uint64_t* const* a, uint64_t* const* b, int width, int height, int t0,
               int steps, int T, bool temporal, int chunk, hipStream_t stream, int* result_in_b) {
    const int P = kernels_of_rule(rule).planes;
    // a depth whose kernel would need scratch memory runs as passes of the deepest one that does not
    if (temporal) T = usable_depth(rule, layout, words_per_row(layout, width), T);
    uint64_t* const* cur = a;
    uint64_t* const* nxt = b;
    int flips = 0;
    int done = 0;
    while (done < steps) {
        const int g = temporal ? usable_depth(rule, layout, words_per_row(layout, width), steps - done < T ? steps - done : T) : 1;
        pass_plan pl;
        const int rc = plan_pass(rule, layout, width, height, temporal ? g : 1, temporal, chunk, pl);
        if (rc != CELLATO_OK) return fail(rc, plan_error());
        for (int p = 0; p < P; ++p) {
            pl.args.in[p] = cur[p];
            pl.args.out[p] = nxt[p];
        }
        pl.args.t0 = t0 + done;
        const int lr = launch(pl, stream);
        if (lr != CELLATO_OK) return lr;
        uint64_t* const* tmp = cur;
        cur = nxt;
        nxt = tmp;
        ++flips;
        done += g;
    }
    *result_in_b = flips % 2;
    return CELLATO_OK;
}

} // namespace
} // namespace cellato::hip

using namespace cellato::hip;

extern "C" {

int cellato_rule_count(void) { return RULES; }

const char* cellato_rule_name(int rule) {
    if (rule < 0 || rule >= RULES) return nullptr;
    return kernels_of_rule(rule).name;
}

int cellato_rule_find(const char* name) {
    if (!name) return CELLATO_E_RULE;
    for (int r = 0; r < RULES; ++r)
        if (std::strcmp(kernels_of_rule(r).name, name) == 0) return r;
    return fail(CELLATO_E_RULE, std::string("unknown automaton ") + name);
}

int cellato_rule_planes(int rule) {
    if (rule < 0 || rule >= RULES) return CELLATO_E_RULE;
    return kernels_of_rule(rule).planes;
}

int cellato_rule_states(int rule, int32_t* values, int capacity) {
    if (rule < 0 || rule >= RULES) return CELLATO_E_RULE;
    const rule_kernels& k = kernels_of_rule(rule);
    if (values)
        for (int i = 0; i < k.nstates && i < capacity; ++i) values[i] = k.state_values[i];
    return k.nstates;
}

int cellato_rule_traits(int rule) {
    if (rule < 0 || rule >= RULES) return CELLATO_E_RULE;
    return static_cast<int>(kernels_of_rule(rule).traits);
}

int cellato_max_temporal_steps(void) { return MAX_T; }

int cellato_hip_simple(int rule, int layout, uint64_t* const* buf_a, uint64_t* const* buf_b, int width, int height,
                       int t0, int steps, int chunk_rows, void* stream, int* result_in_b) {
    const int rc = check_common(rule, layout, buf_a, buf_b, width, height, steps, result_in_b);
    if (rc != CELLATO_OK) return rc;
    return run_passes(rule, layout, buf_a, buf_b, width, height, t0, steps, 1, false, chunk_rows,
                      static_cast<hipStream_t>(stream), result_in_b);
}

int cellato_hip_temporal(int rule, int layout, uint64_t* const* buf_a, uint64_t* const* buf_b, int width, int height,
                         int t0, int steps, int temporal_steps, int chunk_rows, void* stream, int* result_in_b) {
    const int rc = check_common(rule, layout, buf_a, buf_b, width, height, steps, result_in_b);
    if (rc != CELLATO_OK) return rc;
    if (temporal_steps < 1 || temporal_steps > MAX_T)
        return fail(CELLATO_E_TEMPORAL, "temporal_steps must be in 1.." + std::to_string(MAX_T) + ", got " +
                                             std::to_string(temporal_steps));
    return run_passes(rule, layout, buf_a, buf_b, width, height, t0, steps, temporal_steps, true, chunk_rows,
                      static_cast<hipStream_t>(stream), result_in_b);
}

int cellato_hip_plan(int rule, int layout, int width, int height, int temporal_steps, int chunk_rows, int* chunk,
                     int* tasks, int* blocks) {
    if (rule < 0 || rule >= RULES) return fail(CELLATO_E_RULE, "unknown rule id " + std::to_string(rule));
    if (layout != CELLATO_LAYOUT_LINEAR && layout != CELLATO_LAYOUT_TILED) return fail(CELLATO_E_LAYOUT, "unknown layout");
    if (width <= 0 || height <= 0 || (layout == CELLATO_LAYOUT_LINEAR && width % 64) ||
        (layout == CELLATO_LAYOUT_TILED && (width % 8 || height % 8)))
        return fail(CELLATO_E_GEOMETRY, "grid size not representable in the layout");
    if (temporal_steps < 0 || temporal_steps > MAX_T) return fail(CELLATO_E_TEMPORAL, "temporal_steps out of range");
    pass_plan pl;
    const bool temporal = temporal_steps > 0;
    const int T = temporal ? usable_depth(rule, layout, words_per_row(layout, width), temporal_steps) : 1;
    const int rc = plan_pass(rule, layout, width, height, T, temporal, chunk_rows, pl);
    if (rc != CELLATO_OK) return fail(rc, plan_error());
    if (chunk) *chunk = pl.rows ? pl.args.chunk / 8 : pl.args.chunk;   // in rows of the layout's words
    if (tasks) *tasks = pl.args.ntasks;
    if (blocks) *blocks = pl.blocks;
    return CELLATO_OK;
}

int cellato_hip_bit_array(int rule, int word_bits, void* buf_a, void* buf_b, int width, int height, int t0, int steps,
                          void* stream, int* result_in_b) {
    if (rule < 0 || rule >= RULES) return fail(CELLATO_E_RULE, "unknown rule id " + std::to_string(rule));
    if (word_bits != 32 && word_bits != 64)
        return fail(CELLATO_E_ARG, "bit_array words are 32 or 64 bits, got " + std::to_string(word_bits));
    if (!buf_a || !buf_b || !result_in_b) return fail(CELLATO_E_ARG, "null buffer or result pointer");
    if (steps < 0) return fail(CELLATO_E_ARG, "negative step count");
    const rule_kernels& k = kernels_of_rule(rule);
    const int cpw = word_bits / k.planes;
    if (width <= 0 || height <= 0 || width % cpw != 0)
        return fail(CELLATO_E_GEOMETRY, "Grid x_size must be divisible by cells_per_word. x_size=" + std::to_string(width) +
                                            ", cells_per_word=" + std::to_string(cpw));
    const long long words = static_cast<long long>(width / cpw) * height;
    const long long blocks = (words + BLOCK_THREADS - 1) / BLOCK_THREADS;
    if (blocks > 0x7FFFFFFFLL) return fail(CELLATO_E_GEOMETRY, "grid too large for one launch");
    bit_array_args a{nullptr, nullptr, width / cpw, height, 0, words};
    void* cur = buf_a;
    void* nxt = buf_b;
    const void* fn = k.bit_array[word_bits == 64 ? 1 : 0];
    for (int s = 0; s < steps; ++s) {
        a.in = cur;
        a.out = nxt;
        a.t = t0 + s;
        void* kargs[] = {&a};
        const hipError_t e = hipLaunchKernel(fn, dim3(static_cast<unsigned>(blocks)), dim3(BLOCK_THREADS), kargs, 0,
                                             static_cast<hipStream_t>(stream));
        if (e != hipSuccess) return fail(CELLATO_E_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
        void* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    *result_in_b = steps % 2;
    return CELLATO_OK;
}

int cellato_hip_kernel_info(int rule, int layout, int width, int temporal_steps, int batch, int* effective_steps,
                            int* private_segment_bytes, int* registers) {
    if (rule < 0 || rule >= RULES) return fail(CELLATO_E_RULE, "unknown rule id " + std::to_string(rule));
    if (layout != CELLATO_LAYOUT_LINEAR && layout != CELLATO_LAYOUT_TILED) return fail(CELLATO_E_LAYOUT, "unknown layout");
    if (batch && layout != CELLATO_LAYOUT_LINEAR) return fail(CELLATO_E_LAYOUT, "batches use the linear layout");
    if (width <= 0 || (layout == CELLATO_LAYOUT_LINEAR && width % 64) || (layout == CELLATO_LAYOUT_TILED && width % 8))
        return fail(CELLATO_E_GEOMETRY, "grid width not representable in the layout");
    if (temporal_steps < 0 || temporal_steps > MAX_T || (batch && temporal_steps < 1))
        return fail(CELLATO_E_TEMPORAL, "temporal_steps out of range");
    const rule_kernels& k = kernels_of_rule(rule);
    const int wx = words_per_row(layout, width);
    const bool temporal = temporal_steps > 0;
    const int T = temporal ? usable_depth(rule, layout, wx, temporal_steps, batch != 0) : 1;
    const void* fn = batch ? k.batch[T] : pass_kernel(k, layout, wx, T, temporal);
    if (!fn) return fail(CELLATO_E_TEMPORAL, "no kernel instantiated for this configuration");
    hipFuncAttributes attr{};
    const hipError_t e = hipFuncGetAttributes(&attr, fn);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(CELLATO_E_HIP, std::string("hipFuncGetAttributes: ") + hipGetErrorString(e));
    }
    if (effective_steps) *effective_steps = temporal ? T : 0;
    if (private_segment_bytes) *private_segment_bytes = static_cast<int>(attr.localSizeBytes);
    if (registers) *registers = attr.numRegs;
    return CELLATO_OK;
}

const char* cellato_hip_last_error(void) { return last_error().c_str(); }

} // extern "C"
