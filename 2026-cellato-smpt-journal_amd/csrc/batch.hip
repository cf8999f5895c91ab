// Batches of independent grids (different automata, sizes, step counts, depths) advanced by
// one hipGraph launch. Host planning: every grid's steps are split into passes of its own
// depth T (the last pass takes the remainder); a grid's passes occupy consecutive rounds,
// starting where the rounds' work stays level (rounds.hpp); in each round the passes are
// grouped by (rule, T) and each group is one persistent batch_kernel launch over the group's
// task list. Rounds are ordered; the groups of a round run on separate streams. The whole
// schedule is captured into one graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cellato_hip.h"
#include "dispatch.hpp"
#include "plan.hpp"
#include "rounds.hpp"

namespace cellato::hip {
namespace {

int bfail(int code, const std::string& m) {
    last_error() = m;
    return code;
}

// Streams the groups of a round are spread over in the graph: one per hardware queue of the
// process (HIP's default is 4). More streams than queues only share queues, where kernels of
// different streams wait for each other.
constexpr int SIDE_STREAMS = 4;

struct launch_rec {
    const void* fn;
    int task_begin, task_count;
    int blocks;
    long long cost;     // rows x generations of the launch, for spreading launches over streams
};

struct batch {
    std::vector<int> result_in_b;
    std::vector<std::vector<launch_rec>> rounds;
    void* dev = nullptr;            // pass table + task lists
    pass_args* dev_passes = nullptr;
    batch_task* dev_tasks = nullptr;
    hipGraphExec_t exec = nullptr;
    long long tasks = 0;
    int launches = 0;

    ~batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (dev) (void)hipFree(dev);
    }
};

// Streams and events used only while the schedule is captured. Events recorded during a
// capture stand for nodes of the captured graph, so they (and the streams) are released
// as soon as the capture ends, while that graph still exists; nothing of the capture is
// kept once the graph has been instantiated.
struct capture_streams {
    hipStream_t origin = nullptr;
    hipStream_t side[SIDE_STREAMS] = {};
    hipEvent_t fork = nullptr, join[SIDE_STREAMS] = {};

    ~capture_streams() {
        for (auto& e : join) if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (auto& st : side) if (st) (void)hipStreamDestroy(st);
        if (origin) (void)hipStreamDestroy(origin);
    }
};

#define BCHECK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return bfail(CELLATO_E_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// Host schedule of a set of grids: pass table, task lists, launches per round.
struct schedule {
    std::vector<pass_args> passes;
    std::vector<batch_task> tasks;
    std::vector<std::vector<launch_rec>> rounds;
    std::vector<int> result_in_b;
    int launches = 0;
};

int plan_schedule(schedule& S, const cellato_batch_grid* grids, int n, int max_waves, std::vector<int>* starts = nullptr) {
    std::vector<pass_args>& passes = S.passes;
    struct pending_pass { int grid, T, done, index; };
    std::vector<std::vector<pending_pass>> chain(n);   // per grid: its passes in order
    S.result_in_b.assign(n, 0);
    int R = 0;
    for (int gi = 0; gi < n; ++gi) {
        const cellato_batch_grid& g = grids[gi];
        if (g.rule < 0 || g.rule >= 10) return bfail(CELLATO_E_RULE, "grid " + std::to_string(gi) + ": unknown rule");
        if (g.layout != CELLATO_LAYOUT_LINEAR) return bfail(CELLATO_E_LAYOUT, "batches use the linear layout");
        if (g.width <= 0 || g.width % 64 || g.height <= 0)
            return bfail(CELLATO_E_GEOMETRY, "grid " + std::to_string(gi) + ": width must be a positive multiple of 64");
        if (g.steps < 0) return bfail(CELLATO_E_ARG, "negative steps");
        if (g.temporal_steps < 1 || g.temporal_steps > MAX_T)
            return bfail(CELLATO_E_TEMPORAL, "grid " + std::to_string(gi) + ": temporal_steps out of range");
        const int P = kernels_of_rule(g.rule).planes;
        for (int p = 0; p < P; ++p)
            if (!g.a[p] || !g.b[p]) return bfail(CELLATO_E_ARG, "null plane pointer");
        // a depth whose batch kernel would need scratch memory is stepped down
        const int Tg = usable_depth(g.rule, g.layout, g.width / 64, g.temporal_steps, true);
        int done = 0, r = 0;
        while (done < g.steps) {
            const int T = std::min(Tg, g.steps - done);
            chain[gi].push_back({gi, T, done, r});
            done += T;
            ++r;
        }
        S.result_in_b[gi] = r % 2;
        R = std::max(R, r);
    }
    // where each grid's chain of passes starts: level the rounds' work (cell updates)
    std::vector<std::vector<double>> work(n);
    for (int gi = 0; gi < n; ++gi)
        for (const auto& e : chain[gi]) work[gi].push_back(static_cast<double>(grids[gi].width) * grids[gi].height * e.T);
    const std::vector<int> start = place_chains(work, R);
    if (starts) *starts = start;
    std::vector<std::vector<pending_pass>> plan_T(R);   // per round: passes to plan
    // per round and (rule, T) group: total row-streams of work. Each group's launch is sized
    // to fill the GPU by itself: groups run on few hardware queues and often one after the
    // other, so a chunk sized for the whole round would leave a lone group's launch short.
    std::vector<std::map<std::pair<int, int>, long long>> group_streams(R);
    for (int gi = 0; gi < n; ++gi)
        for (const auto& e : chain[gi]) {
            const int r = start[gi] + e.index;
            const cellato_batch_grid& g = grids[gi];
            plan_T[r].push_back(e);
            group_streams[r][{g.rule, e.T}] += pass_streams(g.rule, g.layout, g.width, g.height, e.T, true);
        }
    // plan every pass with the chunk that fills the GPU for its group's work
    std::vector<std::map<std::pair<int, int>, std::vector<int>>> rounds(R);   // (rule, T) -> passes
    for (int r = 0; r < R; ++r) {
        for (const auto& e : plan_T[r]) {
            const cellato_batch_grid& g = grids[e.grid];
            const int P = kernels_of_rule(g.rule).planes;
            pass_plan pl;
            const int rc = plan_pass(g.rule, g.layout, g.width, g.height, e.T, true, g.chunk_rows, pl,
                                     group_streams[r][{g.rule, e.T}]);
            if (rc != CELLATO_OK) return bfail(rc, plan_error());
            // pass k of a grid reads a and writes b when k is even, the other way when odd
            for (int p = 0; p < P; ++p) {
                pl.args.in[p] = e.index % 2 == 0 ? g.a[p] : g.b[p];
                pl.args.out[p] = e.index % 2 == 0 ? g.b[p] : g.a[p];
            }
            pl.args.t0 = g.t0 + e.done;
            rounds[r][{g.rule, e.T}].push_back(static_cast<int>(passes.size()));
            passes.push_back(pl.args);
        }
    }
    S.rounds.resize(R);
    for (int r = 0; r < R; ++r) {
        std::vector<std::pair<long long, launch_rec>> recs;
        for (auto& [key, plist] : rounds[r]) {
            launch_rec L;
            L.fn = kernels_of_rule(key.first).batch[key.second];
            L.task_begin = static_cast<int>(S.tasks.size());
            long long rows = 0;
            for (int pi : plist) {
                for (int t = 0; t < passes[pi].ntasks; ++t) S.tasks.push_back({pi, t});
                rows += static_cast<long long>(passes[pi].ntasks) * passes[pi].chunk;
            }
            L.task_count = static_cast<int>(S.tasks.size()) - L.task_begin;
            const int waves = std::min(L.task_count, max_waves);
            L.blocks = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
            L.cost = rows * key.second;
            recs.push_back({L.cost, L});
            ++S.launches;
        }
        // the launch with the most work first, so that it starts first
        std::stable_sort(recs.begin(), recs.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        for (auto& rec : recs) S.rounds[r].push_back(rec.second);
    }
    return CELLATO_OK;
}

int build(batch& B, const cellato_batch_grid* grids, int n, int max_waves) {
    schedule S;
    const int prc = plan_schedule(S, grids, n, max_waves);
    if (prc != CELLATO_OK) return prc;
    B.result_in_b = S.result_in_b;
    B.rounds = S.rounds;
    B.launches = S.launches;
    B.tasks = static_cast<long long>(S.tasks.size());
    const std::vector<pass_args>& passes = S.passes;
    const std::vector<batch_task>& tasks = S.tasks;

    const size_t pbytes = passes.size() * sizeof(pass_args);
    const size_t tbytes = tasks.size() * sizeof(batch_task);
    BCHECK(hipMalloc(&B.dev, pbytes + tbytes + 256));
    B.dev_passes = static_cast<pass_args*>(B.dev);
    B.dev_tasks = reinterpret_cast<batch_task*>(static_cast<char*>(B.dev) + ((pbytes + 255) / 256) * 256);
    if (pbytes) BCHECK(hipMemcpy(B.dev_passes, passes.data(), pbytes, hipMemcpyHostToDevice));
    if (tbytes) BCHECK(hipMemcpy(B.dev_tasks, tasks.data(), tbytes, hipMemcpyHostToDevice));

    // capture the schedule into one graph
    hipGraph_t graph = nullptr;
    int status = CELLATO_OK;
    hipError_t end = hipSuccess;
    {
        capture_streams C;
        for (auto& st : C.side) BCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        for (auto& e : C.join) BCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        BCHECK(hipEventCreateWithFlags(&C.fork, hipEventDisableTiming));
        BCHECK(hipStreamCreateWithFlags(&C.origin, hipStreamNonBlocking));
        BCHECK(hipStreamBeginCapture(C.origin, hipStreamCaptureModeThreadLocal));
        for (auto& round : B.rounds) {
            if (hipEventRecord(C.fork, C.origin) != hipSuccess) { status = CELLATO_E_HIP; break; }
            // launches come largest first; each goes to the stream with the least work so far
            long long load[SIDE_STREAMS] = {};
            for (size_t j = 0; j < round.size() && status == CELLATO_OK; ++j) {
                const launch_rec& L = round[j];
                const int k = j < SIDE_STREAMS ? static_cast<int>(j)
                                               : static_cast<int>(std::min_element(load, load + SIDE_STREAMS) - load);
                load[k] += L.cost;
                hipStream_t st = C.side[k];
                if (j < SIDE_STREAMS && hipStreamWaitEvent(st, C.fork, 0) != hipSuccess) { status = CELLATO_E_HIP; break; }
                batch_args args{B.dev_passes, B.dev_tasks + L.task_begin, L.task_count};
                void* kargs[] = {&args};
                if (hipLaunchKernel(L.fn, dim3(L.blocks), dim3(BLOCK_THREADS), kargs, 0, st) != hipSuccess)
                    status = CELLATO_E_HIP;
            }
            for (size_t j = 0; j < std::min(round.size(), static_cast<size_t>(SIDE_STREAMS)) && status == CELLATO_OK; ++j) {
                if (hipEventRecord(C.join[j], C.side[j]) != hipSuccess || hipStreamWaitEvent(C.origin, C.join[j], 0) != hipSuccess)
                    status = CELLATO_E_HIP;
            }
            if (status != CELLATO_OK) break;
        }
        end = hipStreamEndCapture(C.origin, &graph);
    }   // capture streams and events released here, the graph still alive
    if (status != CELLATO_OK || end != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return bfail(CELLATO_E_HIP, "graph capture of the batch failed");
    }
    const hipError_t inst = hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);   // the executable graph does not need it
    if (inst != hipSuccess) {
        B.exec = nullptr;
        return bfail(CELLATO_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(inst));
    }
    return CELLATO_OK;
}

// Tables of cellato_hip_multi: two device regions and two pinned host staging buffers used
// alternately; a region is rewritten only after the kernels that read it have finished.
// The groups of the call's rounds are spread over SIDE_STREAMS streams forked from and joined
// back into the caller's stream, as in the graph of a batch.
struct multi_state {
    void* dev[2] = {nullptr, nullptr};
    void* host[2] = {nullptr, nullptr};
    size_t cap[2] = {0, 0};
    hipEvent_t done[2] = {nullptr, nullptr};
    int next = 0;
    hipStream_t side[SIDE_STREAMS] = {};
    hipEvent_t fork = nullptr, join[SIDE_STREAMS] = {};
};

multi_state& mstate() {
    static thread_local multi_state m;
    return m;
}

int run_multi(const cellato_batch_grid* grids, int n, int max_waves, hipStream_t stream, int* result_in_b) {
    schedule S;
    const int prc = plan_schedule(S, grids, n, max_waves);
    if (prc != CELLATO_OK) return prc;
    for (int i = 0; i < n; ++i) result_in_b[i] = S.result_in_b[i];
    if (S.tasks.empty()) return CELLATO_OK;
    multi_state& M = mstate();
    const int k = M.next;
    M.next ^= 1;
    const size_t pbytes = ((S.passes.size() * sizeof(pass_args) + 255) / 256) * 256;
    const size_t bytes = pbytes + S.tasks.size() * sizeof(batch_task);
    if (!M.done[k]) BCHECK(hipEventCreateWithFlags(&M.done[k], hipEventDisableTiming));
    BCHECK(hipEventSynchronize(M.done[k]));   // kernels of the call before last are finished
    if (bytes > M.cap[k]) {
        if (M.dev[k]) BCHECK(hipFree(M.dev[k]));
        if (M.host[k]) BCHECK(hipHostFree(M.host[k]));
        M.dev[k] = M.host[k] = nullptr;
        M.cap[k] = 0;
        const size_t want = bytes * 2;
        BCHECK(hipMalloc(&M.dev[k], want));
        BCHECK(hipHostMalloc(&M.host[k], want, hipHostMallocDefault));
        M.cap[k] = want;
    }
    std::memcpy(M.host[k], S.passes.data(), S.passes.size() * sizeof(pass_args));
    std::memcpy(static_cast<char*>(M.host[k]) + pbytes, S.tasks.data(), S.tasks.size() * sizeof(batch_task));
    BCHECK(hipMemcpyAsync(M.dev[k], M.host[k], bytes, hipMemcpyHostToDevice, stream));
    const auto* dpass = static_cast<const pass_args*>(M.dev[k]);
    const auto* dtask = reinterpret_cast<const batch_task*>(static_cast<char*>(M.dev[k]) + pbytes);
    if (!M.fork) {
        for (auto& sd : M.side) BCHECK(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
        for (auto& e : M.join) BCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        BCHECK(hipEventCreateWithFlags(&M.fork, hipEventDisableTiming));
    }
    for (auto& round : S.rounds) {
        if (round.size() == 1) {       // one group: no fork
            const launch_rec& L = round[0];
            batch_args args{dpass, dtask + L.task_begin, L.task_count};
            void* kargs[] = {&args};
            BCHECK(hipLaunchKernel(L.fn, dim3(L.blocks), dim3(BLOCK_THREADS), kargs, 0, stream));
            continue;
        }
        BCHECK(hipEventRecord(M.fork, stream));
        long long load[SIDE_STREAMS] = {};
        const size_t used = std::min(round.size(), static_cast<size_t>(SIDE_STREAMS));
        for (size_t j = 0; j < used; ++j) BCHECK(hipStreamWaitEvent(M.side[j], M.fork, 0));
        for (size_t j = 0; j < round.size(); ++j) {
            const launch_rec& L = round[j];
            const int q = j < SIDE_STREAMS ? static_cast<int>(j)
                                           : static_cast<int>(std::min_element(load, load + SIDE_STREAMS) - load);
            load[q] += L.cost;
            batch_args args{dpass, dtask + L.task_begin, L.task_count};
            void* kargs[] = {&args};
            BCHECK(hipLaunchKernel(L.fn, dim3(L.blocks), dim3(BLOCK_THREADS), kargs, 0, M.side[q]));
        }
        for (size_t j = 0; j < used; ++j) {
            BCHECK(hipEventRecord(M.join[j], M.side[j]));
            BCHECK(hipStreamWaitEvent(stream, M.join[j], 0));
        }
    }
    BCHECK(hipEventRecord(M.done[k], stream));
    return CELLATO_OK;
}

// ---- cellato_hip_multi_static: the same schedule with its tables in static device memory ----
// The pass table and task lists of one call live in `scratch` (file scope, below), a device
// array that each GPU holds for the life of the process (one copy per device). The host's name
// `scratch` is not a device address: the address of the current device's copy comes from
// hipGetSymbolAddress, taken once per device on the first call there, and only that address is
// passed to copies and kernels. Every call on a device shares its copy: calls that overlap on
// one GPU (other streams, other host threads) race.
constexpr size_t STATIC_SCRATCH_BYTES = CELLATO_MULTI_STATIC_SCRATCH_BYTES;
static_assert(sizeof(pass_args) == 144 && sizeof(batch_task) == 8, "table sizes stated in cellato_hip.h");

} // namespace
} // namespace cellato::hip

static __device__ __align__(256) unsigned char scratch[cellato::hip::STATIC_SCRATCH_BYTES];

namespace cellato::hip {
namespace {

int static_scratch(void** out) {
    static std::mutex mu;
    static std::vector<void*> by_device;
    int dev = 0;
    BCHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (static_cast<int>(by_device.size()) <= dev) by_device.resize(dev + 1, nullptr);
    if (!by_device[dev]) {
        void* p = nullptr;
        BCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        by_device[dev] = p;
    }
    *out = by_device[dev];
    return CELLATO_OK;
}

int run_multi_static(const cellato_batch_grid* grids, int n, int max_waves, hipStream_t stream, int* result_in_b) {
    schedule S;
    const int prc = plan_schedule(S, grids, n, max_waves);
    if (prc != CELLATO_OK) return prc;
    const size_t pbytes = ((S.passes.size() * sizeof(pass_args) + 255) / 256) * 256;
    const size_t bytes = pbytes + S.tasks.size() * sizeof(batch_task);
    if (bytes > STATIC_SCRATCH_BYTES)
        return bfail(CELLATO_E_ARG, "the call's pass and task tables need " + std::to_string(bytes) +
                                        " bytes, more than the " + std::to_string(STATIC_SCRATCH_BYTES) +
                                        " of cellato_hip_multi_static's scratch: split the grids over several calls "
                                        "or use cellato_hip_multi");
    for (int i = 0; i < n; ++i) result_in_b[i] = S.result_in_b[i];
    if (S.tasks.empty()) return CELLATO_OK;
    void* base = nullptr;
    const int src = static_scratch(&base);
    if (src != CELLATO_OK) return src;
    // host image of the tables; the copy from pageable memory is staged before the call returns
    std::vector<unsigned char> image(bytes);
    std::memcpy(image.data(), S.passes.data(), S.passes.size() * sizeof(pass_args));
    std::memcpy(image.data() + pbytes, S.tasks.data(), S.tasks.size() * sizeof(batch_task));
    BCHECK(hipMemcpyAsync(base, image.data(), bytes, hipMemcpyHostToDevice, stream));
    const auto* dpass = static_cast<const pass_args*>(base);
    const auto* dtask = reinterpret_cast<const batch_task*>(static_cast<unsigned char*>(base) + pbytes);
    // every launch on the caller's stream, rounds in order (no side streams or events to create)
    for (auto& round : S.rounds)
        for (const launch_rec& L : round) {
            batch_args args{dpass, dtask + L.task_begin, L.task_count};
            void* kargs[] = {&args};
            BCHECK(hipLaunchKernel(L.fn, dim3(L.blocks), dim3(BLOCK_THREADS), kargs, 0, stream));
        }
    return CELLATO_OK;
}

} // namespace
} // namespace cellato::hip

using namespace cellato::hip;

extern "C" {

int cellato_hip_multi_static(const cellato_batch_grid* grids, int n, int max_waves, void* stream, int* result_in_b) {
    if ((n > 0 && (!grids || !result_in_b)) || n < 0) return bfail(CELLATO_E_ARG, "null argument");
    return run_multi_static(grids, n, max_waves > 0 ? max_waves : 8192, static_cast<hipStream_t>(stream), result_in_b);
}

int cellato_batch_create(const cellato_batch_grid* grids, int n, int max_waves, void** handle) {
    if (!handle || (n > 0 && !grids) || n < 0) return bfail(CELLATO_E_ARG, "null argument");
    *handle = nullptr;
    auto B = std::make_unique<batch>();
    const int rc = build(*B, grids, n, max_waves > 0 ? max_waves : 8192);
    if (rc != CELLATO_OK) return rc;
    *handle = B.release();
    return CELLATO_OK;
}

int cellato_batch_launch(void* handle, void* stream) {
    if (!handle) return bfail(CELLATO_E_ARG, "null batch");
    auto* B = static_cast<batch*>(handle);
    const hipError_t e = hipGraphLaunch(B->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return bfail(CELLATO_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return CELLATO_OK;
}

int cellato_batch_result_in_b(void* handle, int grid) {
    if (!handle) return bfail(CELLATO_E_ARG, "null batch");
    auto* B = static_cast<batch*>(handle);
    if (grid < 0 || grid >= static_cast<int>(B->result_in_b.size())) return bfail(CELLATO_E_ARG, "grid index out of range");
    return B->result_in_b[grid];
}

int cellato_batch_stats(void* handle, int* rounds, int* launches, long long* tasks) {
    if (!handle) return bfail(CELLATO_E_ARG, "null batch");
    auto* B = static_cast<batch*>(handle);
    if (rounds) *rounds = static_cast<int>(B->rounds.size());
    if (launches) *launches = B->launches;
    if (tasks) *tasks = B->tasks;
    return CELLATO_OK;
}

int cellato_hip_multi(const cellato_batch_grid* grids, int n, int max_waves, void* stream, int* result_in_b) {
    if ((n > 0 && (!grids || !result_in_b)) || n < 0) return bfail(CELLATO_E_ARG, "null argument");
    return run_multi(grids, n, max_waves > 0 ? max_waves : 8192, static_cast<hipStream_t>(stream), result_in_b);
}

int cellato_batch_plan(const cellato_batch_grid* grids, int n, int max_waves, int* start_round, int* rounds,
                       int* launches, long long* tasks) {
    if ((n > 0 && !grids) || n < 0) return bfail(CELLATO_E_ARG, "null argument");
    schedule S;
    std::vector<int> starts;
    const int rc = plan_schedule(S, grids, n, max_waves > 0 ? max_waves : 8192, &starts);
    if (rc != CELLATO_OK) return rc;
    if (start_round)
        for (int i = 0; i < n; ++i) start_round[i] = starts[i];
    if (rounds) *rounds = static_cast<int>(S.rounds.size());
    if (launches) *launches = S.launches;
    if (tasks) *tasks = static_cast<long long>(S.tasks.size());
    return CELLATO_OK;
}

int cellato_batch_destroy(void* handle) {
    delete static_cast<batch*>(handle);
    return CELLATO_OK;
}

} // extern "C"
