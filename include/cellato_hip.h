/* C ABI of the MI355X (gfx950) bit-plane cellular-automaton step.
 *
 * Plain C: no C++ types, no HIP types, no exceptions across the boundary. Every function
 * returns CELLATO_OK (0) or a negative status; cellato_hip_last_error() then holds a
 * message. Grids are bit planes of 64-bit words, P planes (P = cellato_rule_planes(rule)),
 * each plane a separate device allocation of (width/64) * height words (linear layout) or
 * (width/8) * (height/8) words (tiled layout), passed as a host array of P device pointers.
 *
 * Reference interfaces replaced (paths in matyas-brabec/2026-cellato-smpt-journal):
 *   cellato_hip_simple    <- traversers::cuda::simple::traverser::run / run_kernel
 *                            (include/traversers/cuda/simple.hpp:61-71, simple.cu:45-87)
 *   cellato_hip_temporal  <- traversers::cuda::temporal::traverser::run / run_kernel
 *                            (include/traversers/cuda/temporal.hpp:99-108, temporal.cu:139-272)
 *   cellato_rule_*        <- the automaton registry of src/main.cpp:284-318 and the
 *                            per-automaton config structs (src/<automaton>/config.hpp)
 *   cellato_hip_last_error<- the CudaError thrown by CUCH (include/traversers/cuda_utils.cuh:42-52)
 */
#ifndef CELLATO_HIP_H
#define CELLATO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CELLATO_OK 0
#define CELLATO_E_RULE (-1)      /* unknown rule id */
#define CELLATO_E_LAYOUT (-2)    /* unknown layout */
#define CELLATO_E_GEOMETRY (-3)  /* width/height not representable in the layout */
#define CELLATO_E_TEMPORAL (-4)  /* temporal_steps outside 1..cellato_max_temporal_steps() */
#define CELLATO_E_ARG (-5)       /* null pointer, negative steps, ... */
#define CELLATO_E_HIP (-6)       /* a HIP runtime call failed */

#define CELLATO_LAYOUT_LINEAR 0  /* word = 64 cells of a row, cell x at bit x % 64 */
#define CELLATO_LAYOUT_TILED 1   /* word = 8x8 block, cell (x, y) at bit (y % 8) * 8 + x % 8 */

/* ---- rule registry ---------------------------------------------------------------------- */
int cellato_rule_count(void);
/* the reference's automaton names: "game-of-life", "wire", "forest-fire", "excitable",
 * "brian", "maze", "traffic", "fluid", "critters", "cyclic"; NULL for an unknown id */
const char* cellato_rule_name(int rule);
int cellato_rule_find(const char* name);          /* id or CELLATO_E_RULE */
int cellato_rule_planes(int rule);                /* bit planes P, or CELLATO_E_RULE */
/* Writes the state value (enum value) of each dictionary index 0..n-1 into values (at most
 * `capacity` of them) and returns n; plane b of a cell is bit b of its dictionary index
 * (reference: include/memory/state_dictionary.hpp:10-70). */
int cellato_rule_states(int rule, int32_t* values, int capacity);
/* Kinds of nodes the rule's AST uses, as flags: 1 Moore count, 2 von Neumann count,
 * 4 Margolus block, 8 direct neighbour read, 16 count with a per-cell target, 32 plus/modulo,
 * 64 alternating rules (include/cellato/evaluators/bit_planes.hpp, rule_traits). */
int cellato_rule_traits(int rule);
int cellato_max_temporal_steps(void);

/* ---- traversers --------------------------------------------------------------------------
 * Both advance the grid held in buf_a by `steps` generations, the first one computed with
 * time step t0 (rules that alternate or use Margolus blocks read its parity). buf_a and buf_b
 * are ping-ponged; on return *result_in_b tells which one holds the result (buf_a's
 * contents are overwritten whenever steps >= 2). Launches are asynchronous on `stream`
 * (a hipStream_t, NULL = default stream).
 * chunk_rows: rows of words a wave streams through per task (tiled: rows of 8x8 blocks);
 * 0 picks a default.
 *
 * simple:   one generation per kernel launch.
 * temporal: temporal_steps generations per kernel launch (register-resident pipeline);
 *           steps need not be a multiple of temporal_steps (the remainder runs as a
 *           shorter last launch). A depth whose kernel would need scratch memory (the
 *           cyclic automaton's linear layout above 4) runs as launches of the deepest depth
 *           that does not; cellato_hip_kernel_info reports the depth used. Over the tiled
 *           layout, rules of one or two planes from a depth measured per automaton run the
 *           planes as rows of 64-cell columns (same results, same arguments).
 * Geometry: linear needs width % 64 == 0; tiled needs width % 8 == 0 and height % 8 == 0. */
int cellato_hip_simple(int rule, int layout, uint64_t* const* buf_a, uint64_t* const* buf_b,
                       int width, int height, int t0, int steps, int chunk_rows,
                       void* stream, int* result_in_b);

int cellato_hip_temporal(int rule, int layout, uint64_t* const* buf_a, uint64_t* const* buf_b,
                         int width, int height, int t0, int steps, int temporal_steps,
                         int chunk_rows, void* stream, int* result_in_b);

/* bit_array layout (reference: evaluators::bit_array::evaluator under the CUDA simple traverser,
 * include/evaluators/bit_array.hpp:73-107, traversers/cuda/simple.cu:20-97): one device array
 * of word_bits-bit words (32 or 64), cells_per_word = word_bits / P cells of a row per word,
 * cell i (row-major) at bits [(i % cells_per_word) * P, + P) of word i / cells_per_word, as
 * its dictionary index; width must be a multiple of cells_per_word. One generation per
 * launch, every cell evaluated with plain integer semantics; ping-pong between buf_a and
 * buf_b as for the other traversers (*result_in_b = steps % 2). */
int cellato_hip_bit_array(int rule, int word_bits, void* buf_a, void* buf_b, int width, int height, int t0, int steps,
                          void* stream, int* result_in_b);

/* The launch one pass would make (host only, no GPU work): rows per task (`chunk`; chunk_rows
 * 0 = the automatic choice), tasks (one wave each) and blocks of 4 waves. temporal_steps 0 =
 * the simple traverser's kernel, 1..max = the temporal traverser's. No counterpart in the
 * reference (whose launch shape is fixed by its tile-size parameters). */
int cellato_hip_plan(int rule, int layout, int width, int height, int temporal_steps, int chunk_rows, int* chunk,
                     int* tasks, int* blocks);

/* The kernel a pass would launch, described without launching it (needs a device): the
 * depth it runs at (*effective_steps; temporal_steps 0 = the simple traverser, reported as 0),
 * the bytes of scratch memory per lane it needs (*private_segment_bytes) and its registers per
 * lane. A temporal depth whose kernel would need scratch memory (registers spilled to memory)
 * runs as passes of the deepest depth whose kernel does not; `batch` != 0 asks for the batch
 * work-list kernel of that depth (linear layout). No counterpart in the reference. */
int cellato_hip_kernel_info(int rule, int layout, int width, int temporal_steps, int batch, int* effective_steps,
                            int* private_segment_bytes, int* registers);

/* ---- batches ------------------------------------------------------------------------------
 * Many independent grids (any of the rules, any size, own step count, start step and depth)
 * advanced together: the steps of every grid are cut into launches of its own depth, the
 * launches of all grids are grouped by (rule, depth) into persistent work-list kernels, and
 * the whole schedule is captured once into a hipGraph; cellato_batch_launch replays it with a
 * single graph launch. Linear layout. The batch keeps pointers to the caller's buffers:
 * they must stay allocated (and be re-filled with initial states before a replay). After a
 * launch, grid i's result is in b if cellato_batch_result_in_b(handle, i) == 1, else in a.
 * No counterpart in the reference (which advances one grid per traverser). */
typedef struct cellato_batch_grid {
    int rule;
    int layout;             /* CELLATO_LAYOUT_LINEAR */
    int width, height;      /* cells */
    int steps;              /* generations to advance */
    int t0;                 /* time step of the first generation */
    int temporal_steps;     /* generations per launch, 1..cellato_max_temporal_steps() */
    int chunk_rows;         /* 0 = default */
    uint64_t* a[5];         /* P device plane pointers holding the initial state */
    uint64_t* b[5];         /* P device plane pointers of the second buffer */
} cellato_batch_grid;

int cellato_batch_create(const cellato_batch_grid* grids, int n, int max_waves, void** handle);
int cellato_batch_launch(void* handle, void* stream);
int cellato_batch_result_in_b(void* handle, int grid);
int cellato_batch_stats(void* handle, int* rounds, int* launches, long long* tasks);
int cellato_batch_destroy(void* handle);   /* after the stream has finished the batch's launches */
/* Schedule of a batch without building it (host only, no GPU work; the plane pointers must
 * be non-null but are not dereferenced): the round in which each grid's chain of passes starts (start_round, n entries, may
 * be NULL), the round count, kernel launches and tasks. */
int cellato_batch_plan(const cellato_batch_grid* grids, int n, int max_waves, int* start_round, int* rounds,
                       int* launches, long long* tasks);

/* The same schedule launched directly (no graph) on `stream`, for callers whose set of grids
 * changes between calls (the serving loop calls it once per tick with one pass per live job).
 * result_in_b[i] receives which buffer holds grid i's result. The pass tables go through two
 * pinned staging buffers used alternately (kept for the life of the thread). */
int cellato_hip_multi(const cellato_batch_grid* grids, int n, int max_waves, void* stream, int* result_in_b);

/* cellato_hip_multi for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same grids, same schedule and kernels, same results; the call allocates no device or pinned
 * memory and creates no streams or events: the pass and task tables go to a static device
 * array of CELLATO_MULTI_STATIC_SCRATCH_BYTES (8 MiB: the tables take 144 bytes per pass and 8
 * per task, e.g. 1000 passes of 1000 tasks) that each GPU holds for the life of the process,
 * and every launch goes to `stream` in round order. A call whose tables do not fit is refused
 * with CELLATO_E_ARG (nothing is launched). Every call on one GPU shares that array, so calls
 * that overlap on a GPU (from other streams or threads) race; calls on one stream are ordered
 * by it. The host tables are copied from pageable memory, staged before the call returns. */
#define CELLATO_MULTI_STATIC_SCRATCH_BYTES (8u << 20)
int cellato_hip_multi_static(const cellato_batch_grid* grids, int n, int max_waves, void* stream, int* result_in_b);

/* ---- conversions on the device -------------------------------------------------------------
 * values: device array of width*height int32 state values (enum values, row-major), the
 * reference's standard grid. Asynchronous on `stream`.
 * pack:     values -> planes (bit_planes_grid.hpp:234-270, tiled :243-288); *bad_values (a
 *           device int, zeroed by the caller) becomes nonzero if a value is not a state.
 * unpack:   planes -> values (get_cell, bit_planes_grid.hpp:97-124).
 * checksum: the 16 tile sums of state values of standard_grid.hpp:179-213, added into the
 *           device array sums16 (zero it first).
 * fill_random: seeded i.i.d. state per cell: dictionary index k with probability probs[k]
 *           (host array, one per index, normalised); cell (x, y) draws the top 32 bits of
 *           splitmix64(seed ^ splitmix64(y * width + x)). */
int cellato_hip_pack(int rule, int layout, const int32_t* values, int width, int height, uint64_t* const* planes,
                     int* bad_values, void* stream);
int cellato_hip_unpack(int rule, int layout, uint64_t* const* planes, int width, int height, int32_t* values,
                       void* stream);
int cellato_hip_checksum(int rule, int layout, uint64_t* const* planes, int width, int height,
                         unsigned long long* sums16, void* stream);
int cellato_hip_fill_random(int rule, int layout, uint64_t* const* planes, int width, int height,
                            unsigned long long seed, const double* probs, int nprobs, void* stream);

/* ---- cost-model network (autotuner) -------------------------------------------------------
 * fp32, row-major, device pointers, asynchronous on `stream`. No counterpart in the reference.
 * linear_forward:  Y[N, Dout] = act(X[N, Din] W[Din, Dout] + b[Dout]); act = ReLU if relu.
 * linear_backward: G = dY * act'(Y); dW = X^T G; db = column sums of G (if db); dX = G W^T (if dX).
 * mse:             dpred = 2 (pred - target) / N; *loss += mean squared error (zero it first).
 * adam:            one Adam step (bias-corrected, step >= 1) on n parameters. */
int cellato_mlp_linear_forward(const float* X, const float* W, const float* b, float* Y, int N, int Din, int Dout,
                               int relu, void* stream);
int cellato_mlp_linear_backward(const float* X, const float* W, const float* Y, const float* dY, float* G, float* dX,
                                float* dW, float* db, int N, int Din, int Dout, int relu, void* stream);
int cellato_mlp_mse(const float* pred, const float* target, float* dpred, float* loss, int N, void* stream);
int cellato_mlp_adam(float* param, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
                     float eps, int step, void* stream);

/* Message of the last failed call on this host thread ("" if none). */
const char* cellato_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CELLATO_HIP_H */
