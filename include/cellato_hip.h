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
int cellato_max_temporal_steps(void);

/* ---- traversers --------------------------------------------------------------------------
 * Both advance the grid held in buf_a by `steps` generations, the first one computed with
 * time step t0 (rules that alternate or use Margolus blocks read its parity). buf_a and buf_b
 * are ping-ponged; on return *result_in_b tells which one holds the result (buf_a's
 * contents are overwritten whenever steps >= 2). Launches are asynchronous on `stream`
 * (a hipStream_t, NULL = default stream).
 * chunk_rows: rows of words a wave streams through per task; 0 picks a default.
 *
 * simple:   one generation per kernel launch.
 * temporal: temporal_steps generations per kernel launch (register-resident pipeline);
 *           steps need not be a multiple of temporal_steps (the remainder runs as a
 *           shorter last launch).
 * Geometry: linear needs width % 64 == 0; tiled needs width % 8 == 0 and height % 8 == 0. */
int cellato_hip_simple(int rule, int layout, uint64_t* const* buf_a, uint64_t* const* buf_b,
                       int width, int height, int t0, int steps, int chunk_rows,
                       void* stream, int* result_in_b);

int cellato_hip_temporal(int rule, int layout, uint64_t* const* buf_a, uint64_t* const* buf_b,
                         int width, int height, int t0, int steps, int temporal_steps,
                         int chunk_rows, void* stream, int* result_in_b);

/* Message of the last failed call on this host thread ("" if none). */
const char* cellato_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CELLATO_HIP_H */
