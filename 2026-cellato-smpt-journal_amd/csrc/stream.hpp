// Register-streaming bit-plane step for gfx950 (wave64).
//
// One wave owns a band of word columns and walks down a chunk of rows. Lane = one 64-bit
// word column (linear layout: 64 cells of a row; tiled layout: an 8x8 cell block). For each
// of the T generations computed per launch the wave keeps a 3-row window of the previous
// generation in VGPRs, so T generations of a row are produced per row loaded from HBM
// (a wavefront pipeline: stage g computes generation g of row s-g while row s is loaded).
// The words left and right of a lane come from the neighbouring lanes through DPP wave
// shifts (full 64-lane segments) or ds_bpermute (rings of fewer lanes); the lanes at the
// ends of a segment either load their outside neighbour (one generation per launch) or are
// halo lanes whose results are discarded (T generations per launch, segments overlap by
// two words). Nothing is staged in LDS; there is no barrier in the loop.
#ifndef CELLATO_HIP_STREAM_HPP
#define CELLATO_HIP_STREAM_HPP

#include <hip/hip_runtime.h>
#include <cstdint>

#include "cellato/evaluators/bit_planes.hpp"

namespace cellato::hip {

using word = std::uint64_t;
using cellato::evaluators::bit_planes::vint;

constexpr int MAX_PLANES = 5;
constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK_THREADS = 64 * WAVES_PER_BLOCK;
constexpr int PREFETCH = 4;   // rows loaded ahead of the row being evaluated

enum layout_kind : int { LINEAR = 0, TILED = 1 };

// How lanes find the word columns to their left and right.
enum xch_kind : int {
    XCH_SEGMENT = 0,   // 64 consecutive columns of a longer row; lanes 0 / 63 use a halo value
    XCH_RING64 = 1,    // a whole row of exactly 64 columns
    XCH_PERMUTE = 2,   // anything else: explicit source lanes, optional halo at either end
};

// How the two ends of a segment get their outside neighbours.
enum halo_kind : int {
    HALO_NONE = 0,     // rings: no outside
    HALO_LOAD = 1,     // the end lanes load the outside column (exact for one generation)
    HALO_LANES = 2,    // the end lanes are halo lanes: computed from garbage, never stored
};

// Kernel arguments of one grid pass (one launch over one grid).
struct pass_args {
    const word* in[MAX_PLANES];
    word* out[MAX_PLANES];
    int wx;          // words per row (linear: W/64, tiled: W/8)
    int h;           // rows of words (linear: H, tiled: H/8)
    int t0;          // time step of the first generation computed by this pass
    int chunk;       // rows per lane group and task
    int n;           // lanes per group (segment width)
    int groups;      // lane groups per wave
    int stride;      // first column of segment s is s*stride - (halo lanes ? 1 : 0)
    int nseg;        // segments per row
    int nchunk;      // row chunks
    int ntasks;      // nseg * nchunk
};

__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t lo32(word v) { return static_cast<uint32_t>(v); }
__device__ __forceinline__ uint32_t hi32(word v) { return static_cast<uint32_t>(v >> 32); }
__device__ __forceinline__ word join(uint32_t lo, uint32_t hi) { return (word(hi) << 32) | lo; }

// DPP controls of the GFX9 wave-wide shifts and rotates (lane i reads lane i-1 / i+1)
constexpr int DPP_WAVE_SHL1 = 0x130;   // lane i <- lane i+1
constexpr int DPP_WAVE_ROL1 = 0x134;   // lane i <- lane i+1, lane 63 <- lane 0
constexpr int DPP_WAVE_SHR1 = 0x138;   // lane i <- lane i-1
constexpr int DPP_WAVE_ROR1 = 0x13C;   // lane i <- lane i-1, lane 0 <- lane 63

template <int ctrl>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t src) {
    // bound_ctrl = false: a lane without a source (lane 0 for shr, 63 for shl) keeps `old`
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(src), ctrl, 0xF, 0xF, false));
}

// ---- lane exchange policies ------------------------------------------------------------------
// left(v, halo): the value v of the column to the left; right(v, halo): to the right.
// 32-bit halves are exchanged separately so that a caller needing one half moves one half.
template <int X>
struct xch;

template <>
struct xch<XCH_SEGMENT> {
    __device__ uint32_t left(uint32_t v, uint32_t halo) const { return dpp<DPP_WAVE_SHR1>(halo, v); }
    __device__ uint32_t right(uint32_t v, uint32_t halo) const { return dpp<DPP_WAVE_SHL1>(halo, v); }
};

template <>
struct xch<XCH_RING64> {
    __device__ uint32_t left(uint32_t v, uint32_t) const { return dpp<DPP_WAVE_ROR1>(0u, v); }
    __device__ uint32_t right(uint32_t v, uint32_t) const { return dpp<DPP_WAVE_ROL1>(0u, v); }
};

template <>
struct xch<XCH_PERMUTE> {
    int src_l, src_r;      // byte address (4 * lane) of the left / right source lane
    bool halo_l, halo_r;   // take the halo value instead
    __device__ uint32_t left(uint32_t v, uint32_t halo) const {
        const uint32_t x = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_l, static_cast<int>(v)));
        return halo_l ? halo : x;
    }
    __device__ uint32_t right(uint32_t v, uint32_t halo) const {
        const uint32_t x = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_r, static_cast<int>(v)));
        return halo_r ? halo : x;
    }
};

// ---- windows ---------------------------------------------------------------------------------
// rows[0..2] = the words of rows y-1, y, y+1 in this lane's column; halo[] the same for
// the column outside the segment end this lane sits on (HALO_LOAD only).
template <int P>
struct rows3 {
    word w[3][P];
};

// Linear layout: a word is 64 cells of one row; cell x at bit x % 64.
template <int P, int X, bool HALO>
struct linear_window {
    const rows3<P>& c;
    const rows3<P>& hl;
    xch<X> x;
    word yo;    // all ones when the row is odd
    int t;

    __device__ word yodd() const { return yo; }

    // neighbour column dx of a per-lane value (own) and its halo counterpart
    template <int dx>
    __device__ word hshift(word own, word halo) const {
        if constexpr (dx == 0) {
            return own;
        } else if constexpr (dx < 0) {
            const uint32_t l = x.left(hi32(own), HALO ? hi32(halo) : 0u);
            return (own << 1) | (l >> 31);
        } else {
            const uint32_t r = x.right(lo32(own), HALO ? lo32(halo) : 0u);
            return (own >> 1) | (word(r) << 63);
        }
    }

    template <int dx, int dy>
    __device__ vint<P> nb() const {
        vint<P> r;
#pragma unroll
        for (int p = 0; p < P; ++p) r.p[p] = hshift<dx>(c.w[dy + 1][p], hl.w[dy + 1][p]);
        return r;
    }

    template <int dx, int dy, class F>
    __device__ word nbf(F f) const {
        vint<P> own, halo;
#pragma unroll
        for (int p = 0; p < P; ++p) { own.p[p] = c.w[dy + 1][p]; halo.p[p] = hl.w[dy + 1][p]; }
        if constexpr (dx == 0) return f(own);
        else return hshift<dx>(f(own), HALO ? f(halo) : word(0));
    }
};

// Tiled layout: a word is an 8x8 block; cell (x, y) at bit (y % 8) * 8 + x % 8.
template <int P, int X, bool HALO>
struct tiled_window {
    const rows3<P>& c;
    const rows3<P>& hl;
    xch<X> x;
    int t;

    static constexpr word COL0 = 0x0101010101010101ULL;
    static constexpr word COL7 = 0x8080808080808080ULL;

    __device__ word yodd() const { return 0xFF00FF00FF00FF00ULL; }   // rows 1, 3, 5, 7 of a block

    // the block shifted by one cell row, filled from the block above / below
    template <int dy>
    __device__ static word vshift(const word (&col)[3][P], int p) {
        if constexpr (dy == 0) return col[1][p];
        else if constexpr (dy < 0) return (col[1][p] << 8) | (col[0][p] >> 56);
        else return (col[1][p] >> 8) | (col[2][p] << 56);
    }

    template <int dx>
    __device__ word hshift(word own, word halo) const {
        if constexpr (dx == 0) {
            return own;
        } else if constexpr (dx < 0) {
            const word l = join(x.left(lo32(own), HALO ? lo32(halo) : 0u), x.left(hi32(own), HALO ? hi32(halo) : 0u));
            return ((own << 1) & ~COL0) | ((l >> 7) & COL0);
        } else {
            const word r = join(x.right(lo32(own), HALO ? lo32(halo) : 0u), x.right(hi32(own), HALO ? hi32(halo) : 0u));
            return ((own >> 1) & ~COL7) | ((r << 7) & COL7);
        }
    }

    template <int dx, int dy>
    __device__ vint<P> nb() const {
        vint<P> r;
#pragma unroll
        for (int p = 0; p < P; ++p) r.p[p] = hshift<dx>(vshift<dy>(c.w, p), HALO ? vshift<dy>(hl.w, p) : word(0));
        return r;
    }

    template <int dx, int dy, class F>
    __device__ word nbf(F f) const {
        vint<P> own, halo;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            own.p[p] = vshift<dy>(c.w, p);
            halo.p[p] = HALO ? vshift<dy>(hl.w, p) : word(0);
        }
        if constexpr (dx == 0) return f(own);
        else return hshift<dx>(f(own), HALO ? f(halo) : word(0));
    }
};

__device__ __forceinline__ int wrap(int v, int m) {
    int r = v % m;
    return r < 0 ? r + m : r;
}

// ---- one task: a band of columns x a chunk of rows, T generations --------------------------
template <class Eval, int LAYOUT, int T, int X, int HALO>
__device__ __forceinline__ void stream_task(const pass_args& a, int task) {
    constexpr int P = Eval::planes;
    constexpr bool LOADS_HALO = HALO == HALO_LOAD;
    static_assert(!(HALO == HALO_LOAD && T != 1), "loaded halos are exact for one generation only");

    const int lane = lane_id();
    const int seg = task % a.nseg;
    const int chunk = task / a.nseg;

    // lane -> (group, column)
    const int g = lane / a.n;
    const int w = lane - g * a.n;
    const bool lane_used = g < a.groups;
    int col;                  // this lane's word column
    bool owner;               // whether this lane stores its column
    int halo_col = 0;         // the outside column of an end lane (HALO_LOAD)
    bool end_lane = false;
    xch<X> xc{};
    if constexpr (HALO == HALO_LANES) {
        // segments overlap by two columns; lanes 0 and 63 are halo lanes
        const int first = seg * a.stride - 1;
        col = wrap(first + lane, a.wx);
        const int owned = seg * a.stride + lane - 1;
        owner = lane >= 1 && lane <= 62 && owned < a.wx;
    } else if constexpr (HALO == HALO_LOAD) {
        // full 64-column segments; the last one is moved back to end at the row's end
        // (columns it shares with its neighbour are written twice with the same value)
        const int first = seg * a.stride < a.wx - 64 ? seg * a.stride : a.wx - 64;
        col = first + lane;
        owner = true;
        end_lane = lane == 0 || lane == 63;
        halo_col = lane == 0 ? wrap(first - 1, a.wx) : wrap(first + 64, a.wx);
    } else {
        // whole rows: groups of n lanes, one row each, wrapping inside the group
        col = w < a.wx ? w : a.wx - 1;
        owner = lane_used;
    }
    if constexpr (X == XCH_PERMUTE) {
        static_assert(HALO == HALO_NONE, "permuted exchange is used for whole-row groups only");
        const int base = g * a.n;
        xc.src_l = 4 * (base + (w == 0 ? a.n - 1 : w - 1));
        xc.src_r = 4 * (base + (w == a.n - 1 ? 0 : w + 1));
        xc.halo_l = xc.halo_r = false;
    }

    // rows handled by this lane's group
    const int r0 = (chunk * a.groups + (lane_used ? g : 0)) * a.chunk;
    const int r1 = r0 + a.chunk < a.h ? r0 + a.chunk : a.h;

    rows3<P> win[T];         // win[k]: rows of generation k around the row stage k+1 computes
    rows3<P> hwin;           // halo column, generation 0 (HALO_LOAD)
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int p = 0; p < P; ++p) win[k].w[i][p] = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int p = 0; p < P; ++p) hwin.w[i][p] = 0;

    const size_t wx = static_cast<size_t>(a.wx);
    // row of the next load, kept wrapped onto [0, h) incrementally (no division in the loop)
    int lrow = wrap(r0 - T, a.h);
    auto load_row = [&](word (&dst)[P], word (&hdst)[P]) {
        const size_t row = static_cast<size_t>(lrow) * wx;
#pragma unroll
        for (int p = 0; p < P; ++p) dst[p] = a.in[p][row + col];
        if constexpr (LOADS_HALO) {
            if (end_lane) {
#pragma unroll
                for (int p = 0; p < P; ++p) hdst[p] = a.in[p][row + halo_col];
            }
        }
        lrow = lrow + 1 == a.h ? 0 : lrow + 1;
    };

    // Step i loads row s = r0 - T + i; every group runs the same number of steps, so the
    // loop is uniform across the wave even when groups walk different rows.
    const int nsteps = a.chunk + 2 * T;
    word pf[PREFETCH][P];
    word hpf[PREFETCH][P];
#pragma unroll
    for (int u = 0; u < PREFETCH; ++u) {
#pragma unroll
        for (int p = 0; p < P; ++p) hpf[u][p] = 0;
        load_row(pf[u], hpf[u]);
    }
    int ycur = wrap(r0 - T - 2, a.h);   // advanced to row s - 1 (the row stage 0 computes) at each step

    for (int ib = 0; ib < nsteps; ib += PREFETCH) {
#pragma unroll
        for (int u = 0; u < PREFETCH; ++u) {
            const int s = r0 - T + ib + u;
            ycur = ycur + 1 == a.h ? 0 : ycur + 1;
            // push the loaded row into generation 0's window
#pragma unroll
            for (int p = 0; p < P; ++p) {
                win[0].w[0][p] = win[0].w[1][p];
                win[0].w[1][p] = win[0].w[2][p];
                win[0].w[2][p] = pf[u][p];
                if constexpr (LOADS_HALO) {
                    hwin.w[0][p] = hwin.w[1][p];
                    hwin.w[1][p] = hwin.w[2][p];
                    hwin.w[2][p] = hpf[u][p];
                }
            }
            load_row(pf[u], hpf[u]);

            // stages: generation k+1 of row s-1-k from generation k rows s-2-k .. s-k
            vint<P> out;
#pragma unroll
            for (int k = 0; k < T; ++k) {
                const int t = a.t0 + k;
                if constexpr (LAYOUT == LINEAR) {
                    int y = ycur - k;          // only rules with a Margolus block read it
                    while (y < 0) y += a.h;
                    const linear_window<P, X, LOADS_HALO> wnd{win[k], hwin, xc, (y & 1) ? ~word(0) : word(0), t};
                    out = Eval::evaluate(wnd);
                } else {
                    const tiled_window<P, X, LOADS_HALO> wnd{win[k], hwin, xc, t};
                    out = Eval::evaluate(wnd);
                }
                if (k + 1 < T) {
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        win[k + 1].w[0][p] = win[k + 1].w[1][p];
                        win[k + 1].w[1][p] = win[k + 1].w[2][p];
                        win[k + 1].w[2][p] = out.p[p];
                    }
                }
            }
            // generation T of row s - T (in [r0, r1) implies in [0, h))
            const int yo = s - T;
            if (yo >= r0 && yo < r1 && owner) {
                const size_t at = static_cast<size_t>(yo) * wx + col;
#pragma unroll
                for (int p = 0; p < P; ++p) a.out[p][at] = out.p[p];
            }
        }
    }
}

// Blocks are dealt round-robin to the 8 XCDs; renumber them so each XCD (own L2) gets a
// contiguous range of tasks, i.e. neighbouring chunks and segments share its L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int q = nblocks / 8, rem = nblocks % 8;
    const int xcd = b % 8, local = b / 8;
    return xcd < rem ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

template <class Eval, int LAYOUT, int T, int X, int HALO>
__global__ void __launch_bounds__(BLOCK_THREADS) stream_kernel(pass_args a) {
    const int blk = xcd_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int task = blk * WAVES_PER_BLOCK + static_cast<int>(threadIdx.x) / 64;
    if (task >= a.ntasks) return;
    stream_task<Eval, LAYOUT, T, X, HALO>(a, task);
}

} // namespace cellato::hip

#endif
