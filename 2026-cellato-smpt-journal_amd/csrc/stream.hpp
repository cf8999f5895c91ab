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
// two words). Nothing is shared through LDS and there is no barrier in the loop (the tiled
// planes processed as rows stage each lane's own 8-row blocks in LDS, see TILED_ROWS).
#ifndef CELLATO_HIP_STREAM_HPP
#define CELLATO_HIP_STREAM_HPP

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "cellato/evaluators/bit_planes.hpp"

namespace cellato::hip {

using word = std::uint64_t;
using cellato::evaluators::bit_planes::klist;
using cellato::evaluators::bit_planes::kindex;
using cellato::evaluators::bit_planes::vint;

constexpr int MAX_PLANES = 5;
constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK_THREADS = 64 * WAVES_PER_BLOCK;

// TILED_ROWS: planes stored in the tiled layout, processed as rows of the linear layout. A lane
// owns a column of 64 cells, i.e. 8 consecutive 8x8 blocks (8 block words, 64 bytes per block
// row, moved with four 16-byte accesses). Each lane stages 8 rows per plane in LDS (its own
// slots, nothing shared) from which the row pipeline reads one row per step; at the first row
// of a block row the block words loaded earlier are transposed (8x8 bytes, 32 byte permutes)
// into its 8 rows. Output rows collect in 8 more slots; after the last row of a block row they
// are transposed back and stored, and the loads of the next block row to stage are issued just
// before those stores. Loads and stores thus always come in that order, in one branch, and the
// compiler's wait for the block words counts the stores after them instead of waiting for the
// stores to complete (with the two bursts in branches of their own, every wait included the
// stores' completion; one 8-byte access per step instead was 2-4x slower). The stores are
// plain write-back stores: a lane's 64 bytes go out as four 16-byte pieces of different
// instructions, which L2 merges into whole lines; written through (device scope, as the other
// layouts store) each piece is a partial write to HBM, 1.9x slower at depth 2. Everything in between (windows, row sums, shifts) is the linear layout's: a
// horizontal neighbour costs a funnel shift instead of an exchange with the neighbouring block
// plus byte-row masks, and the pipeline's halo is T cell rows, not T block rows. Tasks cover
// whole block rows (chunks of a multiple of 8 rows).
enum layout_kind : int { LINEAR = 0, TILED = 1, TILED_ROWS = 2 };

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
    HALO_LANES = 2,    // lane groups whose end lanes are halo lanes: computed from garbage,
                       // never stored. A row is cut into 62-column segments (one 64-lane group
                       // per wave) plus a tail of the columns left over (groups of columns + 2
                       // lanes, each group another band of rows); a wave shift across a group
                       // boundary reaches only halo lanes.
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
    int ntasks;      // main_tasks + the tail's tasks
    int main_tasks;  // nseg * nchunk; tasks from here on cover the tail columns
    int first;       // halo-lane groups: column of lane 0 of segment 0 (a halo column)
    int tail_first;  // halo-lane segments: first column (a halo column) of the tail groups
    int tail_n;      // lanes per tail group (tail columns + 2), 0 = no tail
    int tail_groups; // tail groups per wave
    int variant;     // lane mapping, for the batch kernel: 0 segments + halo lanes, 1 ring of 64, 2 groups
};

enum lane_variant : int { VARIANT_SEGMENT = 0, VARIANT_RING64 = 1, VARIANT_GROUPS = 2 };

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

// lanes without a source read 0 (bound_ctrl set): no `old` register to prepare
template <int ctrl>
__device__ __forceinline__ uint32_t dpp0(uint32_t src) {
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(src), ctrl, 0xF, 0xF, true));
}

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
    // segment ends that are halo lanes: nothing outside, zero is as good as anything
    __device__ uint32_t left0(uint32_t v) const { return dpp0<DPP_WAVE_SHR1>(v); }
    __device__ uint32_t right0(uint32_t v) const { return dpp0<DPP_WAVE_SHL1>(v); }
};

template <>
struct xch<XCH_RING64> {
    __device__ uint32_t left(uint32_t v, uint32_t) const { return dpp0<DPP_WAVE_ROR1>(v); }
    __device__ uint32_t right(uint32_t v, uint32_t) const { return dpp0<DPP_WAVE_ROL1>(v); }
    __device__ uint32_t left0(uint32_t v) const { return left(v, 0u); }
    __device__ uint32_t right0(uint32_t v) const { return right(v, 0u); }
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
    __device__ uint32_t left0(uint32_t v) const { return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_l, static_cast<int>(v))); }
    __device__ uint32_t right0(uint32_t v) const { return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(src_r, static_cast<int>(v))); }
};

// ---- windows ---------------------------------------------------------------------------------
// top / mid / bot = the words of rows y-1, y, y+1 in this lane's column (P planes each);
// htop / hmid / hbot the same for the column outside the segment end this lane sits on
// (only read when HALO is set).
template <int P>
using row_t = word[P];

// A row of the linear layout as held in a ring slot: the P plane words of this lane's column
// plus the edge halves of the neighbouring columns that the 1-cell shifts need (bits 32..63
// of the column to the left, bits 0..31 of the column to the right), fetched from the
// neighbouring lanes once when the row enters the ring; and, for each of the NS states the
// rule counts over the Moore neighbourhood, the row's 3-cell horizontal sums of that
// state's mask (sum bit, carry), also formed once at entry and read by the three output rows
// that use the row.
template <int P, int NS = 0>
struct lrow {
    word w[P];
    uint32_t l[P];   // hi half of the left column's word
    uint32_t r[P];   // lo half of the right column's word
    word rs_s[NS > 0 ? NS : 1];
    word rs_c[NS > 0 ? NS : 1];
};

// the word of a cell row shifted by one cell: dx = -1 gives, at bit i, the cell at x - 1
template <int dx>
__device__ __forceinline__ word shift1(word own, uint32_t left_hi, uint32_t right_lo) {
    if constexpr (dx == 0) {
        return own;
    } else if constexpr (dx < 0) {
        return join(__builtin_amdgcn_alignbit(lo32(own), left_hi, 31), __builtin_amdgcn_alignbit(hi32(own), lo32(own), 31));
    } else {
        return join(__builtin_amdgcn_alignbit(hi32(own), lo32(own), 1), __builtin_amdgcn_alignbit(right_lo, hi32(own), 1));
    }
}

// Linear layout: a word is 64 cells of one row; cell x at bit x % 64. Targets = the counted
// states whose row sums the ring slots carry; VnTargets = those of them kept as (3-cell OR,
// mask) for a von Neumann "any" test.
template <int P, class Targets = klist<>, class VnTargets = klist<>>
struct linear_window {
    static constexpr int NS = Targets::size;
    static constexpr bool row_sums = NS > 0;
    using vn_targets = VnTargets;
    const lrow<P, NS>& top;
    const lrow<P, NS>& mid;
    const lrow<P, NS>& bot;
    word yo;    // all ones when the row is odd
    int t;

    __device__ word yodd() const { return yo; }

    template <long long K, int dy>
    __device__ void rsum(word& s, word& c) const {
        constexpr int j = kindex<K, Targets>::value;
        static_assert(j >= 0, "row sums are kept for the rule's counted states only");
        s = row<dy>().rs_s[j];
        c = row<dy>().rs_c[j];
    }

    template <int dy>
    __device__ const lrow<P, NS>& row() const {
        if constexpr (dy < 0) return top;
        else if constexpr (dy == 0) return mid;
        else return bot;
    }

    template <int dx, int dy>
    __device__ vint<P> nb() const {
        const lrow<P, NS>& r = row<dy>();
        vint<P> v;
#pragma unroll
        for (int p = 0; p < P; ++p) v.p[p] = shift1<dx>(r.w[p], r.l[p], r.r[p]);
        return v;
    }

    // F is per cell, so F of the neighbour's edge half is the edge half of F of the neighbour
    template <int dx, int dy, class F>
    __device__ word nbf(F f) const {
        const lrow<P, NS>& r = row<dy>();
        vint<P> own;
#pragma unroll
        for (int p = 0; p < P; ++p) own.p[p] = r.w[p];
        const word fo = f(own);
        if constexpr (dx == 0) {
            return fo;
        } else {
            vint<P> side;
#pragma unroll
            for (int p = 0; p < P; ++p) side.p[p] = dx < 0 ? join(0u, r.l[p]) : join(r.r[p], 0u);
            const word fs = f(side);
            return shift1<dx>(fo, hi32(fs), lo32(fs));
        }
    }
};

// Row sums of a tiled ring slot: per counted state, the 3-cell horizontal sum of its mask for
// each cell of the 8x8 block (sum bit, carry), formed once when the block enters the ring.
template <int NS>
struct tsums {
    word s[NS > 0 ? NS : 1];
    word c[NS > 0 ? NS : 1];
};

// The cell to the left (right) of every cell of a tiled word (8 cells per byte row): the byte
// rows shifted by one bit, the edge column from the block to the left (right). Per 32-bit
// half: two shifts and one three-input select (truth table 0xE4 = c ? a : b) with the column
// mask, where the compiler made 64-bit shifts and four mask ops of the same expression.
__device__ __forceinline__ word tiled_left(word own, word left) {
    return join(__builtin_amdgcn_bitop3_b32(lo32(own) << 1, lo32(left) >> 7, 0xFEFEFEFEu, 0xE4),
                __builtin_amdgcn_bitop3_b32(hi32(own) << 1, hi32(left) >> 7, 0xFEFEFEFEu, 0xE4));
}
__device__ __forceinline__ word tiled_right(word own, word right) {
    return join(__builtin_amdgcn_bitop3_b32(lo32(own) >> 1, lo32(right) << 7, 0x7F7F7F7Fu, 0xE4),
                __builtin_amdgcn_bitop3_b32(hi32(own) >> 1, hi32(right) << 7, 0x7F7F7F7Fu, 0xE4));
}

// The edge word of a tiled block: column 0 holds column 7 of the block to the left, column 7
// holds column 0 of the block to the right (the other columns are unused). Formed once when a
// block enters a ring slot, from the neighbouring lanes' words; every later read of a
// horizontal neighbour combines the block with it (three-input select c ? a : b per half).
__device__ __forceinline__ word tiled_edges(word left, word right) {
    return join(__builtin_amdgcn_bitop3_b32(lo32(left) >> 7, lo32(right) << 7, 0x01010101u, 0xE4),
                __builtin_amdgcn_bitop3_b32(hi32(left) >> 7, hi32(right) << 7, 0x01010101u, 0xE4));
}
// the cell to the left / right of every cell of a block, from the block and its edge word
__device__ __forceinline__ word tiled_left_e(word own, word e) {
    return join(__builtin_amdgcn_bitop3_b32(lo32(own) << 1, lo32(e), 0xFEFEFEFEu, 0xE4),
                __builtin_amdgcn_bitop3_b32(hi32(own) << 1, hi32(e), 0xFEFEFEFEu, 0xE4));
}
__device__ __forceinline__ word tiled_right_e(word own, word e) {
    return join(__builtin_amdgcn_bitop3_b32(lo32(own) >> 1, lo32(e), 0x7F7F7F7Fu, 0xE4),
                __builtin_amdgcn_bitop3_b32(hi32(own) >> 1, hi32(e), 0x7F7F7F7Fu, 0xE4));
}

// Whether the tiled ring slots of a rule carry edge words (formed once per block and stage)
// or the kernel exchanges neighbouring words between lanes where a horizontal neighbour is
// read. Edge words pay for rules that read neighbours beyond the Moore counts of fixed states
// (those come from row sums, which exchange one mask per counted state): Margolus blocks,
// direct neighbour reads, von Neumann counts, counts with a per-cell target. They cost 2 VGPRs
// per plane and ring slot, so only while P x T <= 10. Measured at 16384^2 per launch: Critters
// T=4 45.4 -> 40.0 us, T=8 96.2 -> 83.3; Cyclic T=2 133.6 -> 116.7 but T=4 241.8 -> 344.9;
// Game of Life and WireWorld slower with them at every depth.
template <class Eval>
constexpr bool tiled_edge_words(int T) {
    using namespace cellato::evaluators::bit_planes;
    constexpr unsigned reads = TRAIT_VON_NEUMANN | TRAIT_MARGOLUS | TRAIT_NEIGHBOR | TRAIT_DYNAMIC_TARGET;
    return (rule_traits<typename Eval::expression>::value & reads) != 0 && Eval::planes * T <= 10;
}

// Tiled layout: a word is an 8x8 block; cell (x, y) at bit (y % 8) * 8 + x % 8. Each ring slot
// holds the block's P plane words and, with EDGES, their P edge words; without, horizontal
// neighbours come from the neighbouring lanes where they are read (`halo`: the block outside
// the segment end, for end lanes that load it).
template <int P, class Targets, int X, bool HALO, bool EDGES>
struct tiled_window {
    static constexpr int NS = Targets::size;
    static constexpr bool row_sums = NS > 0;
    const row_t<P>& top;
    const row_t<P>& mid;
    const row_t<P>& bot;
    const row_t<P>& etop;   // edge words (EDGES) or the halo column's words (HALO)
    const row_t<P>& emid;
    const row_t<P>& ebot;
    xch<X> x;
    int t;
    const tsums<NS>& stop;
    const tsums<NS>& smid;
    const tsums<NS>& sbot;

    __device__ word yodd() const { return 0xFF00FF00FF00FF00ULL; }   // rows 1, 3, 5, 7 of a block

    // the block of sums shifted by one cell row (dy = -1: each cell gets the row above)
    template <int dy>
    __device__ static word vsum(word t_, word m_, word b_) {
        if constexpr (dy == 0) return m_;
        else if constexpr (dy < 0)
            return join(__builtin_amdgcn_alignbit(lo32(m_), hi32(t_), 24), __builtin_amdgcn_alignbit(hi32(m_), lo32(m_), 24));
        else
            return join(__builtin_amdgcn_alignbit(hi32(m_), lo32(m_), 8), __builtin_amdgcn_alignbit(lo32(b_), hi32(m_), 8));
    }

    template <long long K, int dy>
    __device__ void rsum(word& s, word& c) const {
        constexpr int j = kindex<K, Targets>::value;
        static_assert(j >= 0, "row sums are kept for the rule's counted states only");
        s = vsum<dy>(stop.s[j], smid.s[j], sbot.s[j]);
        c = vsum<dy>(stop.c[j], smid.c[j], sbot.c[j]);
    }

    // the block shifted by one cell row, filled from the block above / below
    template <int dy>
    __device__ static word vshift(const row_t<P>& t_, const row_t<P>& m_, const row_t<P>& b_, int p) {
        if constexpr (dy == 0) return m_[p];
        else if constexpr (dy < 0) return (m_[p] << 8) | (t_[p] >> 56);
        else return (m_[p] >> 8) | (b_[p] << 56);
    }

    // e: the vertically shifted edge word (EDGES) or halo word (HALO) of the same plane
    template <int dx>
    __device__ word hshift(word own, word e) const {
        if constexpr (dx == 0) {
            return own;
        } else if constexpr (EDGES) {
            return dx < 0 ? tiled_left_e(own, e) : tiled_right_e(own, e);
        } else if constexpr (dx < 0) {
            return tiled_left(own, join(x.left(lo32(own), HALO ? lo32(e) : 0u), x.left(hi32(own), HALO ? hi32(e) : 0u)));
        } else {
            return tiled_right(own, join(x.right(lo32(own), HALO ? lo32(e) : 0u), x.right(hi32(own), HALO ? hi32(e) : 0u)));
        }
    }

    static constexpr bool SIDE = EDGES || HALO;   // the second set of slot words is read

    template <int dx, int dy>
    __device__ vint<P> nb() const {
        vint<P> r;
#pragma unroll
        for (int p = 0; p < P; ++p)
            r.p[p] = hshift<dx>(vshift<dy>(top, mid, bot, p),
                                dx != 0 && SIDE ? vshift<dy>(etop, emid, ebot, p) : word(0));
        return r;
    }

    // F is per cell, so F of the edge (halo) words gives F of the neighbouring blocks' cells
    template <int dx, int dy, class F>
    __device__ word nbf(F f) const {
        vint<P> own, side;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            own.p[p] = vshift<dy>(top, mid, bot, p);
            side.p[p] = dx != 0 && SIDE ? vshift<dy>(etop, emid, ebot, p) : word(0);
        }
        if constexpr (dx == 0) return f(own);
        else return hshift<dx>(f(own), SIDE ? f(side) : word(0));
    }
};

__device__ __forceinline__ int wrap(int v, int m) {
    int r = v % m;
    return r < 0 ? r + m : r;
}

// Plane access through buffer resources with 32-bit offsets: a row load of a whole-segment
// wave takes the row offset as a scalar (SGPR) operand and the column offset as a per-lane
// constant (no address arithmetic); lane groups on different rows, stores and the outside
// columns pass one per-lane offset (one add). An offset past the plane reads 0 instead of
// faulting. Planes are < 4 GiB (checked on the host).
constexpr int BUFFER_DWORD3 = 0x00020000;   // raw, untyped 32-bit-element access on gfx9
// An offset past every plane (planes hold at most 2^32 - 8 bytes): a load there reads 0 and a
// store there is dropped. Lanes that have nothing to load or store issue the instruction with
// it instead of branching around it, so that the compiler's memory wait counts stay exact
// (a load or store under a branch makes it wait for the whole prefetch queue).
constexpr uint32_t OOB_OFFSET = 0xFFFFFFF8u;
// The same for the 16-byte accesses of TILED_ROWS, which add offsets up to 48: those planes are
// at most OOB_WIDE bytes (checked on the host), so OOB_WIDE + 63 stays past them.
constexpr uint32_t OOB_WIDE = 0xFFFFFF00u;

using u32x4 = uint32_t __attribute__((ext_vector_type(4)));

// 4x4 byte transpose: o[i] byte k = a[k] byte i (v_perm_b32: bytes 0-3 of the second operand,
// 4-7 of the first)
__device__ __forceinline__ void transpose4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t (&o)[4]) {
    const uint32_t t0 = __builtin_amdgcn_perm(a1, a0, 0x06020400u);   // a0.0 a1.0 a0.2 a1.2
    const uint32_t t1 = __builtin_amdgcn_perm(a1, a0, 0x07030501u);   // a0.1 a1.1 a0.3 a1.3
    const uint32_t t2 = __builtin_amdgcn_perm(a3, a2, 0x06020400u);
    const uint32_t t3 = __builtin_amdgcn_perm(a3, a2, 0x07030501u);
    o[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
    o[1] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
    o[2] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
    o[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}
// 8x8 byte transpose of 8 words: r[i] byte j = w[j] byte i (its own inverse). For a lane's 8
// block words (byte i = cell row i of the block) it gives the 8 row words of its 64-cell column.
__device__ __forceinline__ void transpose8(const word (&w)[8], word (&r)[8]) {
    uint32_t a[4], b[4], c[4], d[4];
    transpose4(lo32(w[0]), lo32(w[1]), lo32(w[2]), lo32(w[3]), a);
    transpose4(lo32(w[4]), lo32(w[5]), lo32(w[6]), lo32(w[7]), b);
    transpose4(hi32(w[0]), hi32(w[1]), hi32(w[2]), hi32(w[3]), c);
    transpose4(hi32(w[4]), hi32(w[5]), hi32(w[6]), hi32(w[7]), d);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        r[i] = join(a[i], b[i]);
        r[4 + i] = join(c[i], d[i]);
    }
}

// The LDS staging areas of TILED_ROWS: [in / out][wave][plane][row of the block][lane]
template <int P>
__device__ __forceinline__ word* rows_stage() {
    __shared__ word stage[2][WAVES_PER_BLOCK][P][8][64];
    return &stage[0][0][0][0][0];
}
// x + d onto [0, m) for x in [0, m), 0 <= d <= 2m
__device__ __forceinline__ int wrap_add(int x, int d, int m) {
    int y = x + d;
    y -= y >= m ? m : 0;
    return y - (y >= m ? m : 0);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const word* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<word*>(p), static_cast<short>(0), static_cast<int>(bytes), BUFFER_DWORD3);
}
__device__ __forceinline__ word buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(word, __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(voff), static_cast<int>(soff), 0));
}
// AUX: cache-policy bits of the store (gfx9.4+: 1 sc0, 2 nt, 16 sc1)
template <int AUX = 0>
__device__ __forceinline__ void buf_store(word v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                          r, static_cast<int>(voff), static_cast<int>(soff), AUX);
}

// Rows per unrolled loop body: a multiple of 3 so that the 3-row windows rotate by renaming
// registers instead of moving them.
constexpr int unroll_rows(int T) { return T == 1 ? 12 : T == 2 ? 6 : 3; }
// Rows loaded ahead of use. At the memory-bound depths a loaded row then lives in one register
// for unroll_rows - prefetch_rows steps in flight and 3 in the window, so registers rotate by
// renaming with no copies; a copy would make the wait for its load come early. Measured at
// 16384^2: for one generation, 12 rows per body with 6 ahead beats 6 / 3 (Game of Life 12.2 ->
// 11.9 us per launch); at depth 2, 6 / 3 beats 9 / 6 for every rule but Critters.
constexpr int prefetch_rows(int T) { return T == 1 ? 6 : T == 2 ? 3 : unroll_rows(T); }
// Cache policy of the stores (gfx9.4+ bits: 1 sc0, 2 nt, 16 sc1). sc1 = device scope: the line
// is written through the XCD's L2, so that no dirty lines wait for the end-of-kernel write-back
// that makes them visible to the other XCDs' next launch. Measured at 16384^2 per launch: one
// generation, Game of Life 12.2 -> 11.6 us, WireWorld 21.7 -> 21.4 us; Game of Life T=4 / T=8
// 6 % faster, Critters 2-7 %, WireWorld 1-2 % (T=2 3 % slower), HPP within 2 %. nt (streaming)
// stores bypass the Infinity Cache that the next launch reads from: 30-50 % slower.
constexpr int store_aux(int) { return 16; }
// TILED_ROWS: rows loaded ahead of use (from LDS, so one is enough), chosen per depth to
// maximise the steps between the loads of a block row (issued with the stores of the block row
// completed at output step 2T - 1 mod 8) and its transpose into the slots (at the load of its
// first row, input step T mod 8, made D steps earlier): 1 - T - D mod 8 steps, 8 for 0.
constexpr int rows_lead(int T, int D) { return ((1 - T - D) % 8 + 16) % 8 == 0 ? 8 : ((1 - T - D) % 8 + 16) % 8; }
constexpr int rows_prefetch(int T) {
    int best = 1;
    for (int d = 1; d <= unroll_rows(T); ++d)
        if (rows_lead(T, d) > rows_lead(T, best)) best = d;
    return best;
}

// Row sums of a linear ring slot: for each counted state K (dictionary index), the mask of
// the slot's row and its two 1-cell shifts, added as a full adder. Called once per row and
// stage, when the row enters the ring. The neighbouring columns' edge bits of the mask come
// from the neighbouring lanes' masks (one lane exchange per direction and counted state,
// instead of one per plane followed by the mask of the exchanged halves); end lanes that load
// the outside column (HALO) take its mask instead. A state the rule reads only as "some
// neighbour is K" (Eval::row_sum_or_only) keeps the 3-cell OR in the sum word and 0 in the
// carry word: one op instead of two, and a carry word that is a constant, not a register. A
// state read only as "some von Neumann neighbour is K" (Eval::row_sum_vn_any) keeps the 3-cell
// OR and the mask itself, the mask of the cells above and below for the rows around.
template <class Eval, class Targets>
struct row_sum_fill;
template <class Eval, long long... K>
struct row_sum_fill<Eval, klist<K...>> {
    template <bool HALO, int P, int NS, class XC>
    __device__ __forceinline__ static void run(lrow<P, NS>& r, const XC& xc, const word (&halo)[P]) {
        (one<K, kindex<K, klist<K...>>::value, HALO>(r, xc, halo), ...);
    }
    template <long long Kj, int j, bool HALO, int P, int NS, class XC>
    __device__ __forceinline__ static void one(lrow<P, NS>& r, const XC& xc, const word (&halo)[P]) {
        using namespace cellato::evaluators::bit_planes;
        const is_index<Kj> f{};
        vint<P> own;
#pragma unroll
        for (int p = 0; p < P; ++p) own.p[p] = r.w[p];
        const word m = f(own);
        uint32_t lh, rl;
        if constexpr (HALO) {
            vint<P> hv;
#pragma unroll
            for (int p = 0; p < P; ++p) hv.p[p] = halo[p];
            const word mh = f(hv);
            lh = xc.left(hi32(m), hi32(mh));
            rl = xc.right(lo32(m), lo32(mh));
        } else {
            lh = xc.left0(hi32(m));
            rl = xc.right0(lo32(m));
        }
        if constexpr (Eval::template row_sum_vn_any<Kj>) {
            r.rs_s[j] = lop3<0xFE>(shift1<-1>(m, lh, rl), m, shift1<1>(m, lh, rl));
            r.rs_c[j] = m;
        } else if constexpr (Eval::template row_sum_or_only<Kj>) {
            r.rs_s[j] = lop3<0xFE>(shift1<-1>(m, lh, rl), m, shift1<1>(m, lh, rl));
            r.rs_c[j] = 0;
        } else {
            fa(shift1<-1>(m, lh, rl), m, shift1<1>(m, lh, rl), r.rs_s[j], r.rs_c[j]);
        }
    }
};

// Row sums of a tiled ring slot: the counted state's mask of the block and of the neighbouring
// blocks' edge cells, shifted by one cell within each 8-cell block row, added as a full adder.
// The neighbours' masks come from the slot's edge words (EDGES), or from the neighbouring
// lanes (two lane exchanges of one word per state; `side` = the block outside the segment end
// for end lanes that load it, HALO).
template <class Targets>
struct tiled_sum_fill;
template <long long... K>
struct tiled_sum_fill<klist<K...>> {
    template <bool EDGES, bool HALO, int P, int NS, class XC>
    __device__ __forceinline__ static void run(const word (&w)[P], const word (&side)[P], const XC& xc, tsums<NS>& out) {
        (one<K, kindex<K, klist<K...>>::value, EDGES, HALO>(w, side, xc, out), ...);
    }
    template <long long Kj, int j, bool EDGES, bool HALO, int P, int NS, class XC>
    __device__ __forceinline__ static void one(const word (&w)[P], const word (&side)[P], const XC& xc, tsums<NS>& out) {
        using namespace cellato::evaluators::bit_planes;
        const is_index<Kj> f{};
        vint<P> own, sv;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            own.p[p] = w[p];
            sv.p[p] = side[p];
        }
        const word m = f(own);
        word at_left, at_right;   // cells x - 1 and x + 1
        if constexpr (EDGES) {
            const word me = f(sv);
            at_left = tiled_left_e(m, me);
            at_right = tiled_right_e(m, me);
        } else {
            word ml, mr;   // the mask of the block to the left / right
            if constexpr (HALO) {
                const word mh = f(sv);
                ml = join(xc.left(lo32(m), lo32(mh)), xc.left(hi32(m), hi32(mh)));
                mr = join(xc.right(lo32(m), lo32(mh)), xc.right(hi32(m), hi32(mh)));
            } else {
                ml = join(xc.left0(lo32(m)), xc.left0(hi32(m)));
                mr = join(xc.right0(lo32(m)), xc.right0(hi32(m)));
            }
            at_left = tiled_left(m, ml);
            at_right = tiled_right(m, mr);
        }
        fa(at_left, m, at_right, out.s[j], out.c[j]);
    }
};

// ---- one task: a band of columns x a chunk of rows, T generations --------------------------
// TPAR: -1, or the parity of a.t0 known at compile time; the windows then carry only the
// parity of each stage's step, and a rule that reads it (alternation, Margolus blocks) keeps
// one of its two forms.
// UR / PF: rows per unrolled loop body and rows loaded ahead (PF <= UR); SAUX: cache-policy
// bits of the stores.
template <class Eval, int LAYOUT, int T, int X, int HALO, int TPAR = -1, int UR = unroll_rows(T), int PF = prefetch_rows(T),
          int SAUX = store_aux(T)>
__device__ __forceinline__ void stream_task(const pass_args& a, int task) {
    constexpr int P = Eval::planes;
    constexpr bool LOADS_HALO = HALO == HALO_LOAD;
    constexpr bool ROWS = LAYOUT == TILED_ROWS;     // tiled planes processed as rows
    constexpr bool LIN = LAYOUT != TILED;           // the linear layout's processing
    static_assert(!(ROWS && LOADS_HALO), "tiled rows are for the temporal traverser (halo lanes or rings)");
    // lane groups walk different rows: rows narrower than a wave, and the halo-lane tail
    constexpr bool GROUPED = X == XCH_PERMUTE || HALO == HALO_LANES;
    constexpr int U = UR;
    constexpr int D = ROWS ? rows_prefetch(T) : PF;
    static_assert(UR % 3 == 0 && PF >= 1 && PF <= UR, "rows per body: a multiple of 3; prefetch at most one body");
    static_assert(!(HALO == HALO_LOAD && T != 1), "loaded halos are exact for one generation only");

    const int lane = lane_id();
    const int seg = task % a.nseg;
    const int chunk = task / a.nseg;

    // lane -> (group, column)
    const int g = GROUPED ? lane / a.n : 0;
    const int w = lane - g * a.n;
    const bool lane_used = g < a.groups;
    int col;                  // this lane's word column
    bool owner;               // whether this lane stores its column
    int halo_col = 0;         // the outside column of an end lane (HALO_LOAD)
    bool end_lane = false;
    xch<X> xc{};
    if constexpr (HALO == HALO_LANES) {
        // groups overlap their neighbours by two columns; lanes 0 and n-1 of a group are halo lanes
        col = wrap(a.first + seg * a.stride + w, a.wx);
        owner = lane_used && w >= 1 && w <= a.n - 2;
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

    const uint32_t row_bytes = static_cast<uint32_t>(a.wx) * 8u;
    const uint32_t plane_bytes = row_bytes * static_cast<uint32_t>(a.h);
    __amdgpu_buffer_rsrc_t rin[P], rout[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        rin[p] = plane_rsrc(a.in[p], plane_bytes);
        rout[p] = plane_rsrc(a.out[p], plane_bytes);
    }
    // a linear row is wx words; in TILED_ROWS a lane's column is a span of 8 block words (64
    // bytes) and a block row holds 8 cell rows (8 x row_bytes)
    const uint32_t col_off = static_cast<uint32_t>(col) * (ROWS ? 64u : 8u);
    // the end lanes' outside column; the other lanes read past the plane. The row offset is
    // added with saturation and capped at OOB_OFFSET: no per-row select, which the compiler
    // had turned into a branch around the load in every row.
    const uint32_t halo_base = end_lane ? static_cast<uint32_t>(halo_col) * 8u : OOB_OFFSET;

    // generation-k rows, a ring of 3 slots per stage (linear rows carry their neighbours'
    // edge halves; tiled rows exchange whole words when used)
    using targets = typename Eval::row_sum_targets;
    constexpr int NS = targets::size;
    constexpr bool EDGES = LAYOUT == TILED && tiled_edge_words<Eval>(T);
    lrow<P, NS> lwin[T][3];
    word win[T][3][P];
    word wedge[T][3][P];   // tiled ring slots: edge words (EDGES), or the halo column of stage 0
    tsums<NS> tsum[T][3];  // row sums of the tiled ring slots
    word pf[U][P];         // prefetched generation-0 rows
    word hpf[U][P];
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int p = 0; p < P; ++p) {
                win[k][i][p] = 0;
                wedge[k][i][p] = 0;
                lwin[k][i].w[p] = 0;
                lwin[k][i].l[p] = 0;
                lwin[k][i].r[p] = 0;
            }
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < (NS > 0 ? NS : 1); ++j) {
                lwin[k][i].rs_s[j] = 0;
                lwin[k][i].rs_c[j] = 0;
            }
#pragma unroll
    for (int k = 0; k < T; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < (NS > 0 ? NS : 1); ++j) {
                tsum[k][i].s[j] = 0;
                tsum[k][i].c[j] = 0;
            }

    // Step i loads row s = r0 - T + i; every group runs the same number of steps, so the
    // loop is uniform across the wave even when groups walk different rows.
    const int nsteps = a.chunk + 2 * T;

    // row of the next load, kept wrapped onto [0, h) incrementally (no division in the loop).
    // The prefetch of the last body reaches past the task's rows (steps >= nsteps): those loads
    // go to the offset past the plane (they read 0 and move no data); no step uses them.
    int lrow_next = wrap(r0 - T, a.h);
    int lstep = 0;
    // TILED_ROWS: the block words of the next block row to stage, loaded ahead
    uint32_t blk[ROWS ? P : 1][16];
    word* stage_in = nullptr;    // this wave's slots: input rows, output rows
    word* stage_out = nullptr;
    // block row starting at row `first` (a multiple of 8) into blk, or nothing
    auto load_block = [&](int first, bool want) {
        const uint32_t off = want ? static_cast<uint32_t>(first >> 3) * (row_bytes * 8u) + col_off : OOB_WIDE;
#pragma unroll
        for (int p = 0; p < (ROWS ? P : 0); ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rin[p], static_cast<int>(off), 16 * q, 0);
#pragma unroll
                for (int e = 0; e < 4; ++e) blk[p][4 * q + e] = v[e];
            }
    };
    auto stage_block = [&]() {   // blk -> its 8 rows, into the input slots
#pragma unroll
        for (int p = 0; p < (ROWS ? P : 0); ++p) {
            word w[8], r[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = join(blk[p][2 * j], blk[p][2 * j + 1]);
            transpose8(w, r);
#pragma unroll
            for (int i = 0; i < 8; ++i) stage_in[(p * 8 + i) * 64 + lane] = r[i];
        }
    };
    // Call c of load_row reads row (c - T) & 7 of its block row (r0 and h are multiples of 8:
    // the same in every lane group); the call reading row 0 (c = 8 - K0 + 8m, made at step
    // c - D) first stages the block row. The loads for a staging are issued at the end of the
    // last output step 2T - 1 + 8m before it, rows_lead(T, D) steps earlier.
    constexpr int K0 = (8 - T % 8) % 8;              // row of call 0 in its block row
    constexpr int S1 = 8 - K0 - D;                   // step of the first staging after call 0's
    constexpr int C0 = (2 * T + 7) % 8;              // first output step that stores a block row
    constexpr int LEAD = rows_lead(T, D);
    if constexpr (ROWS) {
        stage_in = rows_stage<P>() + (static_cast<int>(threadIdx.x) / 64) * (P * 512);
        stage_out = stage_in + WAVES_PER_BLOCK * P * 512;
        const int first = lrow_next - K0;
        load_block(first, true);
        stage_block();
        // the block row of the first staging, when it comes before the first output block
        if constexpr (S1 <= C0) load_block(wrap_add(first, 8, a.h), 8 - K0 < nsteps);
        // stores after the loads, as in the loop, so that the wait for blk is the same on
        // every path into it (these go past the plane)
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                __builtin_amdgcn_raw_buffer_store_b128(u32x4{0u, 0u, 0u, 0u}, rout[p], static_cast<int>(OOB_WIDE), 16 * q, 0);
    }
    auto load_row = [&](word (&dst)[P], word (&hdst)[P]) {
        const bool live = lstep < nsteps;   // uniform across the wave
        if constexpr (ROWS) {
            const int k = (lstep - T) & 7;
            if (lstep > 0 && k == 0) stage_block();
#pragma unroll
            for (int p = 0; p < P; ++p) {
                dst[p] = stage_in[(p * 8 + k) * 64 + lane];
                hdst[p] = 0;
            }
            lrow_next = lrow_next + 1 == a.h ? 0 : lrow_next + 1;
            ++lstep;
            return;
        }
        const uint32_t roff = static_cast<uint32_t>(lrow_next) * row_bytes;
        const uint32_t coff = live ? col_off : OOB_OFFSET;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if constexpr (GROUPED) dst[p] = buf_load(rin[p], live ? roff + col_off : OOB_OFFSET, 0u);
            else dst[p] = buf_load(rin[p], coff, roff);
            hdst[p] = 0;
        }
        if constexpr (LOADS_HALO) {
            const uint32_t hoff = live ? __builtin_elementwise_min(__builtin_elementwise_add_sat(roff, halo_base), OOB_OFFSET)
                                       : OOB_OFFSET;
#pragma unroll
            for (int p = 0; p < P; ++p) hdst[p] = buf_load(rin[p], hoff, 0u);
        }
        lrow_next = lrow_next + 1 == a.h ? 0 : lrow_next + 1;
        ++lstep;
    };
#pragma unroll
    for (int u = 0; u < D; ++u) load_row(pf[u], hpf[u]);
    int ycur = wrap(r0 - T - 2, a.h);   // advanced to row s - 1 (the row stage 0 computes) at each step
    int kmod[T];
#pragma unroll
    for (int k = 0; k < T; ++k) kmod[k] = k % a.h;
    // parity of stage k's row y = ycur - k mod h: that of ycur, flipped for odd k mod h, and
    // flipped again where the subtraction wraps (ycur < k mod h) on a torus of odd height
    const bool hodd = (a.h & 1) != 0;
    uint32_t kflip[T];
#pragma unroll
    for (int k = 0; k < T; ++k) kflip[k] = (kmod[k] & 1) ? ~0u : 0u;

    // One block of U steps. FILL: the block may fall in the pipeline fill (steps below 2T),
    // where stages that no stored row needs are skipped.
    auto block = [&](const int ib, auto fill_tag) {
        constexpr bool FILL = decltype(fill_tag)::value;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = r0 - T + ib + u;
            const int nw = u % 3, tp = (u + 1) % 3, md = (u + 2) % 3;   // slots: newest, y-1, y
            ycur = ycur + 1 == a.h ? 0 : ycur + 1;
            const uint32_t ypar = (ycur & 1) ? ~0u : 0u;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if constexpr (LIN) {
                    const word v = pf[u][p];
                    lwin[0][nw].w[p] = v;
                    if constexpr (LOADS_HALO) {
                        lwin[0][nw].l[p] = xc.left(hi32(v), hi32(hpf[u][p]));
                        lwin[0][nw].r[p] = xc.right(lo32(v), lo32(hpf[u][p]));
                    } else {
                        lwin[0][nw].l[p] = xc.left0(hi32(v));
                        lwin[0][nw].r[p] = xc.right0(lo32(v));
                    }
                } else {
                    const word v = pf[u][p];
                    win[0][nw][p] = v;
                    if constexpr (EDGES) {
                        word l, r;   // the blocks to the left and right (the outside block for end lanes)
                        if constexpr (LOADS_HALO) {
                            const word hv = hpf[u][p];
                            l = join(xc.left(lo32(v), lo32(hv)), xc.left(hi32(v), hi32(hv)));
                            r = join(xc.right(lo32(v), lo32(hv)), xc.right(hi32(v), hi32(hv)));
                        } else {
                            l = join(xc.left0(lo32(v)), xc.left0(hi32(v)));
                            r = join(xc.right0(lo32(v)), xc.right0(hi32(v)));
                        }
                        wedge[0][nw][p] = tiled_edges(l, r);
                    } else if constexpr (LOADS_HALO) {
                        wedge[0][nw][p] = hpf[u][p];
                    }
                }
            }
            if constexpr (LIN && NS > 0) row_sum_fill<Eval, targets>::template run<LOADS_HALO>(lwin[0][nw], xc, hpf[u]);
            if constexpr (LAYOUT == TILED && NS > 0)
                tiled_sum_fill<targets>::template run<EDGES, LOADS_HALO>(win[0][nw], wedge[0][nw], xc, tsum[0][nw]);
            // the load stays after this step's use of the window (hoisted above it, it would
            // overlap the register it refills with a live row and force a copy)
            if constexpr (T <= 2) __builtin_amdgcn_sched_barrier(0);
            load_row(pf[(u + D) % U], hpf[(u + D) % U]);   // the row of step ib + u + D

            // stage k: generation k+1 of row s-1-k from generation-k rows s-2-k .. s-k. The
            // stored rows [r0, r1) need stage k from row r0 - (T-1-k) on, i.e. from step 2k + 2;
            // before that (the pipeline fill) stage k is skipped: what it would write to stage
            // k+1's ring is read only by stage k+1 steps that are skipped too.
            vint<P> out;
#pragma unroll
            for (int k = 0; k < T; ++k) {
                if (FILL && ib + u < 2 * k + 2) continue;
                const int t = TPAR < 0 ? a.t0 + k : (TPAR + k) & 1;
                if constexpr (LIN) {
                    // the parity of the row of stage k (only rules with a Margolus block read
                    // it): one compare and one select per stage (forming the row, y - k wrapped,
                    // took five more ops per word and stage)
                    const uint32_t yk = ypar ^ kflip[k] ^ ((hodd & (ycur < kmod[k])) ? ~0u : 0u);
                    const linear_window<P, targets, typename Eval::vn_sum_targets> wnd{lwin[k][tp], lwin[k][md], lwin[k][nw],
                                                                                        join(yk, yk), t};
                    out = Eval::evaluate(wnd);
                } else {
                    // without edge words, the halo column (HALO_LOAD) is kept in stage 0's slots only
                    const int ks = EDGES ? k : 0;
                    const tiled_window<P, targets, X, LOADS_HALO, EDGES> wnd{
                        win[k][tp], win[k][md], win[k][nw], wedge[ks][tp], wedge[ks][md], wedge[ks][nw], xc, t,
                        tsum[k][tp], tsum[k][md], tsum[k][nw]};
                    out = Eval::evaluate(wnd);
                }
                if (k + 1 < T) {
                    const int kn = k + 1 < T ? k + 1 : 0;
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        if constexpr (LIN) {
                            lwin[kn][nw].w[p] = out.p[p];
                            lwin[kn][nw].l[p] = xc.left0(hi32(out.p[p]));
                            lwin[kn][nw].r[p] = xc.right0(lo32(out.p[p]));
                        } else {
                            const word v = out.p[p];
                            win[kn][nw][p] = v;
                            if constexpr (EDGES)
                                wedge[kn][nw][p] = tiled_edges(join(xc.left0(lo32(v)), xc.left0(hi32(v))),
                                                               join(xc.right0(lo32(v)), xc.right0(hi32(v))));
                        }
                    }
                    if constexpr (LIN && NS > 0) row_sum_fill<Eval, targets>::template run<false>(lwin[kn][nw], xc, hpf[u]);
                    if constexpr (LAYOUT == TILED && NS > 0)
                        tiled_sum_fill<targets>::template run<EDGES, false>(win[kn][nw], wedge[kn][nw], xc, tsum[kn][nw]);
                }
            }
            // generation T of row s - T (in [r0, r1) implies in [0, h))
            const int yo = s - T;
            const bool stored = (yo >= r0) & (yo < r1) & owner;   // no short circuit (see OOB_OFFSET)
            if constexpr (ROWS) {
                // output row o of its block row into slot o; after row 7 the loads of the next
                // block row to stage (the one whose row 0 is read at step + LEAD + D, i.e. row
                // yo + T + LEAD + D), then the block row's 8 words out
                const int o = (ib + u - 2 * T) & 7;
#pragma unroll
                for (int p = 0; p < P; ++p) stage_out[(p * 8 + o) * 64 + lane] = out.p[p];
                if (o == 7) {
                    load_block(wrap(yo + T + LEAD + D, a.h), ib + u + LEAD + D < nsteps);
                    const uint32_t boff = stored ? static_cast<uint32_t>(yo >> 3) * (row_bytes * 8u) + col_off : OOB_WIDE;
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        word r[8], w[8];
#pragma unroll
                        for (int i = 0; i < 8; ++i) r[i] = stage_out[(p * 8 + i) * 64 + lane];
                        transpose8(r, w);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const u32x4 v = {lo32(w[2 * q]), hi32(w[2 * q]), lo32(w[2 * q + 1]), hi32(w[2 * q + 1])};
                            __builtin_amdgcn_raw_buffer_store_b128(v, rout[p], static_cast<int>(boff), 16 * q, 0);
                        }
                    }
                }
            } else {
                const uint32_t soff = stored ? static_cast<uint32_t>(yo) * row_bytes + col_off : OOB_OFFSET;
#pragma unroll
                for (int p = 0; p < P; ++p) buf_store<SAUX>(out.p[p], rout[p], soff, 0u);
            }
        }
    };
    // The fill runs in a copy of the loop that tests each stage, the steady state in one
    // without the tests (the tests cost scalar issue slots and keep skipped ring slots alive
    // as extra registers).
    int ib = 0;
    for (; ib < 2 * T && ib < nsteps; ib += U) block(ib, std::true_type{});
    for (; ib < nsteps; ib += U) block(ib, std::false_type{});
}

// Blocks are dealt round-robin to the 8 XCDs; renumber them so each XCD (own L2) gets a
// contiguous range of tasks, i.e. neighbouring chunks and segments share its L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    const int q = nblocks / 8, rem = nblocks % 8;
    const int xcd = b % 8, local = b / 8;
    return xcd < rem ? xcd * (q + 1) + local : rem * (q + 1) + (xcd - rem) * q + local;
}

// One task of a pass whose rows are cut into 62-column halo-lane segments plus a tail of
// narrow lane groups. Both kinds run the same loop (one copy of the code): only the lane
// mapping differs, and it is set up here, uniformly across the wave.
template <class Eval, int LAYOUT, int T, int TPAR = -1, int SAUX = store_aux(T)>
__device__ __forceinline__ void segmented_task(const pass_args& a, int task) {
    pass_args b = a;
    int t = task;
    if (task >= a.main_tasks) {
        b.n = a.tail_n;
        b.groups = a.tail_groups;
        b.first = a.tail_first;
        b.stride = 0;
        b.nseg = 1;
        t = task - a.main_tasks;
    }
    stream_task<Eval, LAYOUT, T, XCH_SEGMENT, HALO_LANES, TPAR, unroll_rows(T), prefetch_rows(T), SAUX>(b, t);
}

template <class Eval, int LAYOUT, int T, int X, int HALO, int UR = unroll_rows(T), int PF = prefetch_rows(T), int SAUX = store_aux(T)>
__global__ void __launch_bounds__(BLOCK_THREADS) stream_kernel(pass_args a) {
    const int blk = xcd_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x));
    const int task = __builtin_amdgcn_readfirstlane(blk * WAVES_PER_BLOCK + static_cast<int>(threadIdx.x) / 64);
    if (task >= a.ntasks) return;
    if constexpr (Eval::reads_step_parity) {
        // one copy of the task per parity of the first step (a uniform branch: a launch
        // runs one of them)
        if (a.t0 & 1) {
            if constexpr (HALO == HALO_LANES) segmented_task<Eval, LAYOUT, T, 1, SAUX>(a, task);
            else stream_task<Eval, LAYOUT, T, X, HALO, 1, UR, PF, SAUX>(a, task);
        } else {
            if constexpr (HALO == HALO_LANES) segmented_task<Eval, LAYOUT, T, 0, SAUX>(a, task);
            else stream_task<Eval, LAYOUT, T, X, HALO, 0, UR, PF, SAUX>(a, task);
        }
    } else if constexpr (HALO == HALO_LANES) {
        segmented_task<Eval, LAYOUT, T, -1, SAUX>(a, task);
    } else {
        stream_task<Eval, LAYOUT, T, X, HALO, -1, UR, PF, SAUX>(a, task);
    }
}

// ---- batches: many grids, one persistent launch per (rule, T) and round ---------------------
struct batch_task {
    int pass;    // index into the pass table
    int task;    // task of that pass (segment + chunk)
};

struct batch_args {
    const pass_args* passes;
    const batch_task* tasks;
    int ntasks;
};

template <class Eval, int LAYOUT, int T, int TPAR, int SAUX = store_aux(T)>
__device__ __forceinline__ void batch_task_run(const pass_args& a, int task) {
    if (a.variant == VARIANT_SEGMENT)
        segmented_task<Eval, LAYOUT, T, TPAR, SAUX>(a, task);
    else if (a.variant == VARIANT_RING64)
        stream_task<Eval, LAYOUT, T, XCH_RING64, HALO_NONE, TPAR, unroll_rows(T), prefetch_rows(T), SAUX>(a, task);
    else
        stream_task<Eval, LAYOUT, T, XCH_PERMUTE, HALO_NONE, TPAR, unroll_rows(T), prefetch_rows(T), SAUX>(a, task);
}

// The pass table entry of a task, copied into scalar registers (the entry is uniform across the
// wave). Read through a reference into global memory instead, its fields are reloaded by vector
// loads after every store of the loop (the stores might alias them): the row count on every
// row. Measured on the 512-grid batch at depth 6: 28.5k -> 33.6k GCUPS.
__device__ __forceinline__ pass_args load_pass(const pass_args* p) {
    static_assert(sizeof(pass_args) % 4 == 0, "pass_args is copied as 32-bit words");
    pass_args a;
    const int* src = reinterpret_cast<const int*>(p);
    int* dst = reinterpret_cast<int*>(&a);
#pragma unroll
    for (int q = 0; q < static_cast<int>(sizeof(pass_args) / 4); ++q) dst[q] = __builtin_amdgcn_readfirstlane(src[q]);
    return a;
}

// A fixed grid of waves walks a task list that mixes passes of different grids (sizes, lane
// mappings, start steps); each task runs exactly as in the single-grid kernel.
template <class Eval, int LAYOUT, int T>
__global__ void __launch_bounds__(BLOCK_THREADS) batch_kernel(batch_args b) {
    const int nwaves = static_cast<int>(gridDim.x) * WAVES_PER_BLOCK;
    const int first = __builtin_amdgcn_readfirstlane(xcd_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)) *
                                                     WAVES_PER_BLOCK + static_cast<int>(threadIdx.x) / 64);
    for (int i = first; i < b.ntasks; i += nwaves) {
        const batch_task t = b.tasks[i];
        const pass_args a = load_pass(b.passes + t.pass);
        if constexpr (Eval::reads_step_parity) {
            // passes of even depth that start at step 0 all start on even steps: only one of
            // the two copies runs
            if (a.t0 & 1) batch_task_run<Eval, LAYOUT, T, 1>(a, t.task);
            else batch_task_run<Eval, LAYOUT, T, 0>(a, t.task);
        } else {
            batch_task_run<Eval, LAYOUT, T, -1>(a, t.task);
        }
    }
}

} // namespace cellato::hip

#endif
