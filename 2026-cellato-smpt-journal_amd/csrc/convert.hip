// Conversions on the device between a grid of state values (int32, one per cell, the
// reference's standard grid) and bit planes, the reference's 16-tile checksum computed on
// the planes, and seeded i.i.d. initial states written straight into planes.
//
// Reference counterparts (CPU loops there): packing bit_planes_grid.hpp:234-270 and
// tiled_bit_planes_grid.hpp:243-288, unpacking via get_cell (:97-124 / :97-133), checksum
// standard_grid.hpp:179-213.
//
// One wave handles one 64-cell word at a time: lane l reads cell l of the word (a row
// segment in the linear layout, row l/8, column l%8 of the block in the tiled layout), and
// a wave-wide ballot of bit b of the lanes' dictionary indices is plane b's word.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "cellato_hip.h"
#include "dispatch.hpp"
#include "plan.hpp"

namespace cellato::hip {
namespace {

struct dict_args {
    int planes;
    int nstates;
    int value_of[32];     // enum value of each dictionary index
    int index_of[64];     // dictionary index of each enum value 0..63 (-1: not a state)
};

struct conv_args {
    word* planes[MAX_PLANES];
    int width, height;    // cells
    int layout;
    dict_args d;
};

// cell coordinates of lane `l` of word `w` (word index within a plane)
__device__ __forceinline__ void cell_of(const conv_args& a, long long w, int l, int& x, int& y) {
    if (a.layout == CELLATO_LAYOUT_LINEAR) {
        const int wx = a.width / 64;
        y = static_cast<int>(w / wx);
        x = static_cast<int>(w % wx) * 64 + l;
    } else {
        const int wx = a.width / 8;
        y = static_cast<int>(w / wx) * 8 + l / 8;
        x = static_cast<int>(w % wx) * 8 + l % 8;
    }
}

__global__ void __launch_bounds__(256) pack_kernel(conv_args a, const int32_t* values, int* bad) {
    const long long words = static_cast<long long>(a.width) * a.height / 64;
    const int lane = lane_id();
    for (long long w = (static_cast<long long>(blockIdx.x) * 256 + threadIdx.x) / 64; w < words;
         w += static_cast<long long>(gridDim.x) * 4) {
        int x, y;
        cell_of(a, w, lane, x, y);
        const int v = values[static_cast<size_t>(y) * a.width + x];
        const int idx = (v >= 0 && v < 64) ? a.d.index_of[v] : -1;
        if (idx < 0) atomicOr(bad, 1);
        for (int b = 0; b < a.d.planes; ++b) {
            const word bits = __ballot((idx >> b) & 1);
            if (lane == b) a.planes[b][w] = bits;
        }
    }
}

__global__ void __launch_bounds__(256) unpack_kernel(conv_args a, int32_t* values) {
    const long long words = static_cast<long long>(a.width) * a.height / 64;
    const int lane = lane_id();
    for (long long w = (static_cast<long long>(blockIdx.x) * 256 + threadIdx.x) / 64; w < words;
         w += static_cast<long long>(gridDim.x) * 4) {
        int x, y;
        cell_of(a, w, lane, x, y);
        int idx = 0;
        for (int b = 0; b < a.d.planes; ++b) idx |= static_cast<int>((a.planes[b][w] >> lane) & 1) << b;
        values[static_cast<size_t>(y) * a.width + x] = idx < a.d.nstates ? a.d.value_of[idx] : -1;
    }
}

// 16 tile sums of state values; tiles of ceil(W/4) x ceil(H/4) cells, row-major tile order
__global__ void __launch_bounds__(256) checksum_kernel(conv_args a, unsigned long long* out) {
    const long long words = static_cast<long long>(a.width) * a.height / 64;
    const int lane = lane_id();
    const int tw = (a.width + 3) / 4, th = (a.height + 3) / 4;
    unsigned long long acc[16];
    for (int i = 0; i < 16; ++i) acc[i] = 0;
    for (long long w = (static_cast<long long>(blockIdx.x) * 256 + threadIdx.x) / 64; w < words;
         w += static_cast<long long>(gridDim.x) * 4) {
        int x, y;
        cell_of(a, w, lane, x, y);
        int idx = 0;
        for (int b = 0; b < a.d.planes; ++b) idx |= static_cast<int>((a.planes[b][w] >> lane) & 1) << b;
        const unsigned long long v = static_cast<unsigned long long>(idx < a.d.nstates ? a.d.value_of[idx] : 0);
        const int t = (y / th) * 4 + x / tw;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] += i == t ? v : 0ULL;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        unsigned long long s = acc[i];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0 && s) atomicAdd(&out[i], s);
    }
}

// splitmix64 finaliser: a counter-based generator, the same on host and device
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct fill_args {
    conv_args c;
    uint64_t seed;
    uint32_t cum[33];     // cumulative thresholds over dictionary indices, scaled to 2^32 (last = 2^32 - 1)
    int ncat;
};

// cell (x, y) gets the first dictionary index k with u < cum[k], u = top 32 bits of
// mix64(seed ^ mix64(y * W + x)).
__global__ void __launch_bounds__(256) fill_kernel(fill_args f) {
    const conv_args& a = f.c;
    const long long words = static_cast<long long>(a.width) * a.height / 64;
    const int lane = lane_id();
    for (long long w = (static_cast<long long>(blockIdx.x) * 256 + threadIdx.x) / 64; w < words;
         w += static_cast<long long>(gridDim.x) * 4) {
        int x, y;
        cell_of(a, w, lane, x, y);
        const uint64_t cell = static_cast<uint64_t>(y) * static_cast<uint64_t>(a.width) + static_cast<uint64_t>(x);
        const uint32_t u = static_cast<uint32_t>(mix64(f.seed ^ mix64(cell)) >> 32);
        int idx = f.ncat - 1;
        for (int k = f.ncat - 2; k >= 0; --k)
            if (u < f.cum[k]) idx = k;
        for (int b = 0; b < a.d.planes; ++b) {
            const word bits = __ballot((idx >> b) & 1);
            if (lane == b) a.planes[b][w] = bits;
        }
    }
}

int cfail(int code, const std::string& m) {
    last_error() = m;
    return code;
}

int setup(int rule, int layout, uint64_t* const* planes, int width, int height, conv_args& a) {
    if (rule < 0 || rule >= 10) return cfail(CELLATO_E_RULE, "unknown rule id");
    if (layout != CELLATO_LAYOUT_LINEAR && layout != CELLATO_LAYOUT_TILED) return cfail(CELLATO_E_LAYOUT, "unknown layout");
    if (width <= 0 || height <= 0 || (layout == CELLATO_LAYOUT_LINEAR && width % 64) ||
        (layout == CELLATO_LAYOUT_TILED && (width % 8 || height % 8)))
        return cfail(CELLATO_E_GEOMETRY, "grid size not representable in the layout");
    if (!planes) return cfail(CELLATO_E_ARG, "null planes");
    const rule_kernels& k = kernels_of_rule(rule);
    a = conv_args{};
    a.width = width;
    a.height = height;
    a.layout = layout;
    a.d.planes = k.planes;
    a.d.nstates = k.nstates;
    for (int i = 0; i < 64; ++i) a.d.index_of[i] = -1;
    for (int i = 0; i < k.nstates; ++i) {
        a.d.value_of[i] = k.state_values[i];
        if (k.state_values[i] >= 0 && k.state_values[i] < 64) a.d.index_of[k.state_values[i]] = i;
    }
    for (int p = 0; p < k.planes; ++p) {
        if (!planes[p]) return cfail(CELLATO_E_ARG, "null plane pointer");
        a.planes[p] = planes[p];
    }
    return CELLATO_OK;
}

int blocks_for(int width, int height) {
    const long long words = static_cast<long long>(width) * height / 64;
    const long long b = (words + 3) / 4;
    return static_cast<int>(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

int launched(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cfail(CELLATO_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return CELLATO_OK;
}

} // namespace
} // namespace cellato::hip

using namespace cellato::hip;

extern "C" {

int cellato_hip_pack(int rule, int layout, const int32_t* values, int width, int height, uint64_t* const* planes,
                     int* bad_values, void* stream) {
    conv_args a;
    const int rc = setup(rule, layout, planes, width, height, a);
    if (rc != CELLATO_OK) return rc;
    if (!values || !bad_values) return cfail(CELLATO_E_ARG, "null values or flag");
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(width, height)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       a, values, bad_values);
    return launched("pack_kernel");
}

int cellato_hip_unpack(int rule, int layout, uint64_t* const* planes, int width, int height, int32_t* values,
                       void* stream) {
    conv_args a;
    const int rc = setup(rule, layout, planes, width, height, a);
    if (rc != CELLATO_OK) return rc;
    if (!values) return cfail(CELLATO_E_ARG, "null values");
    hipLaunchKernelGGL(unpack_kernel, dim3(blocks_for(width, height)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       a, values);
    return launched("unpack_kernel");
}

int cellato_hip_checksum(int rule, int layout, uint64_t* const* planes, int width, int height,
                         unsigned long long* sums16, void* stream) {
    conv_args a;
    const int rc = setup(rule, layout, planes, width, height, a);
    if (rc != CELLATO_OK) return rc;
    if (!sums16) return cfail(CELLATO_E_ARG, "null output");
    hipLaunchKernelGGL(checksum_kernel, dim3(blocks_for(width, height)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       a, sums16);
    return launched("checksum_kernel");
}

int cellato_hip_fill_random(int rule, int layout, uint64_t* const* planes, int width, int height,
                            unsigned long long seed, const double* probs, int nprobs, void* stream) {
    fill_args f{};
    const int rc = setup(rule, layout, planes, width, height, f.c);
    if (rc != CELLATO_OK) return rc;
    if (!probs || nprobs < 1 || nprobs > f.c.d.nstates) return cfail(CELLATO_E_ARG, "probabilities: one per dictionary index");
    double acc = 0.0, total = 0.0;
    for (int i = 0; i < nprobs; ++i) total += probs[i];
    if (!(total > 0.0)) return cfail(CELLATO_E_ARG, "probabilities must sum to a positive value");
    for (int i = 0; i < nprobs; ++i) {
        acc += probs[i] / total;
        const double t = acc * 4294967296.0;
        f.cum[i] = t >= 4294967295.0 ? 0xFFFFFFFFu : static_cast<uint32_t>(t);
    }
    f.cum[nprobs - 1] = 0xFFFFFFFFu;
    f.ncat = nprobs;
    f.seed = seed;
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(width, height)), dim3(256), 0, static_cast<hipStream_t>(stream), f);
    return launched("fill_kernel");
}

} // extern "C"
