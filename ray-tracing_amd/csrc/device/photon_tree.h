// Photon kd tree built on the device: the reference's median tree
// (photon_map.h:83-114) over positions already in HBM, with no host copies.
//
// Shape: a node of c > 8 photons splits into floor(c / 2) and c - floor(c / 2),
// so every node of level l holds floor(n / 2^l) or that + 1 photons and the
// shape, the pre-order numbering and the tree positions depend on n only.
// The node's axis is its set's largest extent (ties to the lower axis), its
// left child takes the floor(c / 2) photons first in (coordinate, index)
// order on that axis, and its split value is the coordinate of the first
// photon of the right child (the median).  That value separates the children
// exactly, so the k-nearest search over the tree is exact; the reference reads
// its split value after the right subtree was built, from whichever photon
// std::nth_element left in the median slot, which can lie above other photons
// of the right child and makes its search miss neighbours now and then.
//
// Build (presorted lists): the photon indices sorted by (coordinate, index) on
// each axis, once.  Every level partitions the three lists stably, segment by
// segment, by the side the chosen axis's order gives each photon, so every
// segment stays sorted on every axis: its bounding box is read from the ends
// of the three lists and its median is element floor(c / 2) of the chosen
// axis's list.  Per level: one thread per segment (axis, node record), one
// pass marking sides, a scan of the left flags, one stable scatter.  A leaf's
// photons take tree positions in x order.
#pragma once

#include "common.h"

namespace lrt {

// order-preserving 32-bit key of a float (-0 before +0)
LRT_HD uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// one level's segments: start position, photon count, pre-order node index,
// chosen axis
struct PkLevel {
    uint32_t *start, *count, *pre, *axis;
};

// sort keys of one axis: keys[i] = key of photon i's coordinate, ids[i] = i
__global__ void k_pk_keys(const float* __restrict__ xyz, uint32_t n, int axis, uint32_t* __restrict__ keys,
                          uint32_t* __restrict__ ids) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = float_key(xyz[3 * (size_t)i + axis]);
        ids[i] = i;
    }
}

LRT_D int pk_axis(const float* xyz, const uint32_t* lists, uint32_t n, uint32_t start, uint32_t c) {
    float e[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t* L = lists + (size_t)a * n;
        e[a] = xyz[3 * (size_t)L[start + c - 1] + a] - xyz[3 * (size_t)L[start] + a];
    }
    int axis = 0;
    if (e[1] > e[axis]) axis = 1;
    if (e[2] > e[axis]) axis = 2;
    return axis;
}

// interior level: per segment the axis, the split and the node record, and the
// next level's two segments (the left child's subtree has s_lo nodes when it
// holds a_next photons, s_hi when a_next + 1)
__global__ void k_pk_choose(const float* __restrict__ xyz, const uint32_t* __restrict__ lists, uint32_t n,
                            PkLevel cur, uint32_t m, PkLevel nxt, uint32_t a_next, uint32_t s_lo, uint32_t s_hi,
                            uint4* __restrict__ nodes) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        const uint32_t start = cur.start[j], c = cur.count[j], p = cur.pre[j], k = c / 2;
        const int axis = pk_axis(xyz, lists, n, start, c);
        const uint32_t med = lists[(size_t)axis * n + start + k];
        const float split = xyz[3 * (size_t)med + axis];
        const uint32_t left = p + 1, right = p + 1 + (k == a_next ? s_lo : s_hi);
        nodes[p] = make_uint4(__float_as_uint(split), (uint32_t)axis, left, right);
        cur.axis[j] = (uint32_t)axis;
        nxt.start[2 * j] = start;
        nxt.count[2 * j] = k;
        nxt.pre[2 * j] = left;
        nxt.start[2 * j + 1] = start + k;
        nxt.count[2 * j + 1] = c - k;
        nxt.pre[2 * j + 1] = right;
    }
}

// side[photon] = 0 (left child) or 1 from its position in its segment's
// chosen-axis list
__global__ void k_pk_side(const uint32_t* __restrict__ lists, uint32_t n, const uint32_t* __restrict__ seg_of,
                          PkLevel cur, uint8_t* __restrict__ side) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t j = seg_of[i];
        const uint32_t a = cur.axis[j];
        side[lists[(size_t)a * n + i]] = (i - cur.start[j] < cur.count[j] / 2) ? 0 : 1;
    }
}

// exclusive scan of the left flags over the three lists (3n entries):
// per-tile sums, a scan of the tile sums, per-tile scans
constexpr int kPkScanBlock = 256, kPkScanItems = 16, kPkTile = kPkScanBlock * kPkScanItems;

LRT_D uint32_t pk_flag(const uint32_t* lists, const uint8_t* side, size_t g, size_t total) {
    return g < total && side[lists[g]] == 0 ? 1u : 0u;
}

__global__ __launch_bounds__(kPkScanBlock) void k_pk_scan_tiles(const uint32_t* __restrict__ lists,
                                                              const uint8_t* __restrict__ side, size_t total,
                                                              uint32_t* __restrict__ tile_sum) {
    __shared__ uint32_t red[kPkScanBlock / 64];
    const size_t base = (size_t)blockIdx.x * kPkTile + (size_t)threadIdx.x * kPkScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int t = 0; t < kPkScanItems; ++t) s += pk_flag(lists, side, base + t, total);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t v = 0;
        for (int w = 0; w < kPkScanBlock / 64; ++w) v += red[w];
        tile_sum[blockIdx.x] = v;
    }
}

// exclusive scan of m tile sums in place, one block of 1024 threads
__global__ __launch_bounds__(1024) void k_pk_scan_sums(uint32_t* __restrict__ a, uint32_t m) {
    __shared__ uint32_t part[1024];
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(m, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kPkScanBlock) void k_pk_scan_apply(const uint32_t* __restrict__ lists,
                                                              const uint8_t* __restrict__ side, size_t total,
                                                              const uint32_t* __restrict__ tile_sum,
                                                              uint32_t* __restrict__ prefix) {
    __shared__ uint32_t part[kPkScanBlock];
    const size_t base = (size_t)blockIdx.x * kPkTile + (size_t)threadIdx.x * kPkScanItems;
    uint32_t f[kPkScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int t = 0; t < kPkScanItems; ++t) {
        f[t] = pk_flag(lists, side, base + t, total);
        s += f[t];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = 1; o < kPkScanBlock; o <<= 1) {
        const uint32_t v = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = tile_sum[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0u);
#pragma unroll
    for (int t = 0; t < kPkScanItems; ++t) {
        if (base + t < total) prefix[base + t] = run;
        run += f[t];
    }
}

// stable partition of every segment of the three lists: left photons first,
// each side in its list order; the next level's segment of every position
__global__ void k_pk_scatter(const uint32_t* __restrict__ lists, uint32_t n, const uint32_t* __restrict__ seg_of,
                             PkLevel cur, const uint8_t* __restrict__ side, const uint32_t* __restrict__ prefix,
                             uint32_t* __restrict__ out, uint32_t* __restrict__ seg_next) {
    const size_t total = 3 * (size_t)n;
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const uint32_t a = (uint32_t)(g / n), i = (uint32_t)(g - (size_t)a * n);
        const uint32_t j = seg_of[i], start = cur.start[j], k = cur.count[j] / 2;
        const uint32_t id = lists[g];
        const bool left = side[id] == 0;
        const uint32_t rank = prefix[g] - prefix[(size_t)a * n + start];  // left photons before position i
        const uint32_t to = left ? start + rank : start + k + (i - start - rank);
        out[(size_t)a * n + to] = id;
        if (a == 0) seg_next[to] = 2 * j + (left ? 0u : 1u);
    }
}

// the last level (every segment holds at most 9 photons): leaves, and for a
// segment of 9 one more split into leaves of 4 and 5; tree positions are the
// photons' x order within each leaf (order[position] = photon index)
__global__ void k_pk_finish(const float* __restrict__ xyz, const uint32_t* __restrict__ lists, uint32_t n, PkLevel cur,
                            uint32_t m, uint4* __restrict__ nodes, uint32_t* __restrict__ order) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        const uint32_t start = cur.start[j], c = cur.count[j], p = cur.pre[j];
        if (c <= 8) {
            nodes[p] = make_uint4(0u, 3u, start, c);
            for (uint32_t t = 0; t < c; ++t) order[start + t] = lists[start + t];
            continue;
        }
        // c == 9: left leaf of 4, right leaf of 5
        const int axis = pk_axis(xyz, lists, n, start, c);
        const uint32_t* L = lists + (size_t)axis * n + start;
        const float split = xyz[3 * (size_t)L[4] + axis];
        nodes[p] = make_uint4(__float_as_uint(split), (uint32_t)axis, p + 1, p + 2);
        nodes[p + 1] = make_uint4(0u, 3u, start, 4u);
        nodes[p + 2] = make_uint4(0u, 3u, start + 4, 5u);
        uint32_t nl = 0, nr = 0;
        for (uint32_t t = 0; t < 9; ++t) {
            const uint32_t id = lists[start + t];
            const bool left = id == L[0] || id == L[1] || id == L[2] || id == L[3];
            if (left) order[start + nl++] = id;
            else order[start + 4 + nr++] = id;
        }
    }
}

}  // namespace lrt
