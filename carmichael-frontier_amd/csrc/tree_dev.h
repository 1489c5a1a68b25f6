// tree_dev.h -- device helpers shared by the enumeration-tree kernels (generation and probe).
#pragma once
#include "cf_internal.h"

__device__ __forceinline__ int node_nxt(const Node &nd) { return (int)(short)(nd.pos & 0xFFFFu); }

// Record id after choosing index x at position `depth` (0-based) of an r-subset.
__device__ __forceinline__ u64 id_update(const SideDesc &s, u64 id, int x, int depth) {
    switch (s.id_mode) {
        case ID_COLEX: return id + s.binom[(size_t)x * (s.r + 1) + depth + 1];
        case ID_PACK_ASC: return id | ((u64)x << (s.id_bits * depth));
        case ID_PACK_DESC: return id | ((u64)x << (s.id_bits * (s.r - 1 - depth)));
        default: return id | (1ULL << x);
    }
}

// Parent of output j: the p in [plo, phi] with off[p] <= j < off[p+1] (off = exclusive scan of
// the per-parent output counts, nn + 1 entries).
__device__ __forceinline__ size_t find_parent(const u64 *__restrict__ off, size_t plo, size_t phi, u64 j) {
    while (plo < phi) {
        const size_t mid = (plo + phi + 1) >> 1;
        if (off[mid] <= j) plo = mid; else phi = mid - 1;
    }
    return plo;
}

// Index chosen by child number k of a node.
__device__ __forceinline__ int child_index(const SideDesc &s, const Node &nd, int k) {
    return (s.side == SIDE_I) ? node_nxt(nd) + k : node_nxt(nd) - k;
}
