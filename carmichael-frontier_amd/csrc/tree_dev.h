// tree_dev.h -- device helpers shared by the enumeration-tree kernels (generation and probe).
#pragma once
#include "cf_internal.h"

template <class N>
__device__ __forceinline__ int node_nxt(const N &nd) { return (int)(short)(nd.pos & 0xFFFFu); }

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
template <class N>
__device__ __forceinline__ int child_index(const SideDesc &s, const N &nd, int k) {
    return (s.side == SIDE_I) ? node_nxt(nd) + k : node_nxt(nd) - k;
}

// Wave-wide scans over the 64 lanes of a wave in DPP moves (no LDS round trip): shifts by 1, 2, 4
// and 8 lanes within each row of 16, then row 0's last lane broadcast into row 1 and row 2's into
// row 3 (row_bcast:15, rows 1 and 3), then lane 31 into rows 2 and 3 (row_bcast:31). A lane whose
// source is outside its row, or whose row is masked off, takes 0, the identity of both the sum and
// the unsigned maximum.
template <int CTRL, int ROWS>
__device__ __forceinline__ u32 dpp_or_zero(u32 v) {
    return (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ u32 wave_incl_sum(u32 v) {
    v += dpp_or_zero<0x111, 0xF>(v);   // row_shr:1
    v += dpp_or_zero<0x112, 0xF>(v);   // row_shr:2
    v += dpp_or_zero<0x114, 0xF>(v);   // row_shr:4
    v += dpp_or_zero<0x118, 0xF>(v);   // row_shr:8
    v += dpp_or_zero<0x142, 0xA>(v);   // row_bcast:15
    v += dpp_or_zero<0x143, 0xC>(v);   // row_bcast:31
    return v;
}
__device__ __forceinline__ u32 umax32(u32 a, u32 b) { return a > b ? a : b; }
__device__ __forceinline__ u32 wave_incl_max(u32 v) {
    v = umax32(v, dpp_or_zero<0x111, 0xF>(v));
    v = umax32(v, dpp_or_zero<0x112, 0xF>(v));
    v = umax32(v, dpp_or_zero<0x114, 0xF>(v));
    v = umax32(v, dpp_or_zero<0x118, 0xF>(v));
    v = umax32(v, dpp_or_zero<0x142, 0xA>(v));
    v = umax32(v, dpp_or_zero<0x143, 0xC>(v));
    return v;
}
// The 64-bit sum: each step moves both halves with the same pattern and adds them as one value.
template <int CTRL, int ROWS>
__device__ __forceinline__ u64 dpp64_or_zero(u64 v) {
    return ((u64)dpp_or_zero<CTRL, ROWS>((u32)(v >> 32)) << 32) | dpp_or_zero<CTRL, ROWS>((u32)v);
}
__device__ __forceinline__ u64 wave_incl_sum64(u64 v) {
    v += dpp64_or_zero<0x111, 0xF>(v);
    v += dpp64_or_zero<0x112, 0xF>(v);
    v += dpp64_or_zero<0x114, 0xF>(v);
    v += dpp64_or_zero<0x118, 0xF>(v);
    v += dpp64_or_zero<0x142, 0xA>(v);
    v += dpp64_or_zero<0x143, 0xC>(v);
    return v;
}
// The value of the lane below (0 in lane 0): wave_shr:1.
__device__ __forceinline__ u32 wave_shift_up1(u32 v) { return dpp_or_zero<0x138, 0xF>(v); }
