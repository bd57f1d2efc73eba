// Wave / block helpers shared by the nlhip kernels (gfx950, wave64).
#pragma once
#include "nlhip_internal.h"

namespace nlh {

// ---------------------------------------------------------------------------
// wave / block helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Inclusive scans across lanes by DPP (data-parallel primitives: the VALU
// reads a neighbouring lane's register in the same instruction) instead of
// ds_bpermute round trips through the LDS crossbar.  GFX9 DPP controls:
// row_shr:n = 0x110 + n shifts within each 16-lane row, lanes shifted in
// from outside the row read 0; row_bcast:15 (0x142) hands lane 15 of a row
// to the next row (row mask 0xa: rows 1 and 3), row_bcast:31 (0x143) hands
// lane 31 to rows 2 and 3 (row mask 0xc).  Rows left out by the row mask
// read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_or_zero(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

// Inclusive scan within each 16-lane row (lanes 0..15 of the wave: the
// first row).
__device__ __forceinline__ int row16_incl_scan(int v)
{
    v += dpp_or_zero<0x111>(v);
    v += dpp_or_zero<0x112>(v);
    v += dpp_or_zero<0x114>(v);
    v += dpp_or_zero<0x118>(v);
    return v;
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    v = row16_incl_scan(v);
    v += dpp_or_zero<0x142, 0xa>(v);
    v += dpp_or_zero<0x143, 0xc>(v);
    return v;
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int64_t dpp64_or_zero(int64_t v)
{
    const uint32_t lo = (uint32_t)dpp_or_zero<CTRL, ROW_MASK>((int)(uint32_t)(uint64_t)v);
    const uint32_t hi = (uint32_t)dpp_or_zero<CTRL, ROW_MASK>((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t wave_incl_scan64(int64_t v)
{
    v += dpp64_or_zero<0x111>(v);
    v += dpp64_or_zero<0x112>(v);
    v += dpp64_or_zero<0x114>(v);
    v += dpp64_or_zero<0x118>(v);
    v += dpp64_or_zero<0x142, 0xa>(v);
    v += dpp64_or_zero<0x143, 0xc>(v);
    return v;
}

__device__ __forceinline__ int64_t readlane64(int64_t v, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Exclusive scan over an NT-thread block; returns this thread's exclusive
// prefix and the block total.
template <int NT>
__device__ __forceinline__ int64_t block_excl_scan_n(int64_t v, int64_t &total)
{
    __shared__ int64_t wsum[NT / kWave];
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int64_t inc = wave_incl_scan64(v);
    if (lane == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    int64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) {
        const int64_t s = wsum[w];
        if (w < wv) before += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t &total)
{
    return block_excl_scan_n<kScanThreads>(v, total);
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Blocks are dealt to the 8 XCDs round-robin; remap so every XCD works on a
// contiguous range of cells and keeps its neighbour-cell data in its own L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb / 8, r = nb % 8;
    const uint32_t xcd = b % 8, idx = b / 8;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

template <typename TI>
__device__ __forceinline__ void store_ti(void *p, int64_t idx, int64_t v)
{
    reinterpret_cast<TI *>(p)[idx] = (TI)v;
}

}  // namespace nlh
