// nlhip device kernels for the sort-based neighbour-list build (gfx950, wave64).
//
// Pipeline (a single build is a batch of one system).  The cell sort
// (compute_cell_ids_kernel!, gpu_kernels.jl:108-125; the histogram and
// offsets, :191-197 and :262-285; sortperm + gathers, cell_list.jl:668-670)
// has two forms with the same result:
//   bucketed (default)
//     bucket_count_kernel   cell id + winding per atom, counts per bucket of
//                           cells in LDS, per block of atoms
//     scan (counts)         every block's place in every bucket
//     bucket_scatter_kernel atoms grouped by bucket
//     bucket_place_kernel   one workgroup per bucket: cell_start / cell_offsets,
//                           stable in-cell rank -> perm, cell_id, sorted records
//   atomic
//     bin_kernel            cell id + winding, cell histogram by global atomics
//     scan (cells)          cell_start / cell_offsets
//     scatter_kernel        group atom indices by cell (arrival order)
//     place_kernel          stable in-cell rank -> perm, cell_id, sorted records
// then
//   scan (pairs)      first / pair offsets (compute_pair_offsets,
//                     gpu_kernels.jl:222-237)
// The pair search itself (count and fill) is in pairs.hip.
#include "device_util.h"

namespace nlh {

// ---------------------------------------------------------------------------
// 0. batches: the system of every atom and every cell, once per build, for
//    the kernels that would otherwise each search the system tables
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void sys_map_kernel(const int64_t *__restrict__ sys_atom_base,
                                                      const int32_t *__restrict__ sys_cell_base, int nsys,
                                                      int64_t nat, int64_t M, int32_t *__restrict__ atom_sys)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nat) atom_sys[g] = find_sys(sys_atom_base, nsys, g);
    else if (g < nat + M) atom_sys[g] = find_sys(sys_cell_base, nsys, (int32_t)(g - nat));
}

// ---------------------------------------------------------------------------
// 1. binning + cell histogram
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void bin_kernel(const double *__restrict__ X, const SysDev *__restrict__ sys,
                                                  const int32_t *__restrict__ atom_sys, int nsys, int64_t nat,
                                                  int32_t *__restrict__ key, int32_t *__restrict__ slot,
                                                  int32_t *__restrict__ cell_count, Status *__restrict__ status)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nat) return;
    const int s = nsys == 1 ? 0 : atom_sys[g];
    const SysDev &sd = sys[s];
    const double x = X[3 * g], y = X[3 * g + 1], z = X[3 * g + 2];
    int32_t c[3], w[3];
    if (!cell_and_winding(sd, x, y, z, c, w)) atomicOr(&status->bad_position, 1);
    const int32_t k = sd.cell_base + (c[0] - 1) + (c[1] - 1) * sd.n[0] + (c[2] - 1) * sd.n[0] * sd.n[1];
    key[g] = k;
    slot[g] = atomicAdd(&cell_count[k], 1);
}

// ---------------------------------------------------------------------------
// 2. scans
// ---------------------------------------------------------------------------
//
// Up to kLookbackTiles tiles of kScanTile items a scan is one kernel with a
// decoupled look-back (one look-back step reaches tile 0); beyond that it is
// reduce -> scan of the tile sums -> downsweep, three kernels, where a long
// chain of look-back steps measured slower (1M pair counts: 51 us single pass
// against 42 us).  The downsweep kernels serve both.
//
// Look-back: desc[0] is a ticket, desc[1 + t] the descriptor of tile t: a state in the
// top two bits (0 not yet published, 1 the tile's aggregate, 2 its inclusive
// prefix) and the value below.  A block takes its tile from the ticket, so
// every tile it waits for belongs to a block that started before it and is
// running or done.  The descriptors are zero before a launch (cleared with
// the status at the start of a build, or by clear_lookback before a repeated
// pair scan).

constexpr uint64_t kDescAgg = 1ull << 62, kDescInc = 2ull << 62, kDescVal = kDescAgg - 1;
constexpr int kLookbackSpins = 1 << 20;   // about a second: reached only if an invariant is broken

__device__ __forceinline__ void load_tile(const int32_t *__restrict__ in, int64_t n, int64_t base, int32_t v[kScanItems],
                                          int32_t mask = -1)
{
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    if (b + kScanItems <= n) {
        const int4 *p = reinterpret_cast<const int4 *>(in + b);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const int4 t = p[q];
            v[4 * q] = t.x & mask; v[4 * q + 1] = t.y & mask; v[4 * q + 2] = t.z & mask; v[4 * q + 3] = t.w & mask;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) v[q] = (b + q < n) ? in[b + q] & mask : 0;
    }
}

// This block's tile, in launch order of the blocks; -1 (and scan_error)
// when the ticket was not zero at the launch, so no descriptor past the grid
// is ever touched.
__device__ __forceinline__ int64_t scan_ticket(uint64_t *__restrict__ desc, Status *__restrict__ status)
{
    __shared__ int64_t tile;
    if (threadIdx.x == 0) {
        tile = (int64_t)atomicAdd(reinterpret_cast<unsigned long long *>(desc), 1ull);
        if (tile >= (int64_t)gridDim.x) {
            status->scan_error = 1;
            tile = -1;
        }
    }
    __syncthreads();
    return tile;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const int32_t *__restrict__ in, int64_t n,
                                                                   int64_t *__restrict__ partials, int32_t mask)
{
    int32_t v[kScanItems];
    load_tile(in, n, (int64_t)blockIdx.x * kScanTile, v, mask);
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    block_excl_scan(s, tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Exclusive scan of nb tile sums in place by one block.  For the pair scan it
// also records the total and the capacity verdict.
__global__ __launch_bounds__(kScanThreads) void scan_partials_kernel(int64_t *__restrict__ partials, int nb,
                                                                     Status *__restrict__ status, int64_t capacity)
{
    int64_t carry = 0;
    for (int base = 0; base < nb; base += kScanThreads) {
        const int i = base + threadIdx.x;
        const int64_t v = i < nb ? partials[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan(v, tot);
        if (i < nb) partials[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && status) {
        status->npairs = carry;
        status->overflow = carry > capacity ? 1 : 0;
    }
}

// The tile of this block and the sum of the tiles before it: from the tile
// sums when given (three-kernel scan), else by ticket and look-back.
// Returns -1 for a block with no tile (see scan_ticket).
__device__ __forceinline__ int64_t scan_tile_begin(uint64_t *__restrict__ desc, const int64_t *__restrict__ partials,
                                                   Status *__restrict__ status)
{
    return partials ? (int64_t)blockIdx.x : scan_ticket(desc, status);
}

__device__ __forceinline__ uint64_t desc_load(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void desc_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Exclusive prefix of tile `tile` whose own sum is `agg`.  The block's first
// wave publishes the aggregate, then reads the descriptors of the 64
// preceding tiles at once (lane l: tile - 1 - l), waits until those up to the
// nearest inclusive prefix are published, and adds their values, moving 64
// tiles further back if none is an inclusive prefix; then it publishes its
// own inclusive prefix.
// Waits are bounded: a wave that waits too long sets status->scan_error (the
// host reports it) rather than spin for ever.
__device__ __forceinline__ int64_t scan_lookback(uint64_t *__restrict__ desc, int64_t tile, int64_t agg,
                                                 Status *__restrict__ status)
{
    __shared__ int64_t prefix;
    uint64_t *d = desc + 1;
    if (threadIdx.x < kWave) {
        const int lane = (int)threadIdx.x;
        if (tile > 0 && lane == 0) desc_store(d + tile, kDescAgg | (uint64_t)agg);
        int64_t ex = 0;
        int spins = 0;
        for (int64_t top = tile - 1; top >= 0; top -= kWave) {
            const int64_t idx = top - lane;
            // before tile 0: an inclusive prefix of zero
            uint64_t v = idx >= 0 ? desc_load(d + idx) : kDescInc;
            for (;;) {
                const bool pending = (v >> 62) == 0;
                // only the lanes before the nearest inclusive prefix matter
                const uint64_t inc = __builtin_amdgcn_ballot_w64((v >> 62) == 2);
                const uint64_t need = inc ? (inc & (~inc + 1)) - 1 : ~0ull;
                if ((__builtin_amdgcn_ballot_w64(pending) & need) == 0) break;
                if (++spins > kLookbackSpins) {
                    if (lane == 0) status->scan_error = 1;
                    if (pending) v = kDescInc;   // give up on it: an inclusive zero
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                if (pending) v = desc_load(d + idx);
            }
            const uint64_t inc = __builtin_amdgcn_ballot_w64((v >> 62) == 2);
            const int64_t val = (int64_t)(v & kDescVal);
            if (inc != 0) {
                const int k = __builtin_ctzll(inc);   // the nearest inclusive prefix
                ex += wave_sum64(lane <= k ? val : 0);
                break;
            }
            ex += wave_sum64(val);
        }
        if (lane == 0) {
            desc_store(d + tile, kDescInc | (uint64_t)(ex + agg));
            prefix = ex;
        }
    }
    __syncthreads();
    return prefix;
}

// A tile's exclusive prefixes and counts pass through LDS: written blocked
// (thread t: items 8t .. 8t + 7, the scan's order), read striped (thread t:
// items t + 256 q), so that consecutive lanes store consecutive elements.
// One padding slot per 8 items keeps the blocked accesses off shared banks.
__device__ __forceinline__ int tile_slot(int i) { return i + (i >> 3); }
constexpr int kTileSlots = kScanTile + kScanTile / 8;

// Cell histogram -> cell_start (0-based), the caller-facing 1-based
// cell_offsets per system; the histogram is cleared for the next build.
template <typename TI>
__global__ __launch_bounds__(kScanThreads) void scan_cells_kernel(
    int32_t *__restrict__ cell_count, int64_t M, uint64_t *__restrict__ desc, const int64_t *__restrict__ partials,
    int32_t *__restrict__ cell_start, const SysDev *__restrict__ sys, const int32_t *__restrict__ cell_sys,
    int nsys, void *__restrict__ offsets_out, Status *__restrict__ status)
{
    const int64_t tile = scan_tile_begin(desc, partials, status);
    if (tile < 0) return;
    int32_t v[kScanItems];
    const int64_t base = tile * kScanTile;
    load_tile(cell_count, M, base, v);
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    const int64_t ex = block_excl_scan(s, tot);
    __shared__ int64_t ex_sh[kTileSlots];
    __shared__ int32_t cnt_sh[kTileSlots];
    int64_t run = ex + (partials ? partials[tile] : scan_lookback(desc, tile, tot, status));
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const int i = tile_slot((int)threadIdx.x * kScanItems + q);
        ex_sh[i] = run;
        cnt_sh[i] = v[q];
        run += v[q];
    }
    __syncthreads();
    for (int q = 0; q < kScanItems; ++q) {
        const int i = q * kScanThreads + (int)threadIdx.x;
        const int64_t c = base + i;
        if (c < M) {
            const int64_t r0 = ex_sh[tile_slot(i)];
            const int32_t n0 = cnt_sh[tile_slot(i)];
            cell_start[c] = (int32_t)r0;
            cell_count[c] = 0;
            const int si = nsys == 1 ? 0 : cell_sys[c];
            const SysDev &sd = sys[si];
            store_ti<TI>(offsets_out, c + si, r0 - sd.atom_base + 1);
            if (c == (int64_t)sd.cell_base + sd.M - 1)   // trailing entry of this system
                store_ti<TI>(offsets_out, c + si + 1, r0 + n0 - sd.atom_base + 1);
            if (c == M - 1) cell_start[M] = (int32_t)(r0 + n0);
        }
    }
}

// Pair counts -> pair_off (global, 0-based, int64), for a single system
// `first` (1-based TI), the fill's tile table, the total, whether it fits the
// output capacity and the largest count.
template <typename TI>
__global__ __launch_bounds__(kScanThreads) void scan_pairs_kernel(
    const int32_t *__restrict__ counts, int64_t nat, uint64_t *__restrict__ desc, const int64_t *__restrict__ partials,
    int64_t *__restrict__ pair_off, int nsys, void *__restrict__ first_out, Status *__restrict__ status,
    int64_t capacity, int32_t *__restrict__ tile_atom, int64_t tile_cap, int tile_shift)
{
    const int64_t tile = scan_tile_begin(desc, partials, status);
    if (tile < 0) return;
    int32_t v[kScanItems];
    const int64_t base = tile * kScanTile;
    load_tile(counts, nat, base, v, kCountMask);   // bit 30: slot-row format
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    const int64_t ex = block_excl_scan(s, tot);
    const int64_t boff = partials ? partials[tile] : scan_lookback(desc, tile, tot, status);
    if (!partials && tile == (int64_t)gridDim.x - 1 && threadIdx.x == 0) {   // the last tile sees the total
        status->npairs = boff + tot;
        status->overflow = boff + tot > capacity ? 1 : 0;
    }
    int32_t mx = 0;   // largest per-atom count (slot capacity check)
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) mx = v[q] > mx ? v[q] : mx;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int32_t o = __shfl_xor(mx, off, kWave);
        mx = o > mx ? o : mx;
    }
    // one atomic per block, and only while it can raise the maximum: atomics
    // on the one status word serialise (one per wave cost 20 us at 1M atoms)
    __shared__ int32_t wmax[kScanThreads / kWave];
    if (lane_id() == 0) wmax[threadIdx.x / kWave] = mx;
    __shared__ int64_t ex_sh[kTileSlots];
    __shared__ int32_t cnt_sh[kTileSlots];
    int64_t run = boff + ex;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const int i = tile_slot((int)threadIdx.x * kScanItems + q);
        ex_sh[i] = run;
        cnt_sh[i] = v[q];
        run += v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t bm = 0;
#pragma unroll
        for (int w = 0; w < kScanThreads / kWave; ++w) bm = wmax[w] > bm ? wmax[w] : bm;
        if (bm > __hip_atomic_load(&status->max_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&status->max_count, bm);
    }
    const int64_t tsz = (int64_t)1 << tile_shift;
    for (int q = 0; q < kScanItems; ++q) {
        const int i = q * kScanThreads + (int)threadIdx.x;
        const int64_t a = base + i;
        if (a < nat) {
            const int64_t r0 = ex_sh[tile_slot(i)];
            const int32_t n0 = cnt_sh[tile_slot(i)];
            pair_off[a] = r0;
            // fill tiles whose first pair is one of this atom's: entry t is
            // the atom owning pair t << tile_shift and the number of its
            // pairs before that one
            for (int64_t t = (r0 + tsz - 1) >> tile_shift; t < tile_cap && (t << tile_shift) < r0 + n0; ++t) {
                tile_atom[2 * t] = (int32_t)a;
                tile_atom[2 * t + 1] = (int32_t)((t << tile_shift) - r0);
            }
            if (nsys == 1) store_ti<TI>(first_out, a, r0 + 1);
            if (a == nat - 1) {
                pair_off[nat] = r0 + n0;
                if (nsys == 1) store_ti<TI>(first_out, nat, r0 + n0 + 1);
            }
        }
    }
}

// Per-system `first` arrays for a batch (local, 1-based) and the per-system
// pair bases.  System s owns entries [atom_base[s] + s, atom_base[s+1] + s] of
// the concatenated first arrays: thread g writes the entry of atom g (its
// system from atom_sys) and, for g in [1, nsys], the trailing entry of
// system g - 1 (its pair count + 1).
template <typename TI>
__global__ __launch_bounds__(256) void batch_first_kernel(const int64_t *__restrict__ pair_off, int64_t nat,
                                                          const int64_t *__restrict__ sys_atom_base,
                                                          const int32_t *__restrict__ atom_sys, int nsys,
                                                          void *__restrict__ first_out, int64_t *__restrict__ sys_pair_base)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g <= nsys) {
        const int64_t e = pair_off[sys_atom_base[g]];
        sys_pair_base[g] = e;
        if (g > 0) store_ti<TI>(first_out, sys_atom_base[g] + g - 1, e - pair_off[sys_atom_base[g - 1]] + 1);
    }
    if (g < nat) {
        const int s = atom_sys[g];
        store_ti<TI>(first_out, g + s, pair_off[g] - pair_off[sys_atom_base[s]] + 1);
    }
}

// ---------------------------------------------------------------------------
// 3. grouping by cell, stable placement
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void scatter_kernel(const int32_t *__restrict__ key, const int32_t *__restrict__ slot,
                                                      const int32_t *__restrict__ cell_start, int64_t nat,
                                                      int32_t *__restrict__ tmp)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nat) return;
    tmp[cell_start[key[g]] + slot[g]] = (int32_t)g;
}

// Stable placement: an atom's position inside its cell is the number of
// atoms of the same cell with a smaller index, so ties keep ascending
// original index -- the order of Julia's stable sortperm.
template <typename TI>
__global__ __launch_bounds__(256) void place_kernel(const double *__restrict__ X, const SysDev *__restrict__ sys,
                                                    const int32_t *__restrict__ cell_sys, int nsys,
                                                    const int32_t *__restrict__ key, const int32_t *__restrict__ cell_start,
                                                    const int32_t *__restrict__ tmp, int64_t nat, Rec *__restrict__ rec,
                                                    int32_t *__restrict__ wbig, void *__restrict__ perm_out,
                                                    void *__restrict__ cell_id_out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nat) return;
    const int32_t e = tmp[q];
    const int32_t c = key[e];
    const int32_t cs = cell_start[c], ce = cell_start[c + 1];
    int32_t rank = 0;
    for (int32_t t = cs; t < ce; ++t) rank += tmp[t] < e;
    const int32_t p = cs + rank;
    const int s = nsys == 1 ? 0 : cell_sys[c];
    const SysDev &sd = sys[s];
    const double x = X[3 * (int64_t)e], y = X[3 * (int64_t)e + 1], z = X[3 * (int64_t)e + 2];
    int32_t cc[3], w[3];
    cell_and_winding(sd, x, y, z, cc, w);
    const int32_t j = (int32_t)(e - sd.atom_base);
    Rec r;
    r.x = x; r.y = y; r.z = z; r.j = j;
    r.w = pack_wind(w[0], w[1], w[2]);
    rec[p] = r;
    if (r.w & kWindBig) {
        wbig[3 * (int64_t)p] = w[0];
        wbig[3 * (int64_t)p + 1] = w[1];
        wbig[3 * (int64_t)p + 2] = w[2];
    }
    store_ti<TI>(perm_out, p, (int64_t)j + 1);
    store_ti<TI>(cell_id_out, p, (int64_t)(c - sd.cell_base) + 1);
}

// Sorted positions X[perm] (SortedCellList.X), produced on request.
__global__ __launch_bounds__(256) void sorted_positions_kernel(const Rec *__restrict__ rec, int64_t nat, double *__restrict__ Xs)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nat) return;
    const Rec r = rec[p];
    Xs[3 * p] = r.x;
    Xs[3 * p + 1] = r.y;
    Xs[3 * p + 2] = r.z;
}

// ---------------------------------------------------------------------------
// 3b. bucketed cell sort (the default form; see the top of the file)
// ---------------------------------------------------------------------------
//
// The global-atomic histogram costs one L2 atomic per atom, and those
// atomics run at a fixed rate (1M atoms: 38 of bin_kernel's 44 us).  Here
// every counter lives in LDS: a block of P.apb consecutive atoms counts them
// per bucket (2^P.bw consecutive cells), one scan over the bucket-major
// P.nb x P.na counts places every (bucket, block) run, and one workgroup per
// bucket orders its few hundred to few thousand atoms by cell and original
// index with LDS counters again.  Only (cell, index) pairs move through the
// scatter; the ordering pass reads the positions from X itself.

__global__ __launch_bounds__(kBinThreads) void bucket_count_kernel(
    const double *__restrict__ X, const SysDev *__restrict__ sys, const int32_t *__restrict__ atom_sys, int nsys,
    int64_t nat, SortPlan P, int32_t *__restrict__ key, int32_t *__restrict__ slot, int32_t *__restrict__ bmat,
    Status *__restrict__ status)
{
    extern __shared__ int32_t hist[];   // P.nb counters
    for (int b = threadIdx.x; b < P.nb; b += kBinThreads) hist[b] = 0;
    __syncthreads();
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);   // consecutive blocks on one XCD
    const int64_t a0 = (int64_t)lb * P.apb;
    const int64_t a1 = a0 + P.apb < nat ? a0 + P.apb : nat;
    bool bad = false;
    for (int64_t g0 = a0 + threadIdx.x; g0 < a1; g0 += kBinThreads * kBinUnroll) {
        double x[kBinUnroll], y[kBinUnroll], z[kBinUnroll];
        int sy[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; ++u) {   // all loads of the thread in flight together
            const int64_t g = g0 + (int64_t)u * kBinThreads;
            x[u] = y[u] = z[u] = 0.0;
            sy[u] = 0;
            if (g < a1) {
                x[u] = X[3 * g];
                y[u] = X[3 * g + 1];
                z[u] = X[3 * g + 2];
                if (nsys != 1) sy[u] = atom_sys[g];
            }
        }
#pragma unroll
        for (int u = 0; u < kBinUnroll; ++u) {
            const int64_t g = g0 + (int64_t)u * kBinThreads;
            if (g < a1) {
                const SysDev &sd = sys[sy[u]];
                int32_t c[3], w[3];
                bad |= !cell_and_winding(sd, x[u], y[u], z[u], c, w);
                const int32_t k = sd.cell_base + (c[0] - 1) + (c[1] - 1) * sd.n[0] + (c[2] - 1) * sd.n[0] * sd.n[1];
                // read once more, by the scatter: streamed past the caches like the
                // sort's other whole-line scratch traffic (the pair pass later relies
                // on the last-level cache holding its slot rows)
                __builtin_nontemporal_store(k, key + g);
                __builtin_nontemporal_store(atomicAdd(&hist[k >> P.bw], 1), slot + g);   // arrival rank in (block, bucket)
            }
        }
    }
    if (bad) atomicOr(&status->bad_position, 1);
    __syncthreads();
    for (int b = threadIdx.x; b < P.nb; b += kBinThreads) bmat[(int64_t)b * P.na + lb] = hist[b];
}

// Exclusive scan of n counts in place (the bucket counts); tile prefixes
// by look-back, or from the tile sums of the three-kernel form.
__global__ __launch_bounds__(kScanThreads) void scan_inplace_kernel(int32_t *__restrict__ a, int64_t n,
                                                                    uint64_t *__restrict__ desc,
                                                                    const int64_t *__restrict__ partials,
                                                                    Status *__restrict__ status)
{
    const int64_t tile = scan_tile_begin(desc, partials, status);
    if (tile < 0) return;
    int32_t v[kScanItems];
    const int64_t base = tile * kScanTile;
    load_tile(a, n, base, v);
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    const int64_t ex = block_excl_scan(s, tot);
    int64_t run = ex + (partials ? partials[tile] : scan_lookback(desc, tile, tot, status));
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    int32_t o[kScanItems];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        o[q] = (int32_t)run;
        run += v[q];
    }
    if (b + kScanItems <= n) {
        int4 *p = reinterpret_cast<int4 *>(a + b);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) p[q] = make_int4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < kScanItems; ++q)
            if (b + q < n) a[b + q] = o[q];
    }
}

__global__ __launch_bounds__(kBinThreads) void bucket_scatter_kernel(int64_t nat, SortPlan P,
                                                                     const int32_t *__restrict__ key,
                                                                     const int32_t *__restrict__ slot,
                                                                     const int32_t *__restrict__ boff,
                                                                     int2 *__restrict__ bke)
{
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);   // the count kernel's block order
    const int64_t a0 = (int64_t)lb * P.apb;
    const int64_t a1 = a0 + P.apb < nat ? a0 + P.apb : nat;
    for (int64_t g0 = a0 + threadIdx.x; g0 < a1; g0 += kBinThreads * kBinUnroll) {
        int32_t k[kBinUnroll], pos[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; ++u) {
            const int64_t g = g0 + (int64_t)u * kBinThreads;
            k[u] = 0;
            pos[u] = 0;
            if (g < a1) {
                k[u] = __builtin_nontemporal_load(key + g);
                pos[u] = __builtin_nontemporal_load(slot + g);
            }
        }
#pragma unroll
        for (int u = 0; u < kBinUnroll; ++u) {
            const int64_t g = g0 + (int64_t)u * kBinThreads;
            // scattered stores: cached, so that the L2 merges them into lines
            if (g < a1) bke[boff[(int64_t)(k[u] >> P.bw) * P.na + lb] + pos[u]] = make_int2(k[u], (int32_t)g);
        }
    }
}

__device__ __forceinline__ int2 load_ke(const int2 *p)
{
    const uint64_t v = __builtin_nontemporal_load(reinterpret_cast<const uint64_t *>(p));
    return make_int2((int32_t)(uint32_t)v, (int32_t)(uint32_t)(v >> 32));
}

// One workgroup per bucket: its cells' counts and starts in LDS, the atoms
// grouped by cell in arrival order, then each atom's rank among its cell's
// atoms by original index -- ties of the cell key keep ascending original
// index, the order of Julia's stable sortperm.  A bucket of at most
// kBucketLdsAtoms atoms keeps its cells, indices and groups in LDS; a larger
// one (clustered atoms) works from global memory (bke and tmp, its own range).
// The sorted record, perm and cell_id of atom e (global index) of cell k at
// sorted place p.
template <typename TI>
__device__ __forceinline__ void place_one(int32_t k, int32_t e, int32_t p, const double *__restrict__ X,
                                          const SysDev *__restrict__ sys, const int32_t *__restrict__ cell_sys, int nsys,
                                          Rec *__restrict__ rec, int32_t *__restrict__ wbig, void *__restrict__ perm_out,
                                          void *__restrict__ cell_id_out)
{
    const int si = nsys == 1 ? 0 : cell_sys[k];
    const SysDev &sd = sys[si];
    const double x = X[3 * (int64_t)e], y = X[3 * (int64_t)e + 1], z = X[3 * (int64_t)e + 2];
    int32_t cc[3], w[3];
    cell_and_winding(sd, x, y, z, cc, w);
    const int32_t j = (int32_t)(e - sd.atom_base);
    Rec o;
    o.x = x;
    o.y = y;
    o.z = z;
    o.j = j;
    o.w = pack_wind(w[0], w[1], w[2]);
    rec[p] = o;
    if (o.w & kWindBig) {
        wbig[3 * (int64_t)p] = w[0];
        wbig[3 * (int64_t)p + 1] = w[1];
        wbig[3 * (int64_t)p + 2] = w[2];
    }
    store_ti<TI>(perm_out, p, (int64_t)j + 1);
    store_ti<TI>(cell_id_out, p, (int64_t)(k - sd.cell_base) + 1);
}

template <typename TI>
__global__ __launch_bounds__(kBucketThreads) void bucket_place_kernel(
    const int2 *__restrict__ bke, const double *__restrict__ X, const int32_t *__restrict__ boff, SortPlan P,
    int64_t nat, int32_t M, const SysDev *__restrict__ sys, const int32_t *__restrict__ cell_sys, int nsys,
    int32_t *__restrict__ tmp, int32_t *__restrict__ cell_start, void *__restrict__ offsets_out,
    Rec *__restrict__ rec, int32_t *__restrict__ wbig, void *__restrict__ perm_out, void *__restrict__ cell_id_out)
{
    __shared__ int32_t kl[kBucketLdsAtoms], el[kBucketLdsAtoms], grp[kBucketLdsAtoms];
    extern __shared__ int32_t lds[];
    const int b = (int)blockIdx.x;
    const int32_t c0 = b << P.bw;
    const int W = (1 << P.bw) < M - c0 ? (1 << P.bw) : M - c0;
    int32_t *co = lds;                        // W + 1: cell starts within the bucket
    int32_t *fill = lds + (1 << P.bw) + 1;    // W: arrivals per cell
    const int32_t bs = boff[(int64_t)b * P.na];
    const int32_t be = b + 1 < P.nb ? boff[(int64_t)(b + 1) * P.na] : (int32_t)nat;
    const int n = be - bs;
    const bool small = n <= kBucketLdsAtoms;   // uniform
    for (int t = threadIdx.x; t < W; t += kBucketThreads) {
        co[t] = 0;
        fill[t] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kBucketThreads) {
        const int2 ke = load_ke(bke + bs + i);
        if (small) {
            kl[i] = ke.x;
            el[i] = ke.y;
        }
        atomicAdd(&co[ke.x - c0], 1);
    }
    __syncthreads();
    // exclusive scan of co[0, W): thread t takes a run of `per` cells
    const int per = (W + kBucketThreads - 1) / kBucketThreads;
    const int t0 = (int)threadIdx.x * per < W ? (int)threadIdx.x * per : W;
    const int t1 = t0 + per < W ? t0 + per : W;
    int32_t sum = 0;
    for (int t = t0; t < t1; ++t) sum += co[t];
    int64_t tot;
    int32_t run = (int32_t)block_excl_scan_n<kBucketThreads>(sum, tot);   // its barriers end every read of co
    for (int t = t0; t < t1; ++t) {
        const int32_t v = co[t];
        co[t] = run;
        run += v;
    }
    if (threadIdx.x == 0) co[W] = (int32_t)tot;
    __syncthreads();
    for (int t = threadIdx.x; t < W; t += kBucketThreads) {
        const int32_t c = c0 + t;
        const int32_t r0 = bs + co[t];
        cell_start[c] = r0;
        const int si = nsys == 1 ? 0 : cell_sys[c];
        const SysDev &sd = sys[si];
        store_ti<TI>(offsets_out, (int64_t)c + si, r0 - sd.atom_base + 1);
        if (c == sd.cell_base + sd.M - 1)   // trailing entry of this system
            store_ti<TI>(offsets_out, (int64_t)c + si + 1, bs + co[t + 1] - sd.atom_base + 1);
    }
    if (b == P.nb - 1 && threadIdx.x == 0) cell_start[M] = (int32_t)nat;
    int32_t *grpg = tmp + bs;   // the large-bucket groups
    for (int i = threadIdx.x; i < n; i += kBucketThreads) {
        int32_t k, e;
        if (small) {
            k = kl[i];
            e = el[i];
        } else {
            const int2 ke = load_ke(bke + bs + i);
            k = ke.x;
            e = ke.y;
        }
        const int c = k - c0;
        const int q = co[c] + atomicAdd(&fill[c], 1);
        if (small) grp[q] = e;
        else grpg[q] = e;
    }
    __syncthreads();   // the groups (LDS, or the block's own tmp stores) are visible to the block
    if (small) {
        // each atom's place in the bucket (its cell's start plus the atoms
        // of its cell with a smaller index), then (cell, index) moved to that
        // place in LDS, so that the records go out in sorted order: whole
        // cache lines instead of scattered 32-byte stores
        constexpr int kPer = kBucketLdsAtoms / kBucketThreads;
        int32_t kq[kPer], eq[kPer], pq[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int i = u * kBucketThreads + (int)threadIdx.x;
            pq[u] = -1;
            if (i < n) {
                const int32_t k = kl[i], e = el[i];
                const int c = k - c0;
                const int s0 = co[c], s1 = co[c + 1];
                int32_t rank = 0;
                for (int t = s0; t < s1; ++t) rank += grp[t] < e;
                kq[u] = k;
                eq[u] = e;
                pq[u] = s0 + rank;
            }
        }
        __syncthreads();   // every read of kl and el is done
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            if (pq[u] >= 0) {
                kl[pq[u]] = kq[u];
                el[pq[u]] = eq[u];
            }
        __syncthreads();
        for (int q = threadIdx.x; q < n; q += kBucketThreads)
            place_one<TI>(kl[q], el[q], bs + q, X, sys, cell_sys, nsys, rec, wbig, perm_out, cell_id_out);
        return;
    }
    for (int i = threadIdx.x; i < n; i += kBucketThreads) {
        int32_t k, e;
        {
            const int2 ke = load_ke(bke + bs + i);
            k = ke.x;
            e = ke.y;
        }
        const int c = k - c0;
        const int s0 = co[c], s1 = co[c + 1];
        int32_t rank = 0;
        for (int t = s0; t < s1; ++t) rank += grpg[t] < e;
        const int32_t p = bs + s0 + rank;
        place_one<TI>(k, e, p, X, sys, cell_sys, nsys, rec, wbig, perm_out, cell_id_out);
    }
}

// Zero the build status (when given) and n scan descriptors.  A kernel
// rather than a memset, so the pipeline can be captured into a HIP graph and
// replayed.
__global__ void clear_status_kernel(Status *__restrict__ status, uint64_t *__restrict__ lookback, int64_t n)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (status && g < (int64_t)(sizeof(Status) / 4)) reinterpret_cast<int32_t *>(status)[g] = 0;
    for (int64_t i = g; i < n; i += (int64_t)gridDim.x * blockDim.x) lookback[i] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void launch_clear_status(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = lookback_entries(b.sort.n_sort, b.nat);
    hipLaunchKernelGGL(clear_status_kernel, dim3(std::min(blocks_for(n, 256), 64u)), dim3(256), 0, st, b.status,
                       b.lookback, n);
}

void launch_zero_words(uint64_t *p, int64_t n, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(clear_status_kernel, dim3(std::min(blocks_for(n, 256), 256u)), dim3(256), 0, st,
                       (Status *)nullptr, p, n);
}

void launch_clear_pair_lookback(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = scan_tiles(b.nat) + 1;
    hipLaunchKernelGGL(clear_status_kernel, dim3(std::min(blocks_for(n, 256), 64u)), dim3(256), 0, st,
                       (Status *)nullptr, b.lookback + lookback_cells(b.sort.n_sort), n);
}

#define NLH_TI_DISPATCH(ti, ...)            \
    do {                                    \
        if ((ti) == 8) {                    \
            using TI = int64_t;             \
            __VA_ARGS__;                    \
        } else {                            \
            using TI = int32_t;             \
            __VA_ARGS__;                    \
        }                                   \
    } while (0)

static void launch_bucket_sort(const PipelineBufs &b, hipStream_t st)
{
    const SortPlan &P = b.sort;
    hipLaunchKernelGGL(bucket_count_kernel, dim3((unsigned)P.na), dim3(kBinThreads), (size_t)P.nb * 4, st, b.X, b.sys,
                       b.atom_sys, b.nsys, b.nat, P, b.key, b.slot, b.bmat, b.status);
    const int64_t nt = scan_tiles(P.n_sort);
    int64_t *part = nullptr;   // three-kernel scan: tile sums in the (then unused) descriptor slots
    if (nt > kSortLookbackTiles) {
        part = reinterpret_cast<int64_t *>(b.lookback);
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nt), dim3(kScanThreads), 0, st, b.bmat, P.n_sort, part,
                           (int32_t)-1);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, part, (int)nt, (Status *)nullptr,
                           (int64_t)0);
    }
    hipLaunchKernelGGL(scan_inplace_kernel, dim3((unsigned)nt), dim3(kScanThreads), 0, st, b.bmat, P.n_sort,
                       b.lookback, part, b.status);
    hipLaunchKernelGGL(bucket_scatter_kernel, dim3((unsigned)P.na), dim3(kBinThreads), 0, st, b.nat, P, b.key, b.slot,
                       b.bmat, b.bke);
    NLH_TI_DISPATCH(b.ti_bytes,
                    hipLaunchKernelGGL(bucket_place_kernel<TI>, dim3((unsigned)P.nb), dim3(kBucketThreads),
                                       (size_t)(2 * (1 << P.bw) + 1) * 4, st, b.bke, b.X, b.bmat, P, b.nat, b.M,
                                       b.sys, b.nsys > 1 ? b.atom_sys + b.nat : nullptr, b.nsys, b.tmp,
                                       b.cell_start, b.cell_offsets_out, b.rec, b.wbig, b.perm_out,
                                       b.cell_id_out));
}

void launch_build(const PipelineBufs &b, hipStream_t st)
{
    if (b.nunits > 0) return;   // the fused build of small systems makes the cell list in the count stage
    if (b.nsys > 1)
        hipLaunchKernelGGL(sys_map_kernel, dim3(blocks_for(b.nat + b.M, 256)), dim3(256), 0, st, b.sys_atom_base,
                           b.sys_cell_base, b.nsys, b.nat, (int64_t)b.M, b.atom_sys);
    if (b.sort.bw >= 0) {
        launch_bucket_sort(b, st);
        return;
    }
    if (b.nat > 0)
        hipLaunchKernelGGL(bin_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.X, b.sys, b.atom_sys,
                           b.nsys, b.nat, b.key, b.slot, b.cell_count, b.status);
    const int64_t nbc = scan_tiles(b.M);
    int64_t *cpart = nullptr;   // three-kernel scan: tile sums in the (then unused) descriptor slots
    if (nbc > kLookbackTiles) {
        cpart = reinterpret_cast<int64_t *>(b.lookback);
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nbc), dim3(kScanThreads), 0, st, b.cell_count,
                           (int64_t)b.M, cpart, (int32_t)-1);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, cpart, (int)nbc,
                           (Status *)nullptr, (int64_t)0);
    }
    NLH_TI_DISPATCH(b.ti_bytes,
                    hipLaunchKernelGGL(scan_cells_kernel<TI>, dim3((unsigned)nbc), dim3(kScanThreads), 0, st,
                                       b.cell_count, (int64_t)b.M, b.lookback, cpart, b.cell_start, b.sys,
                                       b.nsys > 1 ? b.atom_sys + b.nat : nullptr, b.nsys, b.cell_offsets_out,
                                       b.status));
    if (b.nat > 0) {
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.key, b.slot,
                           b.cell_start, b.nat, b.tmp);
        NLH_TI_DISPATCH(b.ti_bytes,
                        hipLaunchKernelGGL(place_kernel<TI>, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.X,
                                           b.sys, b.nsys > 1 ? b.atom_sys + b.nat : nullptr, b.nsys, b.key,
                                           b.cell_start, b.tmp, b.nat, b.rec,
                                           b.wbig, b.perm_out, b.cell_id_out));
    }
}

void launch_scan_pairs(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = b.nat;
    const int64_t nbp = scan_tiles(n);
    uint64_t *desc = b.lookback + lookback_cells(b.sort.n_sort);
    int64_t *ppart = nullptr;   // three-kernel scan: tile sums in the (then unused) descriptor slots
    if (nbp > kLookbackTiles) {
        ppart = reinterpret_cast<int64_t *>(desc);
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nbp), dim3(kScanThreads), 0, st, b.pair_count, n, ppart,
                           kCountMask);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, ppart, (int)nbp, b.status,
                           b.pair_capacity);
    }
    NLH_TI_DISPATCH(b.ti_bytes,
                    hipLaunchKernelGGL(scan_pairs_kernel<TI>, dim3((unsigned)nbp), dim3(kScanThreads), 0, st,
                                       b.pair_count, n, desc, ppart, b.pair_off, b.nsys, b.first_out, b.status,
                                       b.pair_capacity, b.tile_atom, b.tile_cap, fill_tile_shift(b.fill_pairs)));
    if (b.nsys > 1)
        NLH_TI_DISPATCH(b.ti_bytes,
                        hipLaunchKernelGGL(batch_first_kernel<TI>,
                                           dim3(blocks_for(std::max<int64_t>(n, (int64_t)b.nsys + 1), 256)),
                                           dim3(256), 0, st, b.pair_off, n, b.sys_atom_base, b.atom_sys, b.nsys,
                                           b.first_out, b.sys_pair_base));
}

void launch_sorted_positions(const PipelineBufs &b, double *Xs, hipStream_t st)
{
    if (b.nat > 0)
        hipLaunchKernelGGL(sorted_positions_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.rec, b.nat, Xs);
}

}  // namespace nlh
