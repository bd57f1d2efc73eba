// nlhip device kernels for the sort-based neighbour-list build (gfx950, wave64).
//
// Pipeline (a single build is a batch of one system):
//   bin_kernel        cell id + winding per atom, cell histogram by atomics
//                     (compute_cell_ids_kernel!, gpu_kernels.jl:108-125, and
//                     _histogram_kernel!, :191-197)
//   scan (cells)      cell_start / cell_offsets (_compute_cell_offsets,
//                     gpu_kernels.jl:262-285)
//   scatter_kernel    group atom indices by cell (arrival order)
//   place_kernel      stable in-cell rank -> perm, cell_id, sorted records
//                     (sortperm + gathers, cell_list.jl:668-670)
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
    if (lane_id() == 0 && mx > 0) atomicMax(&status->max_count, mx);
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
// pair bases.
template <typename TI>
__global__ __launch_bounds__(256) void batch_first_kernel(const int64_t *__restrict__ pair_off, int64_t nat,
                                                          const int64_t *__restrict__ sys_atom_base, int nsys,
                                                          void *__restrict__ first_out, int64_t *__restrict__ sys_pair_base)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nsys + 1) sys_pair_base[g] = pair_off[sys_atom_base[g]];
    if (g >= nat + nsys) return;
    // entry g of the concatenated first arrays: system s owns entries
    // [atom_base[s] + s, atom_base[s+1] + s]
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sys_atom_base[mid] + mid <= g) lo = mid;
        else hi = mid - 1;
    }
    const int64_t a = g - lo;   // global atom index (or one past the system's end)
    store_ti<TI>(first_out, g, pair_off[a] - pair_off[sys_atom_base[lo]] + 1);
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
                                                    int32_t *__restrict__ wbig, int32_t *__restrict__ inv_perm,
                                                    void *__restrict__ perm_out, void *__restrict__ cell_id_out)
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
    inv_perm[e] = p;
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
    const int64_t n = lookback_entries(b.M, b.nat);
    hipLaunchKernelGGL(clear_status_kernel, dim3(std::min(blocks_for(n, 256), 64u)), dim3(256), 0, st, b.status,
                       b.lookback, n);
}

void launch_clear_pair_lookback(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = scan_tiles(b.nat) + 1;
    hipLaunchKernelGGL(clear_status_kernel, dim3(std::min(blocks_for(n, 256), 64u)), dim3(256), 0, st,
                       (Status *)nullptr, b.lookback + lookback_cells(b.M), n);
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

void launch_build(const PipelineBufs &b, hipStream_t st)
{
    if (b.nsys > 1)
        hipLaunchKernelGGL(sys_map_kernel, dim3(blocks_for(b.nat + b.M, 256)), dim3(256), 0, st, b.sys_atom_base,
                           b.sys_cell_base, b.nsys, b.nat, (int64_t)b.M, b.atom_sys);
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
                                           b.wbig, b.inv_perm, b.perm_out, b.cell_id_out));
    }
}

void launch_scan_pairs(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = b.nat;
    const int64_t nbp = scan_tiles(n);
    uint64_t *desc = b.lookback + lookback_cells(b.M);
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
                        hipLaunchKernelGGL(batch_first_kernel<TI>, dim3(blocks_for(n + b.nsys + 1, 256)), dim3(256),
                                           0, st, b.pair_off, n, b.sys_atom_base, b.nsys, b.first_out,
                                           b.sys_pair_base));
}

void launch_sorted_positions(const PipelineBufs &b, double *Xs, hipStream_t st)
{
    if (b.nat > 0)
        hipLaunchKernelGGL(sorted_positions_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.rec, b.nat, Xs);
}

}  // namespace nlh
