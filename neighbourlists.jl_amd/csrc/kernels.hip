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
// 1. binning + cell histogram
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void bin_kernel(const double *__restrict__ X, const SysDev *__restrict__ sys,
                                                  const int64_t *__restrict__ sys_atom_base, int nsys, int64_t nat,
                                                  int32_t *__restrict__ key, int32_t *__restrict__ slot,
                                                  int32_t *__restrict__ cell_count, Status *__restrict__ status)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nat) return;
    const int s = nsys == 1 ? 0 : find_sys(sys_atom_base, nsys, g);
    const SysDev &sd = sys[s];
    const double x = X[3 * g], y = X[3 * g + 1], z = X[3 * g + 2];
    int32_t c[3], w[3];
    if (!cell_and_winding(sd, x, y, z, c, w)) atomicOr(&status->bad_position, 1);
    const int32_t k = sd.cell_base + (c[0] - 1) + (c[1] - 1) * sd.n[0] + (c[2] - 1) * sd.n[0] * sd.n[1];
    key[g] = k;
    slot[g] = atomicAdd(&cell_count[k], 1);
}

// ---------------------------------------------------------------------------
// 2. scans (reduce -> scan partials -> downsweep)
// ---------------------------------------------------------------------------

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

// Exclusive scan of nb partials in place by one block; partials[nb] = total.
// For the pair scan it also records the total and the capacity verdict.
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
    if (threadIdx.x == 0) {
        partials[nb] = carry;
        if (status) {
            status->npairs = carry;
            status->overflow = carry > capacity ? 1 : 0;
        }
    }
}

// Downsweep for the cell histogram: cell_start (0-based), the caller-facing
// 1-based cell_offsets per system, and the histogram is cleared for the next
// build.
template <typename TI>
__global__ __launch_bounds__(kScanThreads) void scan_cells_down_kernel(
    int32_t *__restrict__ cell_count, int64_t M, const int64_t *__restrict__ partials, int32_t *__restrict__ cell_start,
    const SysDev *__restrict__ sys, const int32_t *__restrict__ sys_cell_base, int nsys, void *__restrict__ offsets_out)
{
    int32_t v[kScanItems];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    load_tile(cell_count, M, base, v);
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    int64_t run = block_excl_scan(s, tot) + (partials ? partials[blockIdx.x] : 0);
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t c = b + q;
        if (c < M) {
            cell_start[c] = (int32_t)run;
            cell_count[c] = 0;
            const int si = nsys == 1 ? 0 : find_sys(sys_cell_base, nsys, (int32_t)c);
            const SysDev &sd = sys[si];
            store_ti<TI>(offsets_out, c + si, run - sd.atom_base + 1);
            if (c == (int64_t)sd.cell_base + sd.M - 1)   // trailing entry of this system
                store_ti<TI>(offsets_out, c + si + 1, run + v[q] - sd.atom_base + 1);
            if (c == M - 1) cell_start[M] = (int32_t)(run + v[q]);
        }
        run += v[q];
    }
}

// Downsweep for the pair counts: pair_off (global, 0-based, int64) and, for a
// single system, `first` (1-based TI).  Also records the total and whether
// it fits the output capacity.
template <typename TI>
__global__ __launch_bounds__(kScanThreads) void scan_pairs_down_kernel(
    const int32_t *__restrict__ counts, int64_t nat, const int64_t *__restrict__ partials, int64_t *__restrict__ pair_off,
    int nsys, void *__restrict__ first_out, Status *__restrict__ status, int64_t capacity,
    int32_t *__restrict__ tile_atom, int64_t tile_cap, int tile_shift)
{
    int32_t v[kScanItems];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    load_tile(counts, nat, base, v, kCountMask);   // bit 30: slot-row format
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    const int64_t ex = block_excl_scan(s, tot);
    const int64_t boff = partials ? partials[blockIdx.x] : 0;
    if (gridDim.x == 1 && threadIdx.x == 0) {   // no partials pass: this block sees the total
        status->npairs = tot;
        status->overflow = tot > capacity ? 1 : 0;
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
    int64_t run = boff + ex;
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t a = b + q;
        if (a < nat) {
            pair_off[a] = run;
            // fill tiles whose first pair is one of this atom's: entry t is
            // the atom owning pair t << tile_shift and the number of its
            // pairs before that one
            const int64_t tsz = (int64_t)1 << tile_shift;
            for (int64_t t = (run + tsz - 1) >> tile_shift; t < tile_cap && (t << tile_shift) < run + v[q]; ++t) {
                tile_atom[2 * t] = (int32_t)a;
                tile_atom[2 * t + 1] = (int32_t)((t << tile_shift) - run);
            }
            if (nsys == 1) store_ti<TI>(first_out, a, run + 1);
            if (a == nat - 1) {
                pair_off[nat] = run + v[q];
                if (nsys == 1) store_ti<TI>(first_out, nat, run + v[q] + 1);
            }
        }
        run += v[q];
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
                                                    const int32_t *__restrict__ sys_cell_base, int nsys,
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
    const int s = nsys == 1 ? 0 : find_sys(sys_cell_base, nsys, c);
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

// Zero the build status.  A kernel rather than a memset, so the pipeline can
// be captured into a HIP graph and replayed.
__global__ void clear_status_kernel(Status *__restrict__ status)
{
    const int t = threadIdx.x;
    if (t < (int)(sizeof(Status) / 4)) reinterpret_cast<int32_t *>(status)[t] = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_clear_status(Status *status, hipStream_t st)
{
    hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(64), 0, st, status);
}

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

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
    if (b.nat > 0)
        hipLaunchKernelGGL(bin_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.X, b.sys, b.sys_atom_base,
                           b.nsys, b.nat, b.key, b.slot, b.cell_count, b.status);
    const int64_t nb = (b.M + kScanTile - 1) / kScanTile;
    if (nb > 1) {
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, b.cell_count,
                           (int64_t)b.M, b.scan_partials, (int32_t)-1);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, b.scan_partials, (int)nb,
                           (Status *)nullptr, (int64_t)0);
    }
    NLH_TI_DISPATCH(b.ti_bytes,
                    hipLaunchKernelGGL(scan_cells_down_kernel<TI>, dim3((unsigned)nb), dim3(kScanThreads), 0, st,
                                       b.cell_count, (int64_t)b.M, nb > 1 ? b.scan_partials : nullptr, b.cell_start,
                                       b.sys, b.sys_cell_base, b.nsys, b.cell_offsets_out));
    if (b.nat > 0) {
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.key, b.slot,
                           b.cell_start, b.nat, b.tmp);
        NLH_TI_DISPATCH(b.ti_bytes,
                        hipLaunchKernelGGL(place_kernel<TI>, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.X,
                                           b.sys, b.sys_cell_base, b.nsys, b.key, b.cell_start, b.tmp, b.nat, b.rec,
                                           b.wbig, b.inv_perm, b.perm_out, b.cell_id_out));
    }
}

void launch_scan_pairs(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = b.nat;
    const int64_t nb = n > 0 ? (n + kScanTile - 1) / kScanTile : 1;
    if (nb > 1) {
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, b.pair_count, n,
                           b.scan_partials, kCountMask);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, b.scan_partials, (int)nb,
                           b.status, b.pair_capacity);
    }
    NLH_TI_DISPATCH(b.ti_bytes,
                    hipLaunchKernelGGL(scan_pairs_down_kernel<TI>, dim3((unsigned)nb), dim3(kScanThreads), 0, st,
                                       b.pair_count, n, nb > 1 ? b.scan_partials : nullptr, b.pair_off, b.nsys,
                                       b.first_out, b.status, b.pair_capacity, b.tile_atom, b.tile_cap,
                                       fill_tile_shift(b.fill_pairs)));
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
