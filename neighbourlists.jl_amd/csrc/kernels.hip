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
//   pairs_kernel<0>   per-atom neighbour counts (count_neighbours_kernel!,
//                     gpu_kernels.jl:132-152)
//   scan (pairs)      first / pair offsets + the i array
//                     (compute_pair_offsets, gpu_kernels.jl:222-237)
//   pairs_kernel<1>   j, S (and optionally R) (fill_pairs_kernel!,
//                     gpu_kernels.jl:159-180)
//
// The pair kernels run one wave per home cell: the wave enumerates the
// neighbour cells in the reference order (dz, dy, dx; gpu_kernels.jl:69-71),
// flattens their atoms into 64-candidate chunks (one candidate per lane) and
// loops over the home atoms; each (home atom, chunk) step is one ballot, so a
// home atom's neighbours come out in exactly the reference traversal order.
#include "nlhip_internal.h"

namespace nlh {

// ---------------------------------------------------------------------------
// wave / block helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ int64_t wave_incl_scan64(int64_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
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

// Exclusive scan over a 256-thread block; returns this thread's exclusive
// prefix and the block total.
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t &total)
{
    __shared__ int64_t wsum[kScanThreads / kWave];
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int64_t inc = wave_incl_scan64(v);
    if (lane == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    int64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; ++w) {
        const int64_t s = wsum[w];
        if (w < wv) before += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
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

__device__ __forceinline__ void load_tile(const int32_t *__restrict__ in, int64_t n, int64_t base, int32_t v[kScanItems])
{
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    if (b + kScanItems <= n) {
        const int4 *p = reinterpret_cast<const int4 *>(in + b);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            const int4 t = p[q];
            v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) v[q] = (b + q < n) ? in[b + q] : 0;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const int32_t *__restrict__ in, int64_t n,
                                                                   int64_t *__restrict__ partials)
{
    int32_t v[kScanItems];
    load_tile(in, n, (int64_t)blockIdx.x * kScanTile, v);
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

// Downsweep for the pair counts: pair_off (global, 0-based, int64), and --
// when the outputs fit the capacity -- the i array, written cooperatively so
// every store is coalesced.  For a single system it also writes `first`.
template <typename TI>
__global__ __launch_bounds__(kScanThreads) void scan_pairs_down_kernel(
    const int32_t *__restrict__ counts, int64_t nat, const int64_t *__restrict__ partials, int64_t *__restrict__ pair_off,
    const int64_t *__restrict__ sys_atom_base, int nsys, void *__restrict__ first_out, void *__restrict__ i_out,
    Status *__restrict__ status, int64_t capacity)
{
    __shared__ int32_t rel[kScanTile + 1];
    int32_t v[kScanItems];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    load_tile(counts, nat, base, v);
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) s += v[q];
    int64_t tot;
    const int64_t ex = block_excl_scan(s, tot);
    const int64_t boff = partials ? partials[blockIdx.x] : 0;
    int overflow;
    if (gridDim.x == 1) {   // no partials pass: this block sees the total
        overflow = tot > capacity;
        if (threadIdx.x == 0) {
            status->npairs = tot;
            status->overflow = overflow;
        }
    } else {
        overflow = partials[gridDim.x] > capacity;
    }
    int64_t run = ex;
    const int64_t b = base + (int64_t)threadIdx.x * kScanItems;
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t a = b + q;
        rel[threadIdx.x * kScanItems + q] = (int32_t)run;
        if (a < nat) {
            pair_off[a] = boff + run;
            if (nsys == 1) store_ti<TI>(first_out, a, boff + run + 1);
            if (a == nat - 1) {
                pair_off[nat] = boff + run + v[q];
                if (nsys == 1) store_ti<TI>(first_out, nat, boff + run + v[q] + 1);
            }
        }
        run += v[q];
    }
    if (threadIdx.x == 0) rel[kScanTile] = (int32_t)tot;
    __syncthreads();
    if (overflow || i_out == nullptr) return;
    // i array: positions [boff, boff + tot) belong to atoms base .. base+tile
    const int64_t nt = nat - base < kScanTile ? nat - base : kScanTile;
    for (int64_t q = threadIdx.x; q < tot; q += kScanThreads) {
        int lo = 0, hi = (int)nt - 1;   // largest t with rel[t] <= q
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (rel[mid] <= q) lo = mid;
            else hi = mid - 1;
        }
        const int64_t a = base + lo;
        const int64_t ab = nsys == 1 ? 0 : sys_atom_base[find_sys(sys_atom_base, nsys, a)];
        store_ti<TI>(i_out, boff + q, a - ab + 1);
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
// 4. pair search: count (FILL=false) and fill (FILL=true)
// ---------------------------------------------------------------------------

struct PairsArgs {
    const SysDev *sys;
    const int32_t *sys_cell_base;
    int nsys;
    int32_t M;
    const int32_t *cell_start;
    const Rec *rec;
    const int32_t *wbig;
    int32_t *pair_count;
    const int64_t *pair_off;
    void *j_out;
    void *S_out;
    double *R_out;
    const Status *status;
};

__device__ __forceinline__ void rec_winding(const Rec &r, const int32_t *__restrict__ wbig, int64_t p, int32_t &wx,
                                            int32_t &wy, int32_t &wz)
{
    if (r.w & kWindBig) {
        wx = wbig[3 * p];
        wy = wbig[3 * p + 1];
        wz = wbig[3 * p + 2];
    } else {
        unpack_wind(r.w, wx, wy, wz);
    }
}

template <bool FILL, bool WANT_R, typename TI>
__global__ __launch_bounds__(kPairsBlock) void pairs_kernel(PairsArgs A)
{
    const int lane = lane_id();
    const int64_t cell = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kWavesPerPairsBlock + (threadIdx.x >> 6);
    if (cell >= A.M) return;
    const int32_t hs = A.cell_start[cell], he = A.cell_start[cell + 1];
    if (hs >= he) return;
    if (FILL && A.status->overflow) return;
    const int s = A.nsys == 1 ? 0 : find_sys(A.sys_cell_base, A.nsys, (int32_t)cell);
    const SysDev &sd = A.sys[s];
    const int n0 = sd.n[0], n1 = sd.n[1], n2 = sd.n[2];
    const int lc = (int)cell - sd.cell_base;
    const int ci0 = lc % n0 + 1, ci1 = (lc / n0) % n1 + 1, ci2 = lc / (n0 * n1) + 1;
    const int nx0 = sd.nxyz[0], nx1 = sd.nxyz[1], nx2 = sd.nxyz[2];
    const int sx = 2 * nx0 + 1, sy = 2 * nx1 + 1;
    const int nseg = sx * sy * (2 * nx2 + 1);
    const bool p0 = sd.pbc & 1, p1 = (sd.pbc >> 1) & 1, p2 = (sd.pbc >> 2) & 1;
    double Cm[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Cm[q] = sd.C[q];
    const double rc2 = sd.rc2;
    const int64_t abase = sd.atom_base;
    const int32_t cbase = sd.cell_base;
    const Rec *__restrict__ rec = A.rec;

    for (int32_t hg = hs; hg < he; hg += kWave) {
        const int nh = min(kWave, he - hg);
        // lane a keeps the running count / write cursor of home atom hg + a
        int64_t mine = 0;
        if (FILL && lane < nh) mine = A.pair_off[abase + rec[hg + lane].j];

        for (int sb = 0; sb < nseg; sb += kWave) {
            // ---- neighbour cells of this batch, one per lane, reference order
            const int g = sb + lane;
            int cnt_g = 0, start_g = 0;
            uint32_t shp = 0;
            if (g < nseg) {
                const int dx = g % sx - nx0;
                const int t = g / sx;
                const int dy = t % sy - nx1;
                const int dz = t / sy - nx2;
                const int a0 = ci0 + dx, a1 = ci1 + dy, a2 = ci2 + dz;
                // _is_cell_in_bounds (gpu_kernels.jl:20-25)
                const bool inb = (p0 || (a0 >= 1 && a0 <= n0)) && (p1 || (a1 >= 1 && a1 <= n1)) &&
                                 (p2 || (a2 >= 1 && a2 <= n2));
                if (inb) {
                    int32_t c0, c1, c2, s0, s1, s2;   // _get_neighbor_cell (:39-47)
                    wrap_and_shift(a0, n0, p0, c0, s0);
                    wrap_and_shift(a1, n1, p1, c1, s1);
                    wrap_and_shift(a2, n2, p2, c2, s2);
                    const int gc = cbase + (c0 - 1) + (c1 - 1) * n0 + (c2 - 1) * n0 * n1;
                    start_g = A.cell_start[gc];
                    cnt_g = A.cell_start[gc + 1] - start_g;
                    shp = pack_wind(s0, s1, s2);
                }
            }
            const int incl = wave_incl_scan(cnt_g);
            const int total = __shfl(incl, kWave - 1, kWave);
            const int excl = incl - cnt_g;

            for (int k0 = 0; k0 < total; k0 += kWave) {
                // ---- candidate k of the flattened list, one per lane
                const int k = k0 + lane;
                const bool kv = k < total;
                int m = 0;
#pragma unroll
                for (int step = 32; step >= 1; step >>= 1) {
                    const int e2 = __shfl(excl, m + step, kWave);
                    if (e2 <= k) m += step;
                }
                const int p = __shfl(start_g, m, kWave) + (k - __shfl(excl, m, kWave));
                const uint32_t shm = __shfl(shp, m, kWave);
                Rec r;
                int32_t wjx = 0, wjy = 0, wjz = 0;
                if (kv) {
                    r = rec[p];
                    rec_winding(r, A.wbig, p, wjx, wjy, wjz);
                } else {
                    r.x = r.y = r.z = 0.0;
                    r.j = -1;
                    r.w = 0;
                }
                int32_t shx, shy, shz;
                unpack_wind(shm, shx, shy, shz);
                const bool zero_shift = (shx == 0) & (shy == 0) & (shz == 0);
                // S = s_loop + wi - wj = (s_loop - wj) + wi
                const int32_t tx = shx - wjx, ty = shy - wjy, tz = shz - wjz;
                uint32_t wcache = 0xffffffffu;
                int32_t Sx = 0, Sy = 0, Sz = 0;
                double csx = 0.0, csy = 0.0, csz = 0.0;

                for (int a = 0; a < nh; ++a) {
                    const Rec h = rec[hg + a];
                    if (h.w != wcache || (h.w & kWindBig)) {   // uniform branch
                        int32_t wix, wiy, wiz;
                        rec_winding(h, A.wbig, hg + a, wix, wiy, wiz);
                        Sx = tx + wix;
                        Sy = ty + wiy;
                        Sz = tz + wiz;
                        const double fx = (double)Sx, fy = (double)Sy, fz = (double)Sz;
                        // (C' * S)_k = sum_m C[m,k] S_m  (gpu_kernels.jl:89)
                        csx = (Cm[0] * fx + Cm[1] * fy) + Cm[2] * fz;
                        csy = (Cm[3] * fx + Cm[4] * fy) + Cm[5] * fz;
                        csz = (Cm[6] * fx + Cm[7] * fy) + Cm[8] * fz;
                        wcache = h.w;
                    }
                    // R = xj - xi + C' * S ; r^2 = dot(R, R)  (gpu_kernels.jl:89-92)
                    const double Rx = (r.x - h.x) + csx;
                    const double Ry = (r.y - h.y) + csy;
                    const double Rz = (r.z - h.z) + csz;
                    const double r2 = (Rx * Rx + Ry * Ry) + Rz * Rz;
                    const bool self = zero_shift & (r.j == h.j);   // _is_self_interaction
                    const bool hit = kv & !self & (r2 < rc2);
                    const uint64_t mask = __ballot(hit);
                    if (!FILL) {
                        if (lane == a) mine += __popcll(mask);
                    } else {
                        const int64_t base = readlane64(mine, a);
                        if (hit) {
                            const int64_t pos = base + mbcnt64(mask);
                            store_ti<TI>(A.j_out, pos, (int64_t)r.j + 1);
                            TI *S = reinterpret_cast<TI *>(A.S_out) + 3 * pos;
                            S[0] = (TI)Sx;
                            S[1] = (TI)Sy;
                            S[2] = (TI)Sz;
                            if (WANT_R) {
                                double *R = A.R_out + 3 * pos;
                                R[0] = Rx;
                                R[1] = Ry;
                                R[2] = Rz;
                            }
                        }
                        if (lane == a) mine += __popcll(mask);
                    }
                }
            }
        }
        if (!FILL && lane < nh) A.pair_count[abase + rec[hg + lane].j] = (int32_t)mine;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

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
                           (int64_t)b.M, b.scan_partials);
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
                                           b.wbig, b.perm_out, b.cell_id_out));
    }
}

static PairsArgs pairs_args(const PipelineBufs &b)
{
    PairsArgs a;
    a.sys = b.sys;
    a.sys_cell_base = b.sys_cell_base;
    a.nsys = b.nsys;
    a.M = b.M;
    a.cell_start = b.cell_start;
    a.rec = b.rec;
    a.wbig = b.wbig;
    a.pair_count = b.pair_count;
    a.pair_off = b.pair_off;
    a.j_out = b.j_out;
    a.S_out = b.S_out;
    a.R_out = b.R_out;
    a.status = b.status;
    return a;
}

void launch_count(const PipelineBufs &b, hipStream_t st)
{
    if (b.nat == 0) return;
    const unsigned nblk = blocks_for(b.M, kWavesPerPairsBlock);
    hipLaunchKernelGGL((pairs_kernel<false, false, int32_t>), dim3(nblk), dim3(kPairsBlock), 0, st, pairs_args(b));
}

void launch_scan_pairs(const PipelineBufs &b, hipStream_t st)
{
    const int64_t n = b.nat;
    const int64_t nb = n > 0 ? (n + kScanTile - 1) / kScanTile : 1;
    if (nb > 1) {
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, b.pair_count, n,
                           b.scan_partials);
        hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, b.scan_partials, (int)nb,
                           b.status, b.pair_capacity);
    }
    NLH_TI_DISPATCH(b.ti_bytes,
                    hipLaunchKernelGGL(scan_pairs_down_kernel<TI>, dim3((unsigned)nb), dim3(kScanThreads), 0, st,
                                       b.pair_count, n, nb > 1 ? b.scan_partials : nullptr, b.pair_off,
                                       b.sys_atom_base, b.nsys, b.first_out, b.i_out, b.status, b.pair_capacity));
    if (b.nsys > 1)
        NLH_TI_DISPATCH(b.ti_bytes,
                        hipLaunchKernelGGL(batch_first_kernel<TI>, dim3(blocks_for(n + b.nsys + 1, 256)), dim3(256),
                                           0, st, b.pair_off, n, b.sys_atom_base, b.nsys, b.first_out,
                                           b.sys_pair_base));
}

void launch_fill(const PipelineBufs &b, bool want_R, hipStream_t st)
{
    if (b.nat == 0) return;
    const unsigned nblk = blocks_for(b.M, kWavesPerPairsBlock);
    PairsArgs a = pairs_args(b);
    if (b.ti_bytes == 8) {
        if (want_R) hipLaunchKernelGGL((pairs_kernel<true, true, int64_t>), dim3(nblk), dim3(kPairsBlock), 0, st, a);
        else hipLaunchKernelGGL((pairs_kernel<true, false, int64_t>), dim3(nblk), dim3(kPairsBlock), 0, st, a);
    } else {
        if (want_R) hipLaunchKernelGGL((pairs_kernel<true, true, int32_t>), dim3(nblk), dim3(kPairsBlock), 0, st, a);
        else hipLaunchKernelGGL((pairs_kernel<true, false, int32_t>), dim3(nblk), dim3(kPairsBlock), 0, st, a);
    }
}

void launch_sorted_positions(const PipelineBufs &b, double *Xs, hipStream_t st)
{
    if (b.nat > 0)
        hipLaunchKernelGGL(sorted_positions_kernel, dim3(blocks_for(b.nat, 256)), dim3(256), 0, st, b.rec, b.nat, Xs);
}

}  // namespace nlh
