// Internal structures shared by the nlhip kernels and the host orchestration.
//
// Conventions (NeighbourLists.jl, src/types.jl:34-82):
//   - cell / inv_cell are 3x3 column-major (Julia SMatrix memory order),
//     element (m,k) at [m + 3k]; ROWS of cell are lattice vectors.
//   - indices handed back to the caller are 1-based; internal indices are
//     0-based.
//   - floating point follows the reference's non-fused left-to-right order;
//     every HIP source is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nlh {

constexpr int kWave = 64;
constexpr int kPairsBlock = 256;          // 4 waves, one home cell per wave
constexpr int kWavesPerPairsBlock = kPairsBlock / kWave;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;             // elements per thread in a scan tile
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kMinFillPairs = 128;        // smallest fill tile (pairs per fill wave)

// Single-pass scan descriptors (a ticket plus one per tile of kScanTile
// items): the cell sort's scan (n_sort items), then the pair scan's.
inline int64_t scan_tiles(int64_t n) { return n > 0 ? (n + kScanTile - 1) / kScanTile : 1; }
constexpr int64_t kLookbackTiles = 64;   // single-pass scans up to this many tiles
constexpr int64_t kSortLookbackTiles = 128;   // the bucketed sort's count scan: single pass up to here
inline int64_t lookback_cells(int64_t n_sort) { return scan_tiles(n_sort) + 1; }
inline int64_t lookback_entries(int64_t n_sort, int64_t nat) { return lookback_cells(n_sort) + scan_tiles(nat) + 1; }

// Cell sort.  Bucketed (the default): buckets are runs of 2^bw consecutive
// global cells.  Blocks of apb consecutive atoms count their atoms per
// bucket in LDS (na blocks, nb buckets), one scan over the nb x na counts
// (bucket-major) gives every block its place in every bucket, the atoms are
// scattered there, and one workgroup per bucket orders its atoms by (cell,
// original index).  Atomic (cell_sort = 1, or grids beyond the bucket
// limits): one global atomic histogram over the cells, scan, grouping and a
// stable in-cell rank.  Both give the same result.
constexpr int kBinThreads = 512;          // threads of the bucket count / scatter blocks
constexpr int kBinUnroll = 4;             // atoms a thread loads before it works on them
constexpr int kBucketThreads = 512;       // threads of a bucket's sorting workgroup
constexpr int kBucketLdsAtoms = 4096;     // buckets up to this size are ordered in LDS
constexpr int kBucketMaxShift = 12;       // at most 4096 cells per bucket
constexpr int kBucketMaxCount = 8192;     // buckets (LDS counters of a count block)
constexpr int kBucketBytes = 8;           // scratch per atom: (cell, index)
struct SortPlan {
    int32_t bw;     // log2 cells per bucket; -1: atomic histogram sort
    int32_t nb;     // buckets
    int32_t na;     // count / scatter blocks
    int32_t apb;    // atoms per count / scatter block (a multiple of kBinThreads)
    int64_t n_sort; // items of the sort's scan: nb * na, or M for the atomic sort
};
inline SortPlan sort_plan(int64_t M, int64_t nat, int cell_sort)
{
    SortPlan p{-1, 0, 0, 0, M};
    if (cell_sort != 0 || nat <= 0 || M <= 0) return p;
    // at least 256 buckets (workgroups of the ordering pass), about 1600
    // atoms each from 400k atoms on, at most 1024
    const int64_t target = nat / 1600 < 256 ? 256 : (nat / 1600 > 1024 ? 1024 : nat / 1600);
    int bw = 0;
    while (bw < kBucketMaxShift && ((M + (1ll << bw) - 1) >> bw) > target) ++bw;
    const int64_t nb = (M + (1ll << bw) - 1) >> bw;
    if (nb > kBucketMaxCount) return p;
    // as many count blocks (up to 512) as keep nb x na within a single-pass scan
    int64_t na = kSortLookbackTiles * kScanTile / nb;
    na = na < 1 ? 1 : (na > 512 ? 512 : na);
    int64_t apb = (nat + na - 1) / na;
    apb = (apb + kBinThreads - 1) / kBinThreads * kBinThreads;
    p.bw = bw;
    p.nb = (int32_t)nb;
    p.apb = (int32_t)apb;
    p.na = (int32_t)((nat + apb - 1) / apb);
    p.n_sort = nb * p.na;
    return p;
}

// log2 of the fill tile (pairs per fill wave: 128, 256 or 512)
inline int fill_tile_shift(int fill_pairs) { return fill_pairs == 128 ? 7 : (fill_pairs == 512 ? 9 : 8); }

// One system of a batch (a single build is a batch of one).
struct SysDev {
    double C[9];        // cell, column-major
    double inv[9];      // inv(cell), column-major
    double rc2;         // cutoff^2
    int32_t n[3];       // ncells per axis
    int32_t nxyz[3];    // neighbour-cell reach per axis
    int32_t pbc;        // bit k set = periodic along lattice vector k
    int32_t nat;        // atoms in this system
    int64_t atom_base;  // first global atom index
    int32_t cell_base;  // first global cell index
    int32_t M;          // cells in this system
    int64_t first_base; // offset of this system's `first` array (atom_base + sys index)
    int32_t brick[3];   // home cells per brick along each axis (pair search work unit)
    int32_t brick_base; // first global brick index
    // Launch order of the bricks: classes of one brick shape (along each
    // axis a full brick, or the partial last one of an axis that brick[k]
    // does not divide), in decreasing home cells, each class in row-major
    // order.  Class c holds the bricks [cls_first[c], cls_first[c + 1]) of
    // the system's launch order; cls_first[c] is the system's brick count
    // for every c past the last class.
    int32_t bfull[3];      // full bricks along each axis (n / brick)
    int32_t cls_first[9];
    int32_t cls_mask[8];   // bit k set: the class's bricks are the partial last brick along axis k
};

// Sorted atom record: position, original (system-local, 0-based) index and
// the packed winding numbers.  32 bytes, one dwordx4 pair per lane.
struct __attribute__((aligned(16))) Rec {
    double x, y, z;
    int32_t j;
    uint32_t w;         // 3 x 10-bit biased windings; bit 31 = stored in wbig
};

constexpr uint32_t kWindBig = 0x80000000u;


// Per-atom neighbour counts carry the format of the atom's slot row in bit
// 30: clear = 4-byte entries (j | S code << 26, every |S_k| <= 1 and j < 2^26),
// set = 8-byte entries (j | packed S << 32).
constexpr int32_t kRow8Bit = 1 << 30;
constexpr int32_t kCountMask = kRow8Bit - 1;
constexpr int32_t kNarrowMaxAtoms = 1 << 26;

// 2-bit code of a shift component in {-1, 0, 1} and back
__host__ __device__ inline uint32_t narrow_code(int32_t sx, int32_t sy, int32_t sz)
{
    return ((uint32_t)sx & 3u) | (((uint32_t)sy & 3u) << 2) | (((uint32_t)sz & 3u) << 4);
}
__host__ __device__ inline int32_t narrow_comp(uint32_t code, int k)
{
    return ((int32_t)(code << (30 - 2 * k))) >> 30;
}

__host__ __device__ inline uint32_t pack_wind(int32_t wx, int32_t wy, int32_t wz)
{
    if (wx < -512 || wx > 511 || wy < -512 || wy > 511 || wz < -512 || wz > 511) return kWindBig;
    return (uint32_t)(wx + 512) | ((uint32_t)(wy + 512) << 10) | ((uint32_t)(wz + 512) << 20);
}

__host__ __device__ inline void unpack_wind(uint32_t w, int32_t &wx, int32_t &wy, int32_t &wz)
{
    wx = (int32_t)(w & 1023u) - 512;
    wy = (int32_t)((w >> 10) & 1023u) - 512;
    wz = (int32_t)((w >> 20) & 1023u) - 512;
}

// Device-side status block, copied to pinned host memory once per build.
struct Status {
    int64_t npairs;          // total pairs over the batch
    int32_t scan_error;      // a scan block's look-back exceeded its bound (internal error)
    int32_t published;       // host copy only: set (last) by the fill that published the rest
    int32_t overflow;        // pair capacity exceeded: outputs not written
    int32_t bad_position;    // non-finite / out-of-range position seen
    int32_t wide;            // a cell shift does not fit a packed slot entry
    int32_t max_count;       // largest per-atom neighbour count
    int32_t small_fallback;  // the fused small-system build met a winding beyond +-511
};

// ---------------------------------------------------------------------------
// Cell-index algebra (src/cell_list.jl:66-147)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int32_t floordiv_pos(int32_t a, int32_t b)
{
    int32_t q = a / b;
    return (a % b != 0 && a < 0) ? q - 1 : q;
}

// wrap_and_shift (cell_list.jl:105-120): wrapped index in [1,n] and shift,
// or clamp with zero shift on an open axis.
__device__ __forceinline__ void wrap_and_shift(int32_t i, int32_t n, bool p, int32_t &w, int32_t &s)
{
    if (!p) {
        w = i < 1 ? 1 : (i > n ? n : i);
        s = 0;
        return;
    }
    if (i >= 1 && i <= n) {
        w = i;
        s = 0;
        return;
    }
    int32_t q = floordiv_pos(i - 1, n);
    w = i - q * n;
    s = q;
}

// position_to_cell_index (cell_list.jl:66-74): frac = inv_cell' * x,
// c_k = trunc(floor(frac_k * n_k + 1)).  Returns false when the value does
// not fit a 32-bit cell index (the reference's unsafe_trunc is undefined
// there).
__device__ __forceinline__ bool cell_index(const SysDev &s, double x, double y, double z, int32_t c[3])
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double f = (s.inv[0 + 3 * k] * x + s.inv[1 + 3 * k] * y) + s.inv[2 + 3 * k] * z;
        const double v = floor(f * (double)s.n[k] + 1.0);
        ok = ok && (v > -1.0e9 && v < 1.0e9);   // also false for NaN
        c[k] = ok ? (int32_t)v : 1;
    }
    return ok;
}

// position_to_cell_index + bin_wrap_and_shift (cell_list.jl:141-147)
__device__ __forceinline__ bool cell_and_winding(const SysDev &s, double x, double y, double z,
                                                 int32_t c[3], int32_t w[3])
{
    int32_t c0[3];
    const bool ok = cell_index(s, x, y, z, c0);
#pragma unroll
    for (int k = 0; k < 3; ++k) wrap_and_shift(c0[k], s.n[k], (s.pbc >> k) & 1, c[k], w[k]);
    return ok;
}

// Largest s in [0, n) with base[s] <= x (base ascending, base[0] == 0).
template <typename T>
__device__ __forceinline__ int find_sys(const T *__restrict__ base, int n, T x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Host-side launchers (kernels.hip)
struct PipelineBufs {
    // inputs
    const double *X;            // 3 * nat_total, AoS
    const SysDev *sys;          // nsys
    const int64_t *sys_atom_base;   // nsys + 1
    const int32_t *sys_cell_base;   // nsys + 1
    const int32_t *sys_brick_base;  // nsys + 1
    int32_t nsys;
    int32_t *atom_sys;          // batches: system of each atom (nat), then of each cell (M); else null
    int64_t nbricks;            // total pair-search bricks
    int64_t nat;                // total atoms
    int32_t M;                  // total cells
    // build
    SortPlan sort;
    int32_t *key;               // nat: global cell id
    int32_t *slot;              // nat: arrival rank (histogram atomics, or the count block's LDS counters)
    int32_t *cell_count;        // M + 1, zero between builds (atomic sort)
    int32_t *bmat;              // nb * na bucket counts, then their exclusive scan (bucketed sort)
    int2 *bke;                  // nat: (global cell, global index) grouped by bucket (bucketed sort)
    int32_t *cell_start;        // M + 1
    int32_t *tmp;               // nat: cell-grouped, not yet stable
    Rec *rec;                   // nat: sorted records
    int32_t *wbig;              // 3 * nat: windings that do not fit the pack
    void *perm_out;             // nat, TI, 1-based
    void *cell_id_out;          // nat, TI, 1-based
    void *cell_offsets_out;     // M + nsys, TI, 1-based per system
    // pairs
    int32_t *pair_count;
    void *slots;                // nat * slot_cap slot entries
    double *slotsR;             // 3 * nat * slot_cap (R requested)
    int32_t slot_cap;
    int64_t *pair_off;          // nat + 1
    int32_t *tile_atom;         // 2 tile_cap: atom owning pair t * fill_pairs, its pairs before it
    int64_t tile_cap;
    uint64_t *lookback;         // single-pass scan descriptors: cells (ticket + tiles), then pairs
    int64_t *sys_pair_base;     // nsys + 1 (device)
    void *first_out;            // nat + nsys, TI, 1-based per system
    void *i_out, *j_out, *S_out;
    double *R_out;
    int64_t pair_capacity;
    Status *status;
    Status *status_pub;         // device address of the handle's pinned host copy of the status
    int ti_bytes;
    // fused build of small systems (units of consecutive systems, one
    // workgroup each): 0 units = the general pipeline
    int32_t nunits = 0;
    int32_t nitems = 0;                   // work items: parts of units, one workgroup each
    const int32_t *units = nullptr;       // nunits + 1: first system of each unit
    const int32_t *unit_order = nullptr;  // 2 nitems: unit, part | parts << 16 (launch order, largest first)
    int fill_block;          // threads per fill workgroup (multiple of 64)
    int fill_pairs;          // output pairs per fill wave (128, 256 or 512)
    int count_block;         // threads per count-pass brick workgroup (256 or 512)
};

void launch_clear_status(const PipelineBufs &b, hipStream_t st);
// Zero n 8-byte words by a kernel (no memset node: safe under graph capture).
void launch_zero_words(uint64_t *p, int64_t n, hipStream_t st);
void launch_clear_pair_lookback(const PipelineBufs &b, hipStream_t st);
void launch_build(const PipelineBufs &b, hipStream_t st);
void launch_count(const PipelineBufs &b, bool want_R, bool wide, hipStream_t st);
void launch_scan_pairs(const PipelineBufs &b, hipStream_t st);
// Returns whether a fill kernel was launched (it then also publishes the
// status to status_pub).
bool launch_fill(const PipelineBufs &b, bool want_R, bool wide, hipStream_t st);
void launch_sorted_positions(const PipelineBufs &b, double *Xs, hipStream_t st);

// Fused build of small systems (b.nunits > 0): the cell list of every
// system, its pair counts and its slot rows, by one workgroup per work item
// (a unit of consecutive systems or one part of one); launched by
// launch_count, while launch_build launches nothing for such a batch.
constexpr int kSmallThreads = 1024;
constexpr int kUnitAtoms = 2048;      // atoms of one unit (and of one system) at most
constexpr int kUnitCells = 2048;      // cells of one unit at most
constexpr int kUnitSystems = 256;     // systems of one unit at most
constexpr int kUnitMaxReach = 3;      // neighbour-cell reach per axis at most
constexpr int kUnitMaxAxisCells = 256;   // cells per axis of one system at most
void launch_small(const PipelineBufs &b, bool want_R, hipStream_t st);

}  // namespace nlh
