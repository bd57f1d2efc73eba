// Pair search of the sort-based neighbour list: count and fill, the MI355X
// counterparts of count_neighbours_kernel! and fill_pairs_kernel!
// (src/gpu_kernels.jl:132-180) around the inner loop _for_each_neighbor_pair
// (src/gpu_kernels.jl:58-101).
//
// count_kernel -- work unit: a brick of up to 4x4x4 home cells per 256-thread
// (or 512-thread) workgroup, the bricks launched largest first.  The brick
// and its one-cell halo are staged in LDS in one load phase (every global
// load of the phase in flight together), then the waves take home cells from
// a workgroup counter and test candidates from LDS only:
//   - the candidates of a home cell are its 27 neighbour cells flattened in
//     the reference order (dz, dy, dx; gpu_kernels.jl:69-71) -- nine
//     contiguous LDS rows -- cut into 64-candidate chunks, one per lane;
//   - the home atoms stream through the wave as uniform values, two at a
//     time; each (home atom, chunk) test is one ballot, and the hit lanes
//     append their (j, S) to the home atom's slot row at its running count,
//     so every row holds the atom's neighbours in exactly the reference
//     traversal order.
// Bricks whose neighbour reach exceeds one cell, whose halo does not fit
// LDS, or whose windings do not pack take a per-wave path that reads global
// memory (one home cell per wave) with the same enumeration and output.
//
// fill_kernel -- one wave per tile of consecutive output pairs, one lane per
// pair: the tile's first atom comes from the pair scan's tile table, and
// the lanes copy the slot-row entries of the atoms the tile covers into the
// PairList arrays (i, j, S and optionally R), so every store is linear.
#include <algorithm>

#include "device_util.h"

namespace nlh {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int sgpr(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The next work item of a wave from a workgroup counter (wave-uniform).
__device__ __forceinline__ int take_next(int32_t *ctr)
{
    int t = 0;
    if (lane_id() == 0) t = atomicAdd(ctr, 1);
    return __builtin_amdgcn_readfirstlane(t);
}

// v with lane `lane` (wave-uniform) set to the wave-uniform x: one
// v_writelane, no exec-mask or compare work.
__device__ __forceinline__ int32_t write_lane(int32_t x, int lane, int32_t v)
{
    // the lane select in M0: the value's SGPR is the one constant-bus read
    __asm__("v_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(sgpr(x)), "{m0}"(sgpr(lane)));
    return v;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    const uint32_t lo = __shfl((uint32_t)v, src, kWave), hi = __shfl((uint32_t)(v >> 32), src, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// (C' * S)_k = sum_m C[m,k] S_m   (gpu_kernels.jl:89)
__device__ __forceinline__ void cell_shift(const double *__restrict__ Cm, int32_t Sx, int32_t Sy, int32_t Sz, double &cx,
                                           double &cy, double &cz)
{
    const double fx = (double)Sx, fy = (double)Sy, fz = (double)Sz;
    cx = (Cm[0] * fx + Cm[1] * fy) + Cm[2] * fz;
    cy = (Cm[3] * fx + Cm[4] * fy) + Cm[5] * fz;
    cz = (Cm[6] * fx + Cm[7] * fy) + Cm[8] * fz;
}

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

constexpr int kChunksPerPass = 4;   // candidate chunks a lane holds at once
constexpr double kFar = 1.0e200;    // invalid candidate: r^2 overflows to +inf, never a hit

struct PairsArgs {
    const SysDev *__restrict__ sys;
    const int32_t *__restrict__ sys_cell_base;
    int nsys;
    int32_t M;
    int64_t nat;
    const int32_t *__restrict__ cell_start;
    const Rec *__restrict__ rec;
    const int32_t *__restrict__ wbig;
};

// ---------------------------------------------------------------------------
// slots: each atom's neighbours (j, S) in traversal order, a row of `cap`
// entries per atom, rows in original atom order.  A row holds 4-byte entries
// (j | S code << 26) when every shift of its brick is within [-1, 1]^3, else
// 8-byte entries (j | S packed in 3 x 10 bits << 32); bit 30 of the atom's
// count tells which.  Wide builds (shifts beyond +-511) use 16-byte entries.
// ---------------------------------------------------------------------------

template <bool WIDE>
struct SlotEntry;
template <>
struct SlotEntry<false> {
    using T = uint64_t;
    __device__ static void get(T e, int32_t &j, int32_t &Sx, int32_t &Sy, int32_t &Sz)
    {
        j = (int32_t)(uint32_t)e;
        unpack_wind((uint32_t)(e >> 32), Sx, Sy, Sz);
    }
};
template <>
struct SlotEntry<true> {
    using T = int4;
    __device__ static void get(T e, int32_t &j, int32_t &Sx, int32_t &Sy, int32_t &Sz)
    {
        j = e.x; Sx = e.y; Sy = e.z; Sz = e.w;
    }
};

struct SlotArgs {
    void *__restrict__ slots;
    double *__restrict__ slotsR;     // 3 doubles per entry (NLH_WANT_R)
    int32_t cap;
    Status *__restrict__ status;
};

// Append the hits of one (home atom, chunk) ballot to the home atom's slot.
// sbase = sorted index of the home atom * cap; ca = its running count
// (wave-uniform).  t_pk: this lane's packed s - w_j (valid when tfits);
// (tx, ty, tz) the same unpacked; w_i the home atom's winding, wzero when 0.
template <bool WANT_R, bool WIDE>
__device__ __forceinline__ void emit_hits(const SlotArgs &O, uint64_t mask, int64_t sbase, int32_t &ca, int32_t cj,
                                          uint32_t t_pk, int32_t tx, int32_t ty, int32_t tz, int32_t wix, int32_t wiy,
                                          int32_t wiz, bool wzero, double Rx, double Ry, double Rz)
{
    if (mask == 0) return;   // wave-uniform
    const int lane = lane_id();
    const int32_t rank = ca + (int32_t)mbcnt64(mask);
    if (((mask >> lane) & 1) && rank < O.cap) {
        const int64_t sidx = sbase + rank;
        if (WIDE) {
            reinterpret_cast<int4 *>(O.slots)[sidx] = make_int4(cj, tx + wix, ty + wiy, tz + wiz);
        } else {
            // S = s_loop + w_i - w_j = t + w_i  (gpu_kernels.jl:88)
            const uint32_t spk = wzero ? t_pk : pack_wind(tx + wix, ty + wiy, tz + wiz);
            if (spk & kWindBig) O.status->wide = 1;
            reinterpret_cast<uint64_t *>(O.slots)[sidx] = (uint64_t)(uint32_t)cj | ((uint64_t)spk << 32);
        }
        if (WANT_R) {
            double *R = O.slotsR + 3 * sidx;
            R[0] = Rx;
            R[1] = Ry;
            R[2] = Rz;
        }
    }
    ca += __popcll(mask);
}

// ---------------------------------------------------------------------------
// per-wave path: one home cell per wave, candidates read from global memory
// ---------------------------------------------------------------------------

struct HomeCell {
    int32_t hs, he;             // sorted atom range
    int ci0, ci1, ci2;          // 1-based cell coordinates
    int n0, n1, n2, nx0, nx1, nx2, sx, sy, nseg;
    bool p0, p1, p2;
    int32_t cbase;
    int64_t abase;
    const SysDev *sd;
};

__device__ __forceinline__ void home_cell(const PairsArgs &A, const SysDev &sd, int cell, HomeCell &H)
{
    H.hs = A.cell_start[cell];
    H.he = A.cell_start[cell + 1];
    H.sd = &sd;
    H.n0 = sd.n[0]; H.n1 = sd.n[1]; H.n2 = sd.n[2];
    H.cbase = sd.cell_base;
    H.abase = sd.atom_base;
    const int lc = cell - H.cbase;
    H.ci0 = lc % H.n0 + 1;
    H.ci1 = (lc / H.n0) % H.n1 + 1;
    H.ci2 = lc / (H.n0 * H.n1) + 1;
    H.nx0 = sd.nxyz[0]; H.nx1 = sd.nxyz[1]; H.nx2 = sd.nxyz[2];
    H.sx = 2 * H.nx0 + 1;
    H.sy = 2 * H.nx1 + 1;
    H.nseg = H.sx * H.sy * (2 * H.nx2 + 1);
    H.p0 = sd.pbc & 1; H.p1 = (sd.pbc >> 1) & 1; H.p2 = (sd.pbc >> 2) & 1;
}

// Neighbour cells sb .. sb+63 of the home cell, one per lane, in the
// reference order.  Returns the number of candidate atoms in the batch
// (wave-uniform); per lane: exclusive prefix, first sorted index, shift.
__device__ __forceinline__ int segment_batch(const PairsArgs &A, const HomeCell &H, int sb, int &excl, int &start_g,
                                             uint32_t &shp)
{
    const int g = sb + lane_id();
    int cnt_g = 0;
    start_g = 0;
    shp = 0;
    if (g < H.nseg) {
        const int dx = g % H.sx - H.nx0;
        const int t = g / H.sx;
        const int dy = t % H.sy - H.nx1;
        const int dz = t / H.sy - H.nx2;
        const int a0 = H.ci0 + dx, a1 = H.ci1 + dy, a2 = H.ci2 + dz;
        // _is_cell_in_bounds (gpu_kernels.jl:20-25)
        const bool inb = (H.p0 || (a0 >= 1 && a0 <= H.n0)) && (H.p1 || (a1 >= 1 && a1 <= H.n1)) &&
                         (H.p2 || (a2 >= 1 && a2 <= H.n2));
        if (inb) {
            int32_t c0, c1, c2, s0, s1, s2;   // _get_neighbor_cell (:39-47)
            wrap_and_shift(a0, H.n0, H.p0, c0, s0);
            wrap_and_shift(a1, H.n1, H.p1, c1, s1);
            wrap_and_shift(a2, H.n2, H.p2, c2, s2);
            const int gc = H.cbase + (c0 - 1) + (c1 - 1) * H.n0 + (c2 - 1) * H.n0 * H.n1;
            start_g = A.cell_start[gc];
            cnt_g = A.cell_start[gc + 1] - start_g;
            shp = pack_wind(s0, s1, s2);
        }
    }
    const int incl = wave_incl_scan(cnt_g);
    excl = incl - cnt_g;
    return sgpr(__shfl(incl, kWave - 1, kWave));
}

struct Cand {
    double x, y, z;
    int32_t j;
    int32_t tx, ty, tz;         // s_loop - w_j  (S = t + w_i)
    bool zero_shift, valid;
};

__device__ __forceinline__ Cand load_candidate(const PairsArgs &A, int k, int total, int excl, int start_g, uint32_t shp)
{
    int m = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const int e2 = __shfl(excl, m + step, kWave);
        if (e2 <= k) m += step;
    }
    const int p = __shfl(start_g, m, kWave) + (k - __shfl(excl, m, kWave));
    const uint32_t shm = __shfl(shp, m, kWave);
    Cand c;
    c.valid = k < total;
    int32_t wjx = 0, wjy = 0, wjz = 0;
    if (c.valid) {
        const Rec r = A.rec[p];
        c.x = r.x; c.y = r.y; c.z = r.z;
        c.j = r.j;
        rec_winding(r, A.wbig, p, wjx, wjy, wjz);
    } else {
        c.x = c.y = c.z = 0.0;
        c.j = -1;
    }
    int32_t sx, sy, sz;
    unpack_wind(shm, sx, sy, sz);
    c.zero_shift = (sx == 0) & (sy == 0) & (sz == 0);
    c.tx = sx - wjx;
    c.ty = sy - wjy;
    c.tz = sz - wjz;
    return c;
}

// Per-wave count of one home cell from global memory: test every (home
// atom, chunk) and append hits to the slots (the brick path's fallback, and
// the path for neighbour reach beyond one cell).
template <bool WANT_R, bool WIDE>
__device__ __forceinline__ void count_cell(const PairsArgs &A, const SysDev &sd, int cell, int32_t *__restrict__ pair_count,
                                        const SlotArgs &O)
{
    HomeCell H;
    home_cell(A, sd, cell, H);
    if (H.hs >= H.he) return;
    const int lane = lane_id();
    const double *__restrict__ Cm = sd.C;
    const double rc2 = sd.rc2;
    for (int32_t hg = H.hs; hg < H.he; hg += kWave) {
        const int nh = sgpr(min(kWave, H.he - hg));
        const int32_t hj_lane = lane < nh ? A.rec[hg + lane].j : 0;
        int32_t cnt_lane = 0;   // lane a: running count of home atom hg + a
        for (int sb = 0; sb < H.nseg; sb += kWave) {
            int excl, start_g;
            uint32_t shp;
            const int total = segment_batch(A, H, sb, excl, start_g, shp);
            for (int k0 = 0; k0 < total; k0 += kChunksPerPass * kWave) {
                const int nch = sgpr(min(kChunksPerPass, (total - k0 + kWave - 1) / kWave));
                Cand c[kChunksPerPass];
                uint32_t tpk[kChunksPerPass];
                int32_t key[kChunksPerPass];
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) {
                    if (g < nch) {
                        c[g] = load_candidate(A, k0 + g * kWave + lane, total, excl, start_g, shp);
                        if (!c[g].valid) c[g].x = c[g].y = c[g].z = kFar;
                        tpk[g] = pack_wind(c[g].tx, c[g].ty, c[g].tz);
                        key[g] = (c[g].valid && c[g].zero_shift) ? c[g].j : -3;
                    } else {
                        c[g].x = c[g].y = c[g].z = kFar;
                        c[g].j = -1;
                        c[g].tx = c[g].ty = c[g].tz = 0;
                        tpk[g] = 0;
                        key[g] = -3;
                    }
                }
                double cx[kChunksPerPass], cy[kChunksPerPass], cz[kChunksPerPass];
                uint32_t wcache = 0xffffffffu;
                int32_t wix = 0, wiy = 0, wiz = 0;
                for (int a = 0; a < nh; ++a) {
                    const Rec h = A.rec[hg + a];   // wave-uniform: scalar loads
                    if (h.w != wcache || (h.w & kWindBig)) {
                        rec_winding(h, A.wbig, hg + a, wix, wiy, wiz);
#pragma unroll
                        for (int g = 0; g < kChunksPerPass; ++g)
                            cell_shift(Cm, c[g].tx + wix, c[g].ty + wiy, c[g].tz + wiz, cx[g], cy[g], cz[g]);
                        wcache = h.w;
                    }
                    const bool wzero = (wix | wiy | wiz) == 0;
                    int32_t ca = __builtin_amdgcn_readlane(cnt_lane, a);
                    const int64_t sbase = (H.abase + __builtin_amdgcn_readlane(hj_lane, a)) * (int64_t)O.cap;
#pragma unroll
                    for (int g = 0; g < kChunksPerPass; ++g) {
                        if (g < nch) {
                            const double Rx = (c[g].x - h.x) + cx[g];
                            const double Ry = (c[g].y - h.y) + cy[g];
                            const double Rz = (c[g].z - h.z) + cz[g];
                            const double r2 = (Rx * Rx + Ry * Ry) + Rz * Rz;
                            const uint64_t mask = __ballot((r2 < rc2) & (key[g] != h.j));
                            emit_hits<WANT_R, WIDE>(O, mask, sbase, ca, c[g].j, tpk[g], c[g].tx, c[g].ty, c[g].tz, wix,
                                                    wiy, wiz, wzero, Rx, Ry, Rz);
                        }
                    }
                    cnt_lane = lane == a ? ca : cnt_lane;
                }
            }
        }
        if (lane < nh) pair_count[H.abase + hj_lane] = cnt_lane | kRow8Bit;
    }
}

// Brick workgroup shapes: BT threads (8 waves for large systems, 4 for
// batches of small ones), 4 BT halo atoms and BT halo cells in LDS.
template <int BT>
struct BrickShape {
    static constexpr int threads = BT;
    static constexpr int waves = BT / kWave;
    static constexpr int cap = 4 * BT;    // halo atoms held in LDS
    static constexpr int halo = BT;       // halo cells, one per thread
    static constexpr int unroll = cap / BT;
};
constexpr uint32_t kPackZero = (512u) | (512u << 10) | (512u << 20);   // pack_wind(0, 0, 0)
constexpr int kBrickChunks = 3;   // candidate chunks a lane holds in the brick pass (192 candidates)

// Quotient of 0 <= x < 2^16 by 1 <= d < 2^16 through a reciprocal: m =
// floor((2^32 - 1) / d) + 1 (0 for d = 1), q = umulhi(x, m).  Exact: m
// exceeds 2^32 / d by less than one, so the product overshoots x / d by less
// than x / 2^32 < 1 / d, which never carries past the next integer.  A few
// instructions instead of the integer-division sequence.
struct SmallDiv {
    uint32_t d, m;
};

__device__ __forceinline__ SmallDiv small_div(int d)
{
    return {(uint32_t)d, d == 1 ? 0u : 0xffffffffu / (uint32_t)d + 1u};
}

__device__ __forceinline__ int small_quot(int x, SmallDiv D)
{
    return D.m ? (int)__umulhi((uint32_t)x, D.m) : x;
}

struct Brick {
    const SysDev *sd;
    int b0[3];     // 1-based coordinates of the first home cell
    int nh[3];     // home cells per axis (clipped at the grid edge)
    int hd[3];     // halo extent per axis (nh + 2)
    int nhalo, nhome;
    bool fast;     // one-cell reach and a small enough halo: the LDS path applies
    SmallDiv dnh0, dnh1, dhd0, dhd1;   // divisors nh[0], nh[1], hd[0], hd[1]
};

// Brick of workgroup w of a count launch: the bricks of the launch in
// system order, each system's in its class order, largest first
// (SysDev::cls_first).  Workgroups are dealt to the eight XCDs in turn, so
// every XCD takes its share of each class and of every region of a system
// (a slab with vacuum above it puts its empty bricks on every XCD, not on
// some of them).
__device__ __forceinline__ void brick_locate(const PairsArgs &A, const int32_t *__restrict__ sys_brick_base, int w,
                                             int halo_max, Brick &B)
{
    const int s = A.nsys == 1 ? 0 : find_sys(sys_brick_base, A.nsys, (int32_t)w);
    const SysDev &sd = A.sys[s];
    B.sd = &sd;
    int li = w - sys_brick_base[s], c = 0;
#pragma unroll
    for (int cc = 1; cc < 8; ++cc) c = li >= sd.cls_first[cc] ? cc : c;
    li -= sd.cls_first[c];
    // brick (ib0, ib1, ib2) of the class: the partial last brick along the
    // axes of its mask, row-major over the full bricks along the others
    const int m = sd.cls_mask[c];
    const int d0 = (m & 1) ? 1 : sd.bfull[0], d1 = (m & 2) ? 1 : sd.bfull[1];
    const int q0 = li / d0;
    const int ib[3] = {(m & 1) ? sd.bfull[0] : li - q0 * d0, (m & 2) ? sd.bfull[1] : q0 % d1,
                       (m & 4) ? sd.bfull[2] : q0 / d1};
    for (int k = 0; k < 3; ++k) {
        B.b0[k] = ib[k] * sd.brick[k] + 1;
        B.nh[k] = min(sd.brick[k], sd.n[k] - B.b0[k] + 1);
        B.hd[k] = B.nh[k] + 2;
    }
    B.nhalo = B.hd[0] * B.hd[1] * B.hd[2];
    B.nhome = B.nh[0] * B.nh[1] * B.nh[2];
    B.fast = sd.nxyz[0] == 1 && sd.nxyz[1] == 1 && sd.nxyz[2] == 1 && B.nhalo <= halo_max;
    B.dnh0 = small_div(B.nh[0]);
    B.dnh1 = small_div(B.nh[1]);
    B.dhd0 = small_div(B.hd[0]);
    B.dhd1 = small_div(B.hd[1]);
}

// wrap_and_shift (cell_list.jl:105-120) of an index within one cell of
// [1, n] on a periodic axis (a halo cell of a one-cell-reach brick): at most
// one wrap, no division.
__device__ __forceinline__ void wrap_near(int32_t a, int32_t n, int32_t &w, int32_t &s)
{
    s = a < 1 ? -1 : (a > n ? 1 : 0);
    w = a - s * n;
}

__device__ __forceinline__ int brick_cell_id(const Brick &B, int hc)
{
    const SysDev &sd = *B.sd;
    const int ix = hc % B.nh[0], iy = (hc / B.nh[0]) % B.nh[1], iz = hc / (B.nh[0] * B.nh[1]);
    return sd.cell_base + (B.b0[0] - 1 + ix) + (B.b0[1] - 1 + iy) * sd.n[0] + (B.b0[2] - 1 + iz) * sd.n[0] * sd.n[1];
}

// One halo atom in LDS, 32 bytes: a candidate is two 16-byte reads.
struct __attribute__((aligned(16))) LAtom {
    double x, y, z;
    int32_t j;                      // j, or j | S code << 26 when every j fits in 26 bits
    uint32_t t;                     // pack(s_halo - w_j); home cells have s_halo = 0
};

constexpr int kPassCands = kBrickChunks * kWave;   // candidates of one pass (three chunks)

template <int BT>
struct BrickLDS {
    using Sh = BrickShape<BT>;
    // entry cap is a far-away atom that lanes past the last candidate read,
    // so candidate loads need no per-lane branch
    LAtom a[Sh::cap + 1];
    int32_t hoff[Sh::halo + 1];     // LDS offset of each halo cell's atoms
    int32_t hstart[Sh::halo];       // first sorted index of each halo cell
    uint32_t hshift[Sh::halo];      // packed cell shift of each halo cell | kHomeCell
    // per wave, for its current home cell: the LDS offset less the candidate
    // prefix of each non-empty candidate row (in row order), and one byte per
    // candidate of the current pass marking where a non-empty row starts
    int32_t rowdl[Sh::waves][16];
    uint32_t rowflag[Sh::waves][kPassCands / 4];
    uint16_t owner[Sh::cap];        // halo cell of each LDS atom slot (while loading)
    int32_t next_cell;              // the next home cell a wave takes
    int32_t total, flag;
    int32_t row8;                   // some shift needs 8-byte slot entries
};

constexpr uint32_t kHomeCell = 1u << 30;   // halo-table mark of the brick's own cells

// Stage the brick and its halo in LDS.  Returns false (for every thread of
// the block) when the brick must take the global-memory path.
template <int BT>
__device__ bool brick_load(const PairsArgs &A, const Brick &B, BrickLDS<BT> &L)
{
    using Sh = BrickShape<BT>;
    const SysDev &sd = *B.sd;
    const int tid = threadIdx.x;
    const bool p[3] = {(sd.pbc & 1) != 0, (sd.pbc & 2) != 0, (sd.pbc & 4) != 0};
    // halo cell table: thread tid owns halo cell tid
    int cnt[1];
    int sum = 0;
#pragma unroll
    for (int u = 0; u < 1; ++u) {
        cnt[u] = 0;
        const int h = tid + u;
        if (h < B.nhalo) {
            const int q0 = small_quot(h, B.dhd0), uz = small_quot(q0, B.dhd1);
            const int ux = h - q0 * B.hd[0], uy = q0 - uz * B.hd[1];
            const int a[3] = {B.b0[0] - 1 + ux, B.b0[1] - 1 + uy, B.b0[2] - 1 + uz};
            bool inb = true;
            for (int k = 0; k < 3; ++k) inb = inb && (p[k] || (a[k] >= 1 && a[k] <= sd.n[k]));
            int st = 0;
            uint32_t sh = kPackZero;
            if (inb) {   // _is_cell_in_bounds / _get_neighbor_cell (gpu_kernels.jl:20-47)
                int32_t c[3], sft[3];
                // a halo cell lies within one cell of the grid: an open axis
                // holds it in bounds here, a periodic one wraps it at most once
                for (int k = 0; k < 3; ++k) wrap_near(a[k], sd.n[k], c[k], sft[k]);
                const int gc = sd.cell_base + (c[0] - 1) + (c[1] - 1) * sd.n[0] + (c[2] - 1) * sd.n[0] * sd.n[1];
                st = A.cell_start[gc];
                cnt[u] = A.cell_start[gc + 1] - st;
                sh = pack_wind(sft[0], sft[1], sft[2]);
            }
            const bool home = ux >= 1 && ux <= B.nh[0] && uy >= 1 && uy <= B.nh[1] && uz >= 1 && uz <= B.nh[2];
            L.hstart[h] = st;
            L.hshift[h] = sh | (home ? kHomeCell : 0u);
        }
        sum += cnt[u];
    }
    int64_t tot;
    int run = (int)block_excl_scan_n<BT>(sum, tot);
#pragma unroll
    for (int u = 0; u < 1; ++u) {
        const int h = tid + u;
        if (h < B.nhalo) {
            L.hoff[h] = run;
            // the cell's atom slots name their cell (a brick whose halo
            // overflows the LDS takes the global-memory path)
            const int e = min(run + cnt[u], Sh::cap);
            for (int q = run; q < e; ++q) L.owner[q] = (uint16_t)h;
        }
        run += cnt[u];
    }
    if (tid == 0) {
        L.hoff[B.nhalo] = (int32_t)tot;
        L.total = (int32_t)tot;
        L.next_cell = Sh::waves;
        L.flag = 0;
        L.row8 = sd.nat > kNarrowMaxAtoms ? 1 : 0;
        L.a[Sh::cap].x = L.a[Sh::cap].y = L.a[Sh::cap].z = kFar;
        L.a[Sh::cap].j = 0;
        L.a[Sh::cap].t = kPackZero;
    }
    __syncthreads();
    const int total = L.total;
    if (total > Sh::cap) return false;
    // halo atoms: all loads of a thread in flight together
    int pg[Sh::unroll];
    uint32_t hs[Sh::unroll];
#pragma unroll
    for (int u = 0; u < Sh::unroll; ++u) {
        const int idx = tid + u * BT;
        pg[u] = 0;
        hs[u] = kPackZero;
        if (idx < total) {
            const int h = L.owner[idx];
            pg[u] = L.hstart[h] + (idx - L.hoff[h]);
            hs[u] = L.hshift[h];
        }
    }
    Rec r[Sh::unroll];
#pragma unroll
    for (int u = 0; u < Sh::unroll; ++u)
        if (tid + u * BT < total) r[u] = A.rec[pg[u]];
#pragma unroll
    for (int u = 0; u < Sh::unroll; ++u) {
        const int idx = tid + u * BT;
        if (idx < total) {
            int32_t wx, wy, wz, sx, sy, sz;
            rec_winding(r[u], A.wbig, pg[u], wx, wy, wz);
            unpack_wind(hs[u] & ~kHomeCell, sx, sy, sz);
            const int32_t tx = sx - wx, ty = sy - wy, tz = sz - wz;
            const uint32_t t = pack_wind(tx, ty, tz);
            if (t & kWindBig) L.flag = 1;
            // 4-byte slot entries need every S = t + w_i within [-1, 1]^3: home
            // atoms inside the cell (w_i = 0) and halo atoms with |t_k| <= 1
            const bool ok = (hs[u] & kHomeCell) ? (tx | ty | tz) == 0
                                                : ((unsigned)(tx + 1) <= 2u && (unsigned)(ty + 1) <= 2u &&
                                                   (unsigned)(tz + 1) <= 2u);
            if (!ok) L.row8 = 1;
            // the 10-bit biased fields of t carry the S codes of a 4-byte entry in
            // their low 2 bits (meaningful only when the brick stays 4-byte)
            const uint32_t code = (t & 3u) | ((t >> 8) & 12u) | ((t >> 16) & 48u);
            LAtom la;
            la.x = r[u].x;
            la.y = r[u].y;
            la.z = r[u].z;
            la.j = sd.nat <= kNarrowMaxAtoms ? (int32_t)((uint32_t)r[u].j | code << 26) : r[u].j;
            la.t = t;
            L.a[idx] = la;
        }
    }
    __syncthreads();
    return L.flag == 0;
}

// Per home cell of a brick: its LDS atom range, first sorted index, the
// nine candidate rows (lane r < 9 holds row r) and the candidate index of
// the home cell's first atom (its own self candidate).
struct HomeRows {
    int h0, h1;      // LDS range of the home atoms (uniform)
    int rstart, rex; // per lane: LDS start and exclusive candidate prefix of row `lane`
    int len;         // per lane: candidates in row `lane`
    int total;       // candidates (uniform)
    int self0;       // candidate index of home atom 0 (uniform)
};

template <int BT>
__device__ __forceinline__ HomeRows brick_home(const Brick &B, const BrickLDS<BT> &L, int hc)
{
    HomeRows R;
    const int q0 = small_quot(hc, B.dnh0), iz = small_quot(q0, B.dnh1);
    const int ix = hc - q0 * B.nh[0], iy = q0 - iz * B.nh[1];
    const int hh = (ix + 1) + B.hd[0] * ((iy + 1) + B.hd[1] * (iz + 1));
    R.h0 = sgpr(L.hoff[hh]);
    R.h1 = sgpr(L.hoff[hh + 1]);
    const int lane = lane_id();
    int len = 0, st = 0;
    if (lane < 9) {   // row (dz, dy) = (lane / 3 - 1, lane % 3 - 1); dx = -1..1 contiguous
        const int dz = lane / 3 - 1, dy = lane % 3 - 1;
        const int hr = ix + B.hd[0] * ((iy + 1 + dy) + B.hd[1] * (iz + 1 + dz));
        st = L.hoff[hr];
        len = L.hoff[hr + 3] - st;
    }
    // inclusive scan over lanes 0..8 (within the first 16-lane row)
    const int incl = row16_incl_scan(len);
    R.rex = incl - len;
    R.rstart = st;
    R.len = len;
    R.total = __builtin_amdgcn_readlane(incl, 8);
    // row 4 is (dz, dy) = (0, 0); the home cell is its middle cell
    R.self0 = __builtin_amdgcn_readlane(R.rex, 4) + (R.h0 - __builtin_amdgcn_readlane(R.rstart, 4));
    return R;
}

// LDS index of candidate k of a home cell (k < total): k plus the offset
// (LDS start - candidate prefix) of the last non-empty row whose first
// candidate is at or before k.  Per home cell the wave publishes the offsets
// of its non-empty rows in row order (rowdl[q], q = the row's rank among the
// non-empty rows).  Per pass of kPassCands candidates it marks with a byte
// each candidate where a non-empty row other than the first starts and turns
// the marks into one 64-bit word per chunk (a ballot); a candidate's rank q
// is then the number of marks at or below it: those of the chunks before (a
// popcount) and those of its own chunk up to its lane (mbcnt).  The row
// maps are the wave's own LDS: written and read by the wave alone, in
// program order.
struct RowMap {
    uint64_t sh[kBrickChunks];   // marks of each chunk, shifted right by one (uniform)
    int base[kBrickChunks];      // marks before each chunk, plus the chunk's own first (uniform)
};

// Publish the row offsets of a home cell; returns this lane's rank among the
// non-empty rows (meaningful on the lanes of non-empty rows, `ne`).
template <int BT>
__device__ __forceinline__ int row_publish(const HomeRows &R, BrickLDS<BT> &L, int wv, bool &ne)
{
    const int lane = lane_id();
    ne = lane < 9 && R.len > 0;
    const int q = (int)mbcnt64(__builtin_amdgcn_ballot_w64(ne));
    if (ne) L.rowdl[wv][q] = R.rstart - R.rex;
    return q;
}

template <int BT>
__device__ __forceinline__ RowMap row_map(const HomeRows &R, BrickLDS<BT> &L, int wv, int k0, int q, bool ne)
{
    const int lane = lane_id();
    uint8_t *fl = reinterpret_cast<uint8_t *>(L.rowflag[wv]);
    if (lane < kPassCands / 4) L.rowflag[wv][lane] = 0u;
    const int p = R.rex - k0;   // pass-relative first candidate of row `lane`
    if (ne && q > 0 && p > 0 && p < kPassCands) fl[p] = 1;
    int acc = (int)__popcll(__builtin_amdgcn_ballot_w64(ne && q > 0 && p <= 0));
    RowMap M;
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g) {
        const uint64_t b = __builtin_amdgcn_ballot_w64(fl[g * kWave + lane] != 0);
        M.sh[g] = b >> 1;
        M.base[g] = acc + (int)(b & 1u);
        acc += (int)__popcll(b);
    }
    return M;
}

template <int BT>
__device__ __forceinline__ int row_index(const RowMap &M, const BrickLDS<BT> &L, int wv, int g, int k)
{
    const int q = M.base[g] + (int)mbcnt64(M.sh[g]);
    return k + L.rowdl[wv][q];
}

// First slot of atom e's row: both factors are below 2^31, so one 32 x 32 -> 64
// bit product.
__device__ __forceinline__ int64_t slot_row(int64_t e, int32_t cap)
{
    return (int64_t)((uint64_t)(uint32_t)e * (uint32_t)cap);
}

// Append one ballot's hits to a home atom's slot row.  `row` and `rrow` are
// the (wave-uniform) slot row and its R row; ca the atom's running count;
// m the ballot of the hit lanes (the self pair already cleared), which also
// serves as the store mask; ent = j | pack(s - w_j) << 32 is the candidate's
// entry when w_i = 0.  Entry r of the row is written only while r < cap; the
// per-lane bound is tested only when the row can overflow in this ballot.
//
// 4-byte rows without R go through a buffer resource spanning the row's cap
// entries: every lane stores, lanes without a hit at an offset past the
// resource's end, and the hardware's range check drops those stores and any
// entry at r >= cap.  No exec-mask or bound branches.
//
// The offset of a lane without a hit: beyond any row resource (cap * 4 <
// 2^30); the bit pattern of 4.0f, an inline constant, so the select that
// applies it needs no literal.
constexpr uint32_t kDropOffset = 0x40800000u;
static_assert(__builtin_bit_cast(uint32_t, 4.0f) == kDropOffset, "the select below writes the offset as 4.0");
constexpr int32_t kRawBufferWord3 = 0x00020000; // raw 32-bit buffer access (GFX9 family)

template <bool NARROW, bool WZERO, bool WANT_R, bool WIDE, bool DENSE = false>
__device__ __forceinline__ void emit_row(void *__restrict__ row, double *__restrict__ rrow, uint32_t cap,
                                         Status *__restrict__ status, uint64_t m, int32_t &ca, uint64_t ent,
                                         int32_t wix, int32_t wiy, int32_t wiz, double Rx, double Ry, double Rz)
{
    const uint32_t n = (uint32_t)__popcll(m);
    if (NARROW && !WANT_R) {
        // every lane stores, each lane without a hit at an offset past the
        // resource's end (a select on the ballot, not an exec-mask region);
        // an empty ballot skips it
        if (!DENSE && n == 0) return;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(row, 0, (int)(cap * 4u), kRawBufferWord3);
        const uint32_t rk = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        uint32_t off;
        // a plain select on the 64-bit ballot (the compiler would otherwise
        // open an exec-mask region around the offset computation)
        __asm__("v_cndmask_b32_e64 %0, 4.0, %1, %2" : "=v"(off) : "v"((rk + (uint32_t)ca) << 2), "s"(m));
        // j | S code << 26 (w_i = 0 in a 4-byte brick)
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)ent, rs, (int)off, 0, 0);
        ca += (int32_t)n;
        return;
    }
    if (n != 0) {
        const uint32_t rk = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        uint64_t st = m;
        if ((uint32_t)ca + n > cap)   // lanes whose entry would fall past the row's end store nothing
            st &= __builtin_amdgcn_ballot_w64((int32_t)rk < (int32_t)cap - ca);
        const uint32_t r = rk + (uint32_t)ca;
        if (__builtin_amdgcn_inverse_ballot_w64(st)) {
            if (NARROW) {
                // j | S code << 26 (w_i = 0 in a 4-byte brick)
                *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(row) + (r << 2)) = (uint32_t)ent;
            } else if (WIDE) {
                int32_t tx, ty, tz;
                unpack_wind((uint32_t)(ent >> 32), tx, ty, tz);
                reinterpret_cast<int4 *>(row)[r] = make_int4((int32_t)(uint32_t)ent, tx + wix, ty + wiy, tz + wiz);
            } else if (WZERO) {
                // S = s_loop + w_i - w_j = s_loop - w_j  (gpu_kernels.jl:88)
                reinterpret_cast<uint64_t *>(row)[r] = ent;
            } else {
                int32_t tx, ty, tz;
                unpack_wind((uint32_t)(ent >> 32), tx, ty, tz);
                const uint32_t spk = pack_wind(tx + wix, ty + wiy, tz + wiz);
                if (spk & kWindBig) status->wide = 1;
                reinterpret_cast<uint64_t *>(row)[r] = (ent & 0xffffffffull) | ((uint64_t)spk << 32);
            }
            if (WANT_R) {
                double *Rd = rrow + 3 * r;
                Rd[0] = Rx;
                Rd[1] = Ry;
                Rd[2] = Rz;
            }
        }
        ca += (int32_t)n;
    }
}

// One (home atom, chunk) test of the brick pass.  SHIFT: add the candidate's
// cell shift C' S (left out when every lane's shift is zero and R is not
// written: r^2 is then the same).  `self` clears the home atom's own
// (j == i, s == 0) candidate from the ballot when it lies in this chunk.
template <bool NARROW, bool SHIFT, bool WZERO, bool WANT_R, bool WIDE>
__device__ __forceinline__ void test_chunk(double cx, double cy, double cz, double sx, double sy, double sz, double hx,
                                           double hy, double hz, double rc2, uint64_t self, void *__restrict__ row,
                                           double *__restrict__ rrow, uint32_t cap, Status *__restrict__ status,
                                           int32_t &ca, uint64_t ent, int32_t wix, int32_t wiy, int32_t wiz)
{
    // R = xj - xi + C' * S ; r^2 = dot(R, R)  (gpu_kernels.jl:89-92)
    double Rx = cx - hx, Ry = cy - hy, Rz = cz - hz;
    if (SHIFT) {
        Rx = Rx + sx;
        Ry = Ry + sy;
        Rz = Rz + sz;
    }
    const double r2 = (Rx * Rx + Ry * Ry) + Rz * Rz;
    const bool in = r2 < rc2;
    const uint64_t m = __builtin_amdgcn_ballot_w64(in) & self;
    emit_row<NARROW, WZERO, WANT_R, WIDE>(row, rrow, cap, status, m, ca, ent, wix, wiy, wiz, Rx, Ry, Rz);
}

// All chunks of one home atom.
template <bool NARROW, bool WZERO, bool WANT_R, bool WIDE>
__device__ __forceinline__ void test_home_atom(int nch, const double *cx, const double *cy, const double *cz,
                                               const double *sx, const double *sy, const double *sz, const uint64_t *csh,
                                               const uint64_t *ent, double hx, double hy, double hz, double rc2,
                                               int ks, void *__restrict__ row, double *__restrict__ rrow, uint32_t cap,
                                               Status *__restrict__ status, int32_t &ca, int32_t wix, int32_t wiy,
                                               int32_t wiz)
{
    // _is_self_interaction: j == i with no cell shift is the home atom's own
    // slot in the home cell, candidate ks of this pass: chunk ks >> 6 (none
    // when ks is out of the pass), lane ks & 63
    const int gs = ks >> 6;
    const uint64_t nself = ~(1ull << (ks & (kWave - 1)));
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g) {
        if (g < nch) {
            const uint64_t self = g == gs ? nself : ~0ull;
            if (WANT_R || csh[g] != 0)
                test_chunk<NARROW, true, WZERO, WANT_R, WIDE>(cx[g], cy[g], cz[g], sx[g], sy[g], sz[g], hx, hy, hz, rc2,
                                                              self, row, rrow, cap, status, ca, ent[g], wix, wiy, wiz);
            else
                test_chunk<NARROW, false, WZERO, WANT_R, WIDE>(cx[g], cy[g], cz[g], sx[g], sy[g], sz[g], hx, hy, hz, rc2,
                                                               self, row, rrow, cap, status, ca, ent[g], wix, wiy, wiz);
        }
    }
}

// Two consecutive home atoms of a 4-byte brick (w_i = 0 for both, so the
// same shifts) against the same chunks: both distance tests of a chunk come
// before both emissions, two independent chains the scheduler interleaves.
// Atom 1's self candidate follows atom 0's.
template <bool WANT_R, bool WIDE, bool DENSE>
__device__ __forceinline__ void test_home_pair(int nch, const double *cx, const double *cy, const double *cz,
                                               const double *sx, const double *sy, const double *sz, const uint64_t *csh,
                                               const uint64_t *ent, double h0x, double h0y, double h0z, double h1x,
                                               double h1y, double h1z, double rc2, int ks0, void *__restrict__ row0,
                                               void *__restrict__ row1, double *__restrict__ rrow0,
                                               double *__restrict__ rrow1, uint32_t cap, Status *__restrict__ status,
                                               int32_t &ca0, int32_t &ca1)
{
    // self candidates (as in test_home_atom): ks0 for atom 0, ks0 + 1 for atom 1
    const int gs0 = ks0 >> 6, gs1 = (ks0 + 1) >> 6;
    const uint64_t nself0 = ~(1ull << (ks0 & (kWave - 1))), nself1 = ~(1ull << ((ks0 + 1) & (kWave - 1)));
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g) {
        if (g < nch) {
            // R = xj - xi + C' * S ; r^2 = dot(R, R)  (gpu_kernels.jl:89-92)
            double R0x = cx[g] - h0x, R0y = cy[g] - h0y, R0z = cz[g] - h0z;
            double R1x = cx[g] - h1x, R1y = cy[g] - h1y, R1z = cz[g] - h1z;
            if (WANT_R || csh[g] != 0) {
                R0x = R0x + sx[g];
                R0y = R0y + sy[g];
                R0z = R0z + sz[g];
                R1x = R1x + sx[g];
                R1y = R1y + sy[g];
                R1z = R1z + sz[g];
            }
            const bool in0 = (R0x * R0x + R0y * R0y) + R0z * R0z < rc2;
            const bool in1 = (R1x * R1x + R1y * R1y) + R1z * R1z < rc2;
            const uint64_t m0 = __builtin_amdgcn_ballot_w64(in0) & (g == gs0 ? nself0 : ~0ull);
            const uint64_t m1 = __builtin_amdgcn_ballot_w64(in1) & (g == gs1 ? nself1 : ~0ull);
            emit_row<true, true, WANT_R, WIDE, DENSE>(row0, rrow0, cap, status, m0, ca0, ent[g], 0, 0, 0, R0x, R0y, R0z);
            emit_row<true, true, WANT_R, WIDE, DENSE>(row1, rrow1, cap, status, m1, ca1, ent[g], 0, 0, 0, R1x, R1y, R1z);
        }
    }
}

// test_home_pair for three full chunks with no shift and 4-byte rows without
// R: the six distance tests of the pair, then its six emissions (every lane
// stores, lanes without a hit at the dropped offset), one branch-free block.
// G: the chunk that holds both self candidates (known for a whole group of
// home atoms: one mask per atom), kBrickChunks when neither lies in the
// pass, or -1 when the chunk is found per atom.
template <int G>
__device__ __forceinline__ void test_pair_plain(const double *cx, const double *cy, const double *cz,
                                                const uint64_t *ent, double h0x, double h0y, double h0z, double h1x,
                                                double h1y, double h1z, double rc2, int ks0, void *__restrict__ row0,
                                                void *__restrict__ row1, uint32_t cap, int32_t &ca0, int32_t &ca1)
{
    const int gs0 = G < 0 ? ks0 >> 6 : G, gs1 = G < 0 ? (ks0 + 1) >> 6 : G;
    const uint64_t nself0 = ~(1ull << (ks0 & (kWave - 1))), nself1 = ~(1ull << ((ks0 + 1) & (kWave - 1)));
    uint64_t m0[kBrickChunks], m1[kBrickChunks];
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g) {
        // R = xj - xi (shift zero) ; r^2 = dot(R, R)  (gpu_kernels.jl:89-92)
        const double R0x = cx[g] - h0x, R0y = cy[g] - h0y, R0z = cz[g] - h0z;
        const double R1x = cx[g] - h1x, R1y = cy[g] - h1y, R1z = cz[g] - h1z;
        const bool in0 = (R0x * R0x + R0y * R0y) + R0z * R0z < rc2;
        const bool in1 = (R1x * R1x + R1y * R1y) + R1z * R1z < rc2;
        m0[g] = __builtin_amdgcn_ballot_w64(in0) & (g == gs0 ? nself0 : ~0ull);
        m1[g] = __builtin_amdgcn_ballot_w64(in1) & (g == gs1 ? nself1 : ~0ull);
    }
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g) {
        emit_row<true, true, false, false, true>(row0, nullptr, cap, nullptr, m0[g], ca0, ent[g], 0, 0, 0, 0.0, 0.0, 0.0);
        emit_row<true, true, false, false, true>(row1, nullptr, cap, nullptr, m1[g], ca1, ent[g], 0, 0, 0, 0.0, 0.0, 0.0);
    }
}

// One home atom against three full chunks with no shift, 4-byte rows
// without R (the odd atom of a pass through test_pair_plain).
__device__ __forceinline__ void test_atom_plain(const double *cx, const double *cy, const double *cz,
                                                const uint64_t *ent, double hx, double hy, double hz, double rc2,
                                                int ks, void *__restrict__ row, uint32_t cap, int32_t &ca)
{
    const int gs = ks >> 6;
    const uint64_t nself = ~(1ull << (ks & (kWave - 1)));
    uint64_t m[kBrickChunks];
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g) {
        // R = xj - xi (shift zero) ; r^2 = dot(R, R)  (gpu_kernels.jl:89-92)
        const double Rx = cx[g] - hx, Ry = cy[g] - hy, Rz = cz[g] - hz;
        m[g] = __builtin_amdgcn_ballot_w64((Rx * Rx + Ry * Ry) + Rz * Rz < rc2) & (g == gs ? nself : ~0ull);
    }
#pragma unroll
    for (int g = 0; g < kBrickChunks; ++g)
        emit_row<true, true, false, false, true>(row, nullptr, cap, nullptr, m[g], ca, ent[g], 0, 0, 0, 0.0, 0.0, 0.0);
}

// The home atoms of a pass through test_pair_plain<G>, two at a time from
// home atom a (left at the first atom not done).  FRESH: the group's first
// pass, every running count still zero.
template <int G, bool FRESH = false>
__device__ __forceinline__ void plain_pairs(const double *cx, const double *cy, const double *cz, const uint64_t *ent,
                                            const LAtom *home, int nh, int32_t hj_l, int kpass, double rc2,
                                            int64_t atom_base, void *__restrict__ slots, int32_t cap, int32_t &cnt_l,
                                            int &a)
{
    for (; a + 1 < nh; a += 2) {
        const LAtom h0 = home[a], h1 = home[a + 1];
        int32_t ca0 = FRESH ? 0 : __builtin_amdgcn_readlane(cnt_l, a);
        int32_t ca1 = FRESH ? 0 : __builtin_amdgcn_readlane(cnt_l, a + 1);
        const int64_t srow0 = slot_row(atom_base + __builtin_amdgcn_readlane(hj_l, a), cap);
        const int64_t srow1 = slot_row(atom_base + __builtin_amdgcn_readlane(hj_l, a + 1), cap);
        test_pair_plain<G>(cx, cy, cz, ent, h0.x, h0.y, h0.z, h1.x, h1.y, h1.z, rc2, kpass + a,
                           reinterpret_cast<uint64_t *>(slots) + srow0, reinterpret_cast<uint64_t *>(slots) + srow1,
                           (uint32_t)cap, ca0, ca1);
        cnt_l = write_lane(ca0, a, cnt_l);
        cnt_l = write_lane(ca1, a + 1, cnt_l);
    }
}

// One pass of a home cell's atoms over loaded candidate chunks: every home
// atom's distance tests against the nch chunks (cx, cy, cz, ent: j |
// pack(s - w_j) << 32, or j | S code << 26 in a 4-byte brick), its hits
// appended to its slot row at its running count (lane a of cnt_l for home
// atom a).  home: the nh home atoms in LDS; ht_l, hj_l: lane a holds
// pack(-w_i) and j of home atom a; kpass: the self candidate of home atom 0
// in this pass (home atom a's is kpass + a); Cm: the cell matrix; fresh:
// the first pass of the home atoms (every running count zero).
template <bool WANT_R, bool WIDE, bool DENSE>
__device__ __forceinline__ void pass_home_atoms(bool narrow, bool fresh, int nch, const double *cx, const double *cy,
                                                const double *cz, const uint64_t *ent, const LAtom *home, int nh,
                                                uint32_t ht_l, int32_t hj_l, int kpass, const double *Cm, double rc2,
                                                int64_t atom_base, void *__restrict__ slots, double *__restrict__ slotsR,
                                                int32_t cap, Status *__restrict__ status, int32_t &cnt_l)
{
    double sx[kBrickChunks], sy[kBrickChunks], sz[kBrickChunks];
    uint32_t wcur = 0xffffffffu;   // home winding the shifts were computed for
    int32_t wix = 0, wiy = 0, wiz = 0;
    // wave-uniform, per chunk: the lanes whose shift is nonzero (the
    // shifts sx, sy, sz are set only where this is nonzero or R is written)
    uint64_t csh[kBrickChunks];
    int a = 0;
    if (narrow && nh >= 2) {
        // every home atom of a 4-byte brick has w_i = 0: the shifts of
        // the candidates are their own cell shifts
#pragma unroll
        for (int g = 0; g < kBrickChunks; ++g)
            csh[g] = g < nch ? __builtin_amdgcn_ballot_w64((uint32_t)(ent[g] >> 32) != kPackZero) : 0;
        if (DENSE && !WANT_R && !WIDE && nch == kBrickChunks && (csh[0] | csh[1] | csh[2]) == 0) {
            // the self candidates kpass .. kpass + nh - 1 of the group: in one
            // chunk (nearly always), in none, or across two
            const int g0 = kpass >> 6, g1 = (kpass + nh - 1) >> 6;
            if (g1 < 0 || g0 >= kBrickChunks)
                plain_pairs<kBrickChunks>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else if (g0 == g1 && g0 == 0 && fresh)
                plain_pairs<0, true>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else if (g0 == g1 && g0 == 1 && fresh)
                plain_pairs<1, true>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else if (g0 == g1 && g0 == 2 && fresh)
                plain_pairs<2, true>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else if (g0 == g1 && g0 == 0)
                plain_pairs<0>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else if (g0 == g1 && g0 == 1)
                plain_pairs<1>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else if (g0 == g1 && g0 == 2)
                plain_pairs<2>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            else
                plain_pairs<-1>(cx, cy, cz, ent, home, nh, hj_l, kpass, rc2, atom_base, slots, cap, cnt_l, a);
            if (a < nh) {   // the odd home atom
                const LAtom hm = home[a];
                int32_t ca = __builtin_amdgcn_readlane(cnt_l, a);
                const int64_t srow = slot_row(atom_base + __builtin_amdgcn_readlane(hj_l, a), cap);
                test_atom_plain(cx, cy, cz, ent, hm.x, hm.y, hm.z, rc2, kpass + a,
                                reinterpret_cast<uint64_t *>(slots) + srow, (uint32_t)cap, ca);
                cnt_l = write_lane(ca, a, cnt_l);
            }
            return;   // the shifts are never needed
        }
#pragma unroll
        for (int g = 0; g < kBrickChunks; ++g) {
            if (g < nch && (WANT_R || csh[g] != 0)) {
                int32_t cx_, cy_, cz_;
                unpack_wind((uint32_t)(ent[g] >> 32), cx_, cy_, cz_);
                cell_shift(Cm, cx_, cy_, cz_, sx[g], sy[g], sz[g]);
            }
        }
        wcur = kPackZero;
        for (; a + 1 < nh; a += 2) {
            // uniform LDS reads: every lane receives the home atom's coordinates
            const LAtom h0 = home[a], h1 = home[a + 1];
            const double h0x = h0.x, h0y = h0.y, h0z = h0.z;
            const double h1x = h1.x, h1y = h1.y, h1z = h1.z;
            int32_t ca0 = __builtin_amdgcn_readlane(cnt_l, a), ca1 = __builtin_amdgcn_readlane(cnt_l, a + 1);
            const int64_t srow0 = slot_row(atom_base + __builtin_amdgcn_readlane(hj_l, a), cap);
            const int64_t srow1 = slot_row(atom_base + __builtin_amdgcn_readlane(hj_l, a + 1), cap);
            void *row0 = WIDE ? (void *)(reinterpret_cast<int4 *>(slots) + srow0)
                              : (void *)(reinterpret_cast<uint64_t *>(slots) + srow0);
            void *row1 = WIDE ? (void *)(reinterpret_cast<int4 *>(slots) + srow1)
                              : (void *)(reinterpret_cast<uint64_t *>(slots) + srow1);
            double *rrow0 = WANT_R ? slotsR + 3 * srow0 : nullptr;
            double *rrow1 = WANT_R ? slotsR + 3 * srow1 : nullptr;
            test_home_pair<WANT_R, WIDE, DENSE>(nch, cx, cy, cz, sx, sy, sz, csh, ent, h0x, h0y, h0z, h1x, h1y,
                                         h1z, rc2, kpass + a, row0, row1, rrow0, rrow1,
                                         (uint32_t)cap, status, ca0, ca1);
            cnt_l = write_lane(ca0, a, cnt_l);
            cnt_l = write_lane(ca1, a + 1, cnt_l);
        }
    }
    for (; a < nh; ++a) {
        const LAtom hm = home[a];
        const double hx = hm.x, hy = hm.y, hz = hm.z;
        const uint32_t ht = (uint32_t)__builtin_amdgcn_readlane((int)ht_l, a);
        if (ht != wcur) {   // first home atom, or a new home winding
            int32_t tx, ty, tz;
            unpack_wind(ht, tx, ty, tz);
            wix = -tx; wiy = -ty; wiz = -tz;
#pragma unroll
            for (int g = 0; g < kBrickChunks; ++g) {
                csh[g] = 0;
                if (g < nch) {
                    int32_t cx_, cy_, cz_;
                    unpack_wind((uint32_t)(ent[g] >> 32), cx_, cy_, cz_);
                    cx_ += wix; cy_ += wiy; cz_ += wiz;
                    csh[g] = __builtin_amdgcn_ballot_w64((cx_ | cy_ | cz_) != 0);
                    if (WANT_R || csh[g] != 0) cell_shift(Cm, cx_, cy_, cz_, sx[g], sy[g], sz[g]);
                }
            }
            wcur = ht;
        }
        const bool wzero = ht == kPackZero;   // the home atom lies inside the cell
        int32_t ca = __builtin_amdgcn_readlane(cnt_l, a);
        // slot rows are indexed by original atom index, so the fill reads them in order
        const int64_t srow = slot_row(atom_base + __builtin_amdgcn_readlane(hj_l, a), cap);
        void *row = WIDE ? (void *)(reinterpret_cast<int4 *>(slots) + srow)
                         : (void *)(reinterpret_cast<uint64_t *>(slots) + srow);
        double *rrow = WANT_R ? slotsR + 3 * srow : nullptr;
        const int ks = kpass + a;   // the self candidate's place in this pass
        if (narrow)   // every home atom of a 4-byte brick has w_i = 0
            test_home_atom<true, true, WANT_R, WIDE>(nch, cx, cy, cz, sx, sy, sz, csh, ent, hx, hy, hz, rc2,
                                                     ks, row, rrow, (uint32_t)cap, status, ca, wix, wiy,
                                                     wiz);
        else if (wzero)
            test_home_atom<false, true, WANT_R, WIDE>(nch, cx, cy, cz, sx, sy, sz, csh, ent, hx, hy, hz, rc2,
                                                      ks, row, rrow, (uint32_t)cap, status, ca, wix, wiy,
                                                      wiz);
        else
            test_home_atom<false, false, WANT_R, WIDE>(nch, cx, cy, cz, sx, sy, sz, csh, ent, hx, hy, hz,
                                                       rc2, ks, row, rrow, (uint32_t)cap, status, ca, wix,
                                                       wiy, wiz);
        cnt_l = write_lane(ca, a, cnt_l);
    }
}

// Brick pass of the count: home atoms stream through the wave as scalars
// (read out of one-per-lane registers), candidates sit one per lane in
// registers, chunk by chunk; each (home atom, chunk) is one FP64 distance
// test and one ballot.  Bricks that cannot take this path (reach beyond one
// cell, halo over the LDS, windings beyond the packed range) are counted
// from global memory, one home cell per wave.
template <bool WANT_R, bool WIDE, int BT>
__global__ __launch_bounds__(BT) void count_kernel(
    const SysDev *__restrict__ sys, const int32_t *__restrict__ sys_cell_base, const int32_t *__restrict__ sys_brick_base,
    int nsys, int32_t M, int64_t nat, const int32_t *__restrict__ cell_start, const Rec *__restrict__ rec,
    const int32_t *__restrict__ wbig, int32_t *__restrict__ pair_count, void *__restrict__ slots,
    double *__restrict__ slotsR, int32_t cap, Status *__restrict__ status)
{
    __shared__ BrickLDS<BT> L;
    __shared__ double Csh[9];
    const PairsArgs A{sys, sys_cell_base, nsys, M, nat, cell_start, rec, wbig};
    Brick B;
    brick_locate(A, sys_brick_base, (int)blockIdx.x, BrickShape<BT>::halo, B);
    if (threadIdx.x < 9) Csh[threadIdx.x] = B.sd->C[threadIdx.x];   // visible after brick_load's barrier
    const bool lds = B.fast && brick_load(A, B, L);
    if (!lds) {
        const SlotArgs O{slots, slotsR, cap, status};
        const int wv0 = sgpr((int)(threadIdx.x >> 6));
        for (int hc = wv0; hc < B.nhome; hc += BrickShape<BT>::waves)
            count_cell<WANT_R, WIDE>(A, *B.sd, brick_cell_id(B, hc), pair_count, O);
        return;
    }
    const SysDev &sd = *B.sd;
    const double rc2 = sd.rc2;
    const int lane = lane_id();
    const int wv = sgpr((int)(threadIdx.x >> 6));
    const bool narrow = !WIDE && sgpr(L.row8) == 0;   // the brick writes 4-byte slot entries
    // L.j holds j | S code << 26 whenever every j fits in 26 bits
    const uint32_t hjmask = sd.nat <= kNarrowMaxAtoms ? (uint32_t)(kNarrowMaxAtoms - 1) : 0xffffffffu;
    const uint32_t jmask = narrow ? 0xffffffffu : hjmask;   // entry low word: j | code << 26, or j
    // home cells: one per wave, then the next one not taken (a wave that
    // ends its cell early takes the next at once)
    for (int hc = wv; hc < B.nhome; hc = take_next(&L.next_cell)) {
        const HomeRows R = brick_home(B, L, hc);
        bool rne;
        const int rq = row_publish(R, L, wv, rne);
        const int nha = R.h1 - R.h0;
        for (int hg = 0; hg < nha; hg += kWave) {
            const int nh = sgpr(min(kWave, nha - hg));
            const int hb = R.h0 + hg;        // LDS index of home atom hg (uniform)
            const int hl = hb + (lane < nh ? lane : 0);
            const uint32_t ht_l = L.a[hl].t;   // pack(-w_i): a home cell's own shift is zero
            const int32_t hj_l = (int32_t)((uint32_t)L.a[hl].j & hjmask);
            const int kself = R.self0 + hg;  // candidate index of home atom hg + a is kself + a
            int32_t cnt_l = 0;                // lane a: running count of home atom hg + a
            for (int k0 = 0; k0 < R.total; k0 += kBrickChunks * kWave) {
                const int nch = sgpr(min(kBrickChunks, (R.total - k0 + kWave - 1) / kWave));
                const RowMap Mp = row_map(R, L, wv, k0, rq, rne);
                double cx[kBrickChunks], cy[kBrickChunks], cz[kBrickChunks];
                uint64_t ent[kBrickChunks];   // j | pack(s - w_j) << 32
#pragma unroll
                for (int g = 0; g < kBrickChunks; ++g) {
                    // every chunk of the pass is loaded (no branch): lanes past the
                    // last candidate read the far sentinel, and chunks past nch
                    // are never tested
                    const int k = k0 + g * kWave + lane;
                    const int lr = row_index(Mp, L, wv, g, k);
                    const int li = k < R.total ? lr : BrickShape<BT>::cap;   // else the far sentinel
                    const LAtom c = L.a[li];
                    cx[g] = c.x;
                    cy[g] = c.y;
                    cz[g] = c.z;
                    ent[g] = (uint64_t)((uint32_t)c.j & jmask) | ((uint64_t)c.t << 32);
                }
                pass_home_atoms<WANT_R, WIDE, true>(narrow, k0 == 0, nch, cx, cy, cz, ent, L.a + hb, nh, ht_l, hj_l,
                                                    kself - k0, Csh, rc2, sd.atom_base, slots, slotsR, cap, status,
                                                    cnt_l);
            }
            if (lane < nh) pair_count[sd.atom_base + hj_l] = cnt_l | (narrow ? 0 : kRow8Bit);
        }
    }
}

// ---------------------------------------------------------------------------
// Fused build of small systems
// ---------------------------------------------------------------------------
//
// One workgroup of kSmallThreads threads per work item: a unit of
// consecutive systems (at most kUnitAtoms atoms, kUnitCells cells and
// kUnitSystems systems) or one of its parts, all of it in LDS: the cell and
// winding of every atom, a stable counting sort by cell (ties in ascending
// original index, Julia's stable sortperm), the cell-list outputs (written
// by part 0), and the neighbour count and slot row of every home atom of
// the part's count work items (items part, part + parts, ..., each a group
// of home atoms of one cell), through the brick pass's tests and emission;
// the waves take those items from a workgroup counter.  A home cell's
// candidates are its (2 nxyz + 1)^3 neighbour cells in the reference order
// (dz, dy, dx outer to inner, gpu_kernels.jl:69-71), each a segment of the
// sorted atoms with its cell shift, taken 64 segments at a time.

// Count work items: a group of at most sys_group home atoms of one cell.
// The group size of a system aims at kItemTests distance tests (chunks of 64
// candidates) per item, at least kMinGroup atoms: a cell whose candidates run
// to thousands (a small periodic cell, its own 26 images) is shared by
// several waves instead of holding one wave for all of it.
constexpr int kItemTests = 512;
constexpr int kMinGroup = 4;
constexpr int kUnitItems = kUnitAtoms / kMinGroup + kUnitCells;   // items of one unit at most

struct SmallSort {                 // LDS of the sort
    int16_t key[kUnitAtoms];       // unit cell of each atom
    int16_t grp[kUnitAtoms];       // atoms grouped by cell, in arrival order
    int32_t fill[kUnitCells];      // per cell: atoms, then arrivals
    uint32_t wk[kUnitAtoms];       // packed winding of each atom
    uint8_t sy[kUnitAtoms];        // unit system of each atom
};
struct SmallRows {                 // LDS of one wave in the count
    int32_t dl[kWave];             // LDS offset less candidate prefix of each non-empty segment
    uint32_t sh[kWave];            // packed cell shift of each non-empty segment
    uint32_t flag[kPassCands / 4]; // one byte per candidate of a pass: a segment starts there
};
struct SmallLDS {
    LAtom a[kUnitAtoms + 1];       // sorted atoms {x, y, z, system-local j, packed winding}; far sentinel last
    int32_t cs[kUnitCells + 1];    // first sorted atom of each unit cell
    int32_t sys_cell0[kUnitSystems + 1];
    int32_t sys_atom0[kUnitSystems + 1];
    int32_t sys_row8[kUnitSystems];   // 8-byte slot rows: some winding, or reach beyond one cell
    int32_t next_item;                // the next count work item a wave takes
    int32_t sys_group[kUnitSystems];  // home atoms per count work item of each system
    uint32_t cell_info[kUnitCells];   // per unit cell: system | (c1 - 1) << 8 | (c2 - 1) << 16 | (c3 - 1) << 24
    int16_t item0[kUnitCells + 1];    // first count work item of each unit cell
    int16_t item_cell[kUnitItems];    // unit cell of each count work item
    union {
        SmallSort s;
        SmallRows r[kSmallThreads / kWave];
    } u;
};

constexpr uint32_t kPackNeg = (1024u) | (1024u << 10) | (1024u << 20);   // pack(-w) = kPackNeg - pack(w)

struct SmallArgs {
    const double *__restrict__ X;
    const SysDev *__restrict__ sys;
    const int64_t *__restrict__ sys_atom_base;
    const int32_t *__restrict__ sys_cell_base;
    const int32_t *__restrict__ units;
    const int32_t *__restrict__ order;
    int32_t *__restrict__ atom_sys;
    Rec *__restrict__ rec;
    void *__restrict__ perm_out;
    void *__restrict__ cell_id_out;
    void *__restrict__ cell_offsets_out;
    int32_t *__restrict__ pair_count;
    void *__restrict__ slots;
    double *__restrict__ slotsR;
    int32_t cap;
    Status *__restrict__ status;
};

// wrap_and_shift (cell_list.jl:105-120) of a neighbour cell index a within a
// few cells of [1, n] on a periodic axis (rn = 1 / n as float: the floor
// of (a - 1/2) / n is exact for |a| < 1024); on an open axis only in-bounds
// indices reach here.
__device__ __forceinline__ void wrap_small(int32_t a, int32_t n, float rn, bool p, int32_t &w, int32_t &s)
{
    if (!p) {
        w = a;
        s = 0;
        return;
    }
    const int32_t q = (int32_t)floorf(((float)a - 0.5f) * rn);
    w = a - q * n;
    s = q;
}

template <bool WANT_R, typename TI>
__global__ __launch_bounds__(kSmallThreads) void small_kernel(SmallArgs A)
{
    __shared__ SmallLDS L;
    const int tid = (int)threadIdx.x, lane = lane_id(), wv = sgpr(tid >> 6);
    const int u = A.order[2 * blockIdx.x], pinfo = A.order[2 * blockIdx.x + 1];
    const int part = pinfo & 0xffff, parts = pinfo >> 16;   // this workgroup counts items part + parts k
    const bool writer = part == 0;                           // the first part writes the cell list
    const int s0 = A.units[u], ns = A.units[u + 1] - s0;
    const int64_t A0 = A.sys_atom_base[s0];
    const int nat_u = (int)(A.sys_atom_base[s0 + ns] - A0);
    const int32_t C0 = A.sys_cell_base[s0];
    const int M_u = A.sys_cell_base[s0 + ns] - C0;
    for (int k = tid; k <= ns; k += kSmallThreads) {
        L.sys_cell0[k] = A.sys_cell_base[s0 + k] - C0;
        L.sys_atom0[k] = (int32_t)(A.sys_atom_base[s0 + k] - A0);
        if (k < ns) {
            const SysDev &sd = A.sys[s0 + k];
            L.sys_row8[k] = (sd.nxyz[0] > 1 || sd.nxyz[1] > 1 || sd.nxyz[2] > 1) ? 1 : 0;
            // candidates per home atom: the system's mean cell occupancy times its neighbour cells
            const int nseg = (2 * sd.nxyz[0] + 1) * (2 * sd.nxyz[1] + 1) * (2 * sd.nxyz[2] + 1);
            const int chunks = max(1, (int)((nseg * (int64_t)sd.nat / max(sd.M, 1) + kWave - 1) / kWave));
            L.sys_group[k] = min(kWave, max(kMinGroup, kItemTests / chunks));
        }
    }
    for (int c = tid; c < M_u; c += kSmallThreads) L.u.s.fill[c] = 0;
    if (tid == 0) {
        LAtom f;
        f.x = f.y = f.z = kFar;
        f.j = 0;
        f.t = kPackZero;
        L.a[kUnitAtoms] = f;
        L.next_item = kSmallThreads / kWave;   // wave w starts at the part's item w
    }
    __syncthreads();
    // cell and winding of every atom (position_to_cell_index + bin_wrap_and_shift,
    // cell_list.jl:66-74, 141-147), atoms per cell
    bool bad = false, big = false;
    for (int i = tid; i < nat_u; i += kSmallThreads) {
        const int k = find_sys(L.sys_atom0, ns, (int32_t)i);
        const SysDev &sd = A.sys[s0 + k];
        const int64_t g = A0 + i;
        int32_t c[3], w[3];
        bad |= !cell_and_winding(sd, A.X[3 * g], A.X[3 * g + 1], A.X[3 * g + 2], c, w);
        const int key = L.sys_cell0[k] + (c[0] - 1) + (c[1] - 1) * sd.n[0] + (c[2] - 1) * sd.n[0] * sd.n[1];
        const uint32_t wp = pack_wind(w[0], w[1], w[2]);
        L.u.s.key[i] = (int16_t)key;
        L.u.s.wk[i] = wp;
        L.u.s.sy[i] = (uint8_t)k;
        if (wp != kPackZero) L.sys_row8[k] = 1;
        big |= (wp & kWindBig) != 0;
        atomicAdd(&L.u.s.fill[key], 1);
        if (writer) A.atom_sys[g] = s0 + k;   // the fill kernel's map of atoms to systems
    }
    if (bad) atomicOr(&A.status->bad_position, 1);
    if (big) atomicOr(&A.status->small_fallback, 1);   // the general pipeline redoes the build
    __syncthreads();
    {   // cell starts: exclusive scan of the counts, two cells per thread
        const int c = 2 * tid;
        const int v0 = c < M_u ? L.u.s.fill[c] : 0, v1 = c + 1 < M_u ? L.u.s.fill[c + 1] : 0;
        int64_t tot;
        const int ex = (int)block_excl_scan_n<kSmallThreads>(v0 + v1, tot);   // its barriers end every read of fill
        if (c < M_u) L.cs[c] = ex;
        if (c + 1 < M_u) L.cs[c + 1] = ex + v0;
        if (tid == 0) L.cs[M_u] = (int32_t)tot;
        if (c < M_u) L.u.s.fill[c] = 0;
        if (c + 1 < M_u) L.u.s.fill[c + 1] = 0;
    }
    __syncthreads();
    for (int i = tid; i < nat_u; i += kSmallThreads) {   // atoms grouped by cell
        const int key = L.u.s.key[i];
        L.u.s.grp[L.cs[key] + atomicAdd(&L.u.s.fill[key], 1)] = (int16_t)i;
    }
    __syncthreads();
    // each atom's place: its cell's start plus the atoms of its cell with a
    // smaller index (the stable order); the sorted record, perm and cell_id
    for (int i = tid; i < nat_u; i += kSmallThreads) {
        const int key = L.u.s.key[i];
        const int st = L.cs[key], en = L.cs[key + 1];
        int rank = 0;
        for (int t = st; t < en; ++t) rank += L.u.s.grp[t] < i;
        const int p = st + rank;
        const int k = L.u.s.sy[i];
        const int64_t g = A0 + i;
        LAtom la;
        la.x = A.X[3 * g];
        la.y = A.X[3 * g + 1];
        la.z = A.X[3 * g + 2];
        la.j = i - L.sys_atom0[k];
        la.t = L.u.s.wk[i];
        L.a[p] = la;
        if (writer) {
            Rec r;
            r.x = la.x;
            r.y = la.y;
            r.z = la.z;
            r.j = la.j;
            r.w = la.t;
            A.rec[A0 + p] = r;
            store_ti<TI>(A.perm_out, A0 + p, (int64_t)la.j + 1);
            store_ti<TI>(A.cell_id_out, A0 + p, (int64_t)(key - L.sys_cell0[k]) + 1);
        }
    }
    // cell offsets: per system, 1-based, its cells then a trailing entry;
    // the system and the grid coordinates of every cell for the count
    for (int c = tid; c < M_u; c += kSmallThreads) {
        const int k = find_sys(L.sys_cell0, ns, (int32_t)c);
        const int lc = c - L.sys_cell0[k], n0 = A.sys[s0 + k].n[0], n1 = A.sys[s0 + k].n[1];
        L.cell_info[c] = (uint32_t)k | (uint32_t)(lc % n0) << 8 | (uint32_t)((lc / n0) % n1) << 16 |
                         (uint32_t)(lc / (n0 * n1)) << 24;
        const int64_t gc = (int64_t)C0 + c + s0 + k;
        if (writer) {
            store_ti<TI>(A.cell_offsets_out, gc, (int64_t)(L.cs[c] - L.sys_atom0[k]) + 1);
            if (c == L.sys_cell0[k + 1] - 1)
                store_ti<TI>(A.cell_offsets_out, gc + 1, (int64_t)(L.sys_atom0[k + 1] - L.sys_atom0[k]) + 1);
        }
    }
    __syncthreads();   // cell starts, cell info and the sorted atoms are complete
    {   // count work items, numbered in cell order: the cells' item counts, scanned
        const int c = 2 * tid;
        int v0 = 0, v1 = 0;
        if (c < M_u) {
            const int g = L.sys_group[L.cell_info[c] & 255u];
            v0 = (L.cs[c + 1] - L.cs[c] + g - 1) / g;
        }
        if (c + 1 < M_u) {
            const int g = L.sys_group[L.cell_info[c + 1] & 255u];
            v1 = (L.cs[c + 2] - L.cs[c + 1] + g - 1) / g;
        }
        int64_t tot;
        const int ex = (int)block_excl_scan_n<kSmallThreads>(v0 + v1, tot);
        if (c < M_u) L.item0[c] = (int16_t)ex;
        if (c + 1 < M_u) L.item0[c + 1] = (int16_t)(ex + v0);
        if (tid == 0) L.item0[M_u] = (int16_t)tot;
        __syncthreads();
        for (int t = 0; t < v0; ++t) L.item_cell[ex + t] = (int16_t)c;
        for (int t = 0; t < v1; ++t) L.item_cell[ex + v0 + t] = (int16_t)(c + 1);
    }
    __syncthreads();   // the items are numbered; the sort's LDS now holds the waves' segment maps
    const int nitems_u = L.item0[M_u];
    SmallRows &Rw = L.u.r[wv];
    uint8_t *fl = reinterpret_cast<uint8_t *>(Rw.flag);
    for (int it = part + parts * wv; it < nitems_u; it = part + parts * take_next(&L.next_item)) {
        const int hc = sgpr(L.item_cell[it]);
        const int ig = it - sgpr(L.item0[hc]);   // the item's group of home atoms in its cell
        const int h0 = sgpr(L.cs[hc]), nha_c = sgpr(L.cs[hc + 1]) - h0;
        const uint32_t info = (uint32_t)sgpr((int)L.cell_info[hc]);
        const int k = (int)(info & 255u);
        const int grp = sgpr(L.sys_group[k]);
        const int hg0 = ig * grp, nha = min(nha_c, hg0 + grp);   // home atoms [hg0, nha) of the cell
        const SysDev &sd = A.sys[s0 + k];
        const int cb = sgpr(L.sys_cell0[k]);
        const int n0 = sd.n[0], n1 = sd.n[1], n2 = sd.n[2];
        const int ci0 = (int)((info >> 8) & 255u) + 1, ci1 = (int)((info >> 16) & 255u) + 1, ci2 = (int)(info >> 24) + 1;
        const int nx0 = sd.nxyz[0], nx1 = sd.nxyz[1], nx2 = sd.nxyz[2];
        const int sxn = 2 * nx0 + 1, syn = 2 * nx1 + 1;
        // quotients of the small integers below by float reciprocals: exact,
        // as every numerator is within a few hundred of zero and a half away
        // from a multiple of the divisor
        const float rsx = 1.0f / (float)sxn, rsy = 1.0f / (float)syn;
        const float rn0 = 1.0f / (float)n0, rn1 = 1.0f / (float)n1, rn2 = 1.0f / (float)n2;
        const int nseg = sxn * syn * (2 * nx2 + 1);
        const int center = (nx2 * syn + nx1) * sxn + nx0;   // (dx, dy, dz) = 0: the home cell, zero shift
        const bool p0 = (sd.pbc & 1) != 0, p1 = (sd.pbc & 2) != 0, p2 = (sd.pbc & 4) != 0;
        const bool narrow = sgpr(L.sys_row8[k]) == 0;   // every S in [-1, 1]^3: 4-byte slot entries
        for (int hg = hg0; hg < nha; hg += kWave) {
            const int nh = sgpr(min(kWave, nha - hg));
            const int hb = h0 + hg;
            const int hl = hb + (lane < nh ? lane : 0);
            const uint32_t ht_l = kPackNeg - L.a[hl].t;   // pack(-w_i)
            const int32_t hj_l = L.a[hl].j;
            int32_t cnt_l = 0;
            for (int sb = 0; sb < nseg; sb += kWave) {
                // lane: neighbour cell sb + lane in the reference order, its
                // sorted atoms and cell shift (_is_cell_in_bounds,
                // _get_neighbor_cell, gpu_kernels.jl:20-47)
                const int gs = sb + lane;
                int st = 0, len = 0;
                uint32_t shp = kPackZero;
                if (gs < nseg) {
                    const int t2 = (int)(((float)gs + 0.5f) * rsx), t3 = (int)(((float)t2 + 0.5f) * rsy);
                    const int dx = gs - t2 * sxn - nx0, dy = t2 - t3 * syn - nx1, dz = t3 - nx2;
                    const int a0 = ci0 + dx, a1 = ci1 + dy, a2 = ci2 + dz;
                    if ((p0 || (a0 >= 1 && a0 <= n0)) && (p1 || (a1 >= 1 && a1 <= n1)) &&
                        (p2 || (a2 >= 1 && a2 <= n2))) {
                        int32_t c0, c1, c2, s0x, s0y, s0z;
                        wrap_small(a0, n0, rn0, p0, c0, s0x);
                        wrap_small(a1, n1, rn1, p1, c1, s0y);
                        wrap_small(a2, n2, rn2, p2, c2, s0z);
                        const int uc = cb + (c0 - 1) + (c1 - 1) * n0 + (c2 - 1) * n0 * n1;
                        st = L.cs[uc];
                        len = L.cs[uc + 1] - st;
                        shp = pack_wind(s0x, s0y, s0z);
                    }
                }
                const int incl = wave_incl_scan(len);
                const int ex = incl - len;
                const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
                const bool ne = len > 0;
                const int q = (int)mbcnt64(__builtin_amdgcn_ballot_w64(ne));
                if (ne) {
                    Rw.dl[q] = st - ex;
                    Rw.sh[q] = shp;
                }
                // home atom hg + a is candidate ex(center) + hg + a of the
                // center segment, when that segment is in this batch
                const int cl = center - sb;
                const int self0 = (cl >= 0 && cl < kWave) ? __builtin_amdgcn_readlane(ex, cl) : -(1 << 28);
                for (int k0 = 0; k0 < total; k0 += kPassCands) {
                    const int nch = sgpr(min(kBrickChunks, (total - k0 + kWave - 1) / kWave));
                    // marks where a non-empty segment other than the first starts
                    if (lane < kPassCands / 4) Rw.flag[lane] = 0u;
                    const int pp = ex - k0;
                    if (ne && q > 0 && pp > 0 && pp < kPassCands) fl[pp] = 1;
                    int acc = (int)__popcll(__builtin_amdgcn_ballot_w64(ne && q > 0 && pp <= 0));
                    double cx[kBrickChunks], cy[kBrickChunks], cz[kBrickChunks];
                    uint64_t ent[kBrickChunks];
#pragma unroll
                    for (int g = 0; g < kBrickChunks; ++g) {
                        const uint64_t bm = __builtin_amdgcn_ballot_w64(fl[g * kWave + lane] != 0);
                        const int qk = acc + (int)(bm & 1u) + (int)mbcnt64(bm >> 1);
                        acc += (int)__popcll(bm);
                        const int kk = k0 + g * kWave + lane;
                        const int li = kk < total ? kk + Rw.dl[qk] : kUnitAtoms;   // else the far sentinel
                        const uint32_t sh = Rw.sh[qk];
                        const LAtom c = L.a[li];
                        cx[g] = c.x;
                        cy[g] = c.y;
                        cz[g] = c.z;
                        const uint32_t t = sh - c.t + kPackZero;   // pack(s - w_j)
                        const uint32_t code = (t & 3u) | ((t >> 8) & 12u) | ((t >> 16) & 48u);
                        const uint32_t lo = narrow ? ((uint32_t)c.j | code << 26) : (uint32_t)c.j;
                        ent[g] = (uint64_t)lo | ((uint64_t)t << 32);
                    }
                    pass_home_atoms<WANT_R, false, false>(narrow, false, nch, cx, cy, cz, ent, L.a + hb, nh, ht_l, hj_l,
                                                   self0 + hg - k0, sd.C, sd.rc2, sd.atom_base, A.slots, A.slotsR,
                                                   A.cap, A.status, cnt_l);
                }
            }
            if (lane < nh) A.pair_count[sd.atom_base + hj_l] = cnt_l | (narrow ? 0 : kRow8Bit);
        }
    }
}

template <bool R, typename TI>
static void small_launch(const PipelineBufs &b, hipStream_t st)
{
    SmallArgs A{b.X,          b.sys,     b.sys_atom_base, b.sys_cell_base, b.units,  b.unit_order,
                b.atom_sys,   b.rec,     b.perm_out,      b.cell_id_out,   b.cell_offsets_out,
                b.pair_count, b.slots,   b.slotsR,        b.slot_cap,      b.status};
    hipLaunchKernelGGL((small_kernel<R, TI>), dim3((unsigned)b.nitems), dim3(kSmallThreads), 0, st, A);
}

void launch_small(const PipelineBufs &b, bool want_R, hipStream_t st)
{
    if (b.nunits <= 0 || b.nitems <= 0) return;
    if (b.ti_bytes == 8) {
        if (want_R) small_launch<true, int64_t>(b, st);
        else small_launch<false, int64_t>(b, st);
    } else {
        if (want_R) small_launch<true, int32_t>(b, st);
        else small_launch<false, int32_t>(b, st);
    }
}

// Fill pass (fill_pairs_kernel!, gpu_kernels.jl:159-180): one wave per tile
// of TP = 64 TW consecutive output pairs, one lane per pair.  Tiles start at
// multiples of TP, so every store instruction covers one aligned 64-pair
// block of i, j, S (and R): the stores write whole cache lines, and lines
// shared between waves occur only at tile edges.  A wave issues the slot
// loads of all its TW blocks before its first store.
//
// The pair scan writes, per tile t, the atom A0 owning pair t TP and how
// many of its pairs come before that one; the next tile's entry bounds the
// tile's atoms from above (A1).  When the tile spans at most 64 atoms the
// wave loads their counts, one atom per lane, rebuilds their run starts by
// a scan, and finds the atom of each of its pairs by a binary search over
// the lanes.  A tile spanning more atoms (possible only when atoms without
// pairs lie in between) is filled by a per-lane search of pair_off instead.

// Slot entries, read raw (8 bytes for every narrow format, no branch) and
// decoded only after all of a wave's loads are issued.
template <bool WIDE>
struct RawEntry;
template <>
struct RawEntry<false> {
    uint64_t e;
    bool w8;
};
template <>
struct RawEntry<true> {
    int4 e;
};

template <bool WIDE>
__device__ __forceinline__ RawEntry<WIDE> load_entry(const void *__restrict__ slots, int64_t rs, int r, bool row8)
{
    RawEntry<WIDE> x;
    if constexpr (WIDE) {
        x.e = reinterpret_cast<const int4 *>(slots)[rs + r];
    } else {
        // 8-byte rows: entry rs + r; 4-byte rows: 4-byte entry 2 rs + r of
        // the same row (the 8 bytes read stay inside the row: 4 r + 8 <= 8 cap)
        const int64_t byte = 8 * rs + ((int64_t)r << (row8 ? 3 : 2));
        __builtin_memcpy(&x.e, reinterpret_cast<const char *>(slots) + byte, 8);
        x.w8 = row8;
    }
    return x;
}

template <bool WIDE>
__device__ __forceinline__ void decode_entry(const RawEntry<WIDE> &x, int32_t &j, int32_t &Sx, int32_t &Sy,
                                             int32_t &Sz)
{
    if constexpr (WIDE) {
        j = x.e.x; Sx = x.e.y; Sy = x.e.z; Sz = x.e.w;
    } else {
        const uint32_t lo = (uint32_t)x.e, hi = (uint32_t)(x.e >> 32);
        int32_t px, py, pz;
        unpack_wind(hi, px, py, pz);
        const uint32_t code = lo >> 26;
        j = x.w8 ? (int32_t)lo : (int32_t)(lo & (uint32_t)(kNarrowMaxAtoms - 1));
        Sx = x.w8 ? px : narrow_comp(code, 0);
        Sy = x.w8 ? py : narrow_comp(code, 1);
        Sz = x.w8 ? pz : narrow_comp(code, 2);
    }
}

// Decode and store one pair at output position pos.  The stores are
// non-temporal: the output streams past the caches, which keep holding the
// slot rows the count pass wrote and the fill reads (with cached stores the
// 1M-atom fill ran at 147 us, with these at 103 us).
template <bool WANT_R, bool WIDE, typename TI>
__device__ __forceinline__ void fill_store_one(const RawEntry<WIDE> &e, int32_t is, int64_t ridx, int64_t pos,
                                               bool live, const double *__restrict__ slotsR, TI *__restrict__ i_out,
                                               TI *__restrict__ j_out, TI *__restrict__ S_out,
                                               double *__restrict__ R_out)
{
    int32_t j, Sx, Sy, Sz;
    decode_entry<WIDE>(e, j, Sx, Sy, Sz);   // outside the branch: it covers the stores only
    if (live) {
        __builtin_nontemporal_store((TI)is, i_out + pos);
        __builtin_nontemporal_store((TI)(j + 1), j_out + pos);
        TI *S = S_out + 3 * pos;
        __builtin_nontemporal_store((TI)Sx, S);
        __builtin_nontemporal_store((TI)Sy, S + 1);
        __builtin_nontemporal_store((TI)Sz, S + 2);
        if (WANT_R) {
            const double *Rs = slotsR + 3 * ridx;
            double *Rd = R_out + 3 * pos;
            __builtin_nontemporal_store(Rs[0], Rd);
            __builtin_nontemporal_store(Rs[1], Rd + 1);
            __builtin_nontemporal_store(Rs[2], Rd + 2);
        }
    }
}

template <bool WANT_R, bool WIDE, typename TI, int TW>
__global__ __launch_bounds__(1024) void fill_kernel(
    const int64_t *__restrict__ sys_atom_base, const int32_t *__restrict__ atom_sys, int nsys, int64_t nat, const int32_t *__restrict__ pair_count,
    const int64_t *__restrict__ pair_off, const int32_t *__restrict__ tile_atom, int64_t tile_cap,
    const void *__restrict__ slots, const double *__restrict__ slotsR, const Status *__restrict__ status,
    Status *__restrict__ status_pub, int32_t cap, int64_t pair_capacity, TI *__restrict__ i_out,
    TI *__restrict__ j_out, TI *__restrict__ S_out, double *__restrict__ R_out)
{
    constexpr int TP = TW * kWave;
    constexpr int32_t kBeyond = 1 << 30;   // window position past every pair of the tile
    const int lane = lane_id();
    // tiles in block order: the waves in flight write one advancing front of
    // the output (an XCD-contiguous block order measured much slower)
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x / kWave) + sgpr((int)(threadIdx.x / kWave));   // wave-uniform
    const int64_t Q0 = w * TP;
    if (blockIdx.x == 0 && threadIdx.x < kWave) {
        // the status is final once the fill runs: publish it to the host's
        // pinned copy (system-scope stores), then mark it published, so the
        // host can go on while the fill writes the pairs
        constexpr int kWords = sizeof(Status) / 4;
        constexpr int kPub = offsetof(Status, published) / 4;
        static_assert(kWords <= kWave, "one wave publishes the status");
        if (threadIdx.x < kWords && threadIdx.x != kPub) {
            const int32_t v = reinterpret_cast<const int32_t *>(status)[threadIdx.x];
            __hip_atomic_store(reinterpret_cast<int32_t *>(status_pub) + threadIdx.x, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");   // system scope: every lane's words land first
        __builtin_amdgcn_wave_barrier();
        if (threadIdx.x == 0)
            __hip_atomic_store(&status_pub->published, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    const Status st = *status;
    // the atoms owning this tile's first pair and the next tile's, read with
    // the status (indexes clamped to the table, which holds at least two
    // entries: waves past the pairs return below without using them)
    const int64_t wc = w + 1 < tile_cap ? w : tile_cap - 2;
    const int32_t a0s = tile_atom[2 * wc], skip0 = tile_atom[2 * wc + 1], a1s = tile_atom[2 * wc + 2];
    __asm__ volatile("" ::"s"(a0s), "s"(skip0), "s"(a1s));   // keeps the loads ahead of the branch: their latency overlaps the status's
    if ((st.npairs > pair_capacity) | (st.max_count > cap) | (!WIDE & (st.wide != 0)) | (Q0 >= st.npairs)) return;
    const int nlive = (int)(st.npairs - Q0 < TP ? st.npairs - Q0 : TP);   // pairs of this tile
    const int64_t A0 = a0s;
    const int64_t A1 = Q0 + TP < st.npairs ? (int64_t)a1s : nat - 1;   // every pair of the tile is owned in [A0, A1]
    if (A1 - A0 < kWave) {
        // the window: atom A0 + k on lane k < span, its run of pairs
        // [fo, fo + cnt) relative to Q0 (the exclusive scan of the counts,
        // less the skip0 pairs of atom A0 before the tile); other lanes hold
        // kBeyond
        const int span = (int)(A1 - A0) + 1;
        const int64_t ak = A0 + lane;
        const bool vk = lane < span;
        const int32_t pck = vk ? pair_count[ak] : 0;
        const int32_t cntk = pck & kCountMask;
        const int32_t incl = wave_incl_scan(cntk);   // lanes past the span hold 0
        // runs past the tile end keep their exact start (<= TP + 64 cap)
        const int32_t fo = vk ? incl - cntk - skip0 : kBeyond;
        const uint64_t row8 = __ballot((pck & kRow8Bit) != 0);
        const int32_t fo0 = -skip0;   // atom A0 owns pair Q0
        const bool batch = nsys > 1;
        int32_t il = 0;   // batches: 1-based system-local index of the lane's atom
        if (batch && vk) il = (int32_t)(ak - sys_atom_base[atom_sys[ak]] + 1);
        // pair q = 64 u + lane belongs to the largest window lane s[u] whose
        // run starts at or before q: one binary search per block, all TW
        // searches advanced together (TW independent lane reads per step),
        // over just enough steps for the tile's span of atoms
        int top = 0;
        while (2 * top < span) top = top ? 2 * top : 1;   // 2 top >= span: s reaches span - 1
        int s[TW], f[TW];
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            s[u] = 0;
            f[u] = fo0;
        }
        for (int step = top; step >= 1; step >>= 1) {
#pragma unroll
            for (int u = 0; u < TW; ++u) {
                const int f2 = __shfl(fo, s[u] + step, kWave);
                if (f2 <= u * kWave + lane) {
                    s[u] += step;
                    f[u] = f2;
                }
            }
        }
        RawEntry<WIDE> e[TW];
        int32_t is[TW];
        int64_t ridx[TW];   // used only with R
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const int q = u * kWave + lane;
            const bool live = q < nlive;
            // pairs past the tile's end read entry 0 of row 0 (in bounds, never
            // stored): no branch around the load
            const int r = live ? q - f[u] : 0;
            const int64_t rs = live ? (A0 + s[u]) * (int64_t)cap : 0;
            const bool w8 = ((row8 >> s[u]) & 1) != 0;
            is[u] = batch ? __shfl(il, s[u], kWave) : (int32_t)(A0 + s[u] + 1);
            ridx[u] = rs + r;
            e[u] = load_entry<WIDE>(slots, rs, r, w8);
        }
#pragma unroll
        for (int u = 0; u < TW; ++u) {
            const int q = u * kWave + lane;
            fill_store_one<WANT_R, WIDE, TI>(e[u], is[u], ridx[u], Q0 + q, q < nlive, slotsR, i_out,
                                             j_out, S_out, R_out);
        }
        return;
    }
    // the tile's pairs span more than 64 atoms: each lane searches pair_off
    // between A0 and A1 for the atom of its pair
    const int64_t hi0 = A1;
    for (int u = 0; u < TW; ++u) {
        const int q = u * kWave + lane;
        if (q < nlive) {
            const int64_t P = Q0 + q;
            int64_t lo = A0, hi = hi0;   // largest atom in [lo, hi] whose run starts at or before P
            while (lo < hi) {
                const int64_t mid = (lo + hi + 1) >> 1;
                if (pair_off[mid] <= P) lo = mid;
                else hi = mid - 1;
            }
            const int32_t pc = pair_count[lo];
            const int r = (int)(P - pair_off[lo]);
            const int64_t rs = lo * (int64_t)cap;
            int32_t isv = (int32_t)(lo + 1);
            if (nsys > 1) isv = (int32_t)(lo - sys_atom_base[atom_sys[lo]] + 1);
            const RawEntry<WIDE> e = load_entry<WIDE>(slots, rs, r, (pc & kRow8Bit) != 0);
            fill_store_one<WANT_R, WIDE, TI>(e, isv, rs + r, P, true, slotsR, i_out, j_out, S_out, R_out);
        }
    }
}


// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <bool R, bool W>
static void count_launch(const PipelineBufs &b, hipStream_t st)
{
    if (b.count_block == 256)
        hipLaunchKernelGGL((count_kernel<R, W, 256>), dim3((unsigned)b.nbricks), dim3(256), 0, st, b.sys,
                           b.sys_cell_base, b.sys_brick_base, b.nsys, b.M, b.nat, b.cell_start, b.rec, b.wbig,
                           b.pair_count, b.slots, b.slotsR, b.slot_cap, b.status);
    else
        hipLaunchKernelGGL((count_kernel<R, W, 512>), dim3((unsigned)b.nbricks), dim3(512), 0, st, b.sys,
                           b.sys_cell_base, b.sys_brick_base, b.nsys, b.M, b.nat, b.cell_start, b.rec, b.wbig,
                           b.pair_count, b.slots, b.slotsR, b.slot_cap, b.status);
}

void launch_count(const PipelineBufs &b, bool want_R, bool wide, hipStream_t st)
{
    if (b.nunits > 0) {   // cell list, counts and slot rows of every unit of small systems
        launch_small(b, want_R, st);
        return;
    }
    if (b.nat == 0 || b.nbricks == 0) return;
    if (want_R) {
        if (wide) count_launch<true, true>(b, st);
        else count_launch<true, false>(b, st);
    } else {
        if (wide) count_launch<false, true>(b, st);
        else count_launch<false, false>(b, st);
    }
}

template <bool R, bool W, typename TI, int TW>
static bool fill_launch(const PipelineBufs &b, hipStream_t st)
{
    // one wave per tile of the pair capacity (waves past the pairs return)
    const int bs = b.fill_block;
    const int64_t waves = (b.pair_capacity + TW * kWave - 1) / (TW * kWave);
    if (waves <= 0) return false;
    const unsigned nblk = (unsigned)((waves * kWave + bs - 1) / bs);
    hipLaunchKernelGGL((fill_kernel<R, W, TI, TW>), dim3(nblk), dim3(bs), 0, st, b.sys_atom_base, b.atom_sys, b.nsys,
                       b.nat,
                       b.pair_count, b.pair_off, b.tile_atom, b.tile_cap, b.slots, b.slotsR, b.status, b.status_pub,
                       b.slot_cap, b.pair_capacity, (TI *)b.i_out, (TI *)b.j_out, (TI *)b.S_out, b.R_out);
    return true;
}

template <bool R, bool W, typename TI>
static bool fill_tiles(const PipelineBufs &b, hipStream_t st)
{
    if (b.fill_pairs == 128) return fill_launch<R, W, TI, 2>(b, st);
    if (b.fill_pairs == 512) return fill_launch<R, W, TI, 8>(b, st);
    return fill_launch<R, W, TI, 4>(b, st);
}

template <typename TI>
static bool fill_dispatch(const PipelineBufs &b, bool want_R, bool wide, hipStream_t st)
{
    if (want_R) return wide ? fill_tiles<true, true, TI>(b, st) : fill_tiles<true, false, TI>(b, st);
    return wide ? fill_tiles<false, true, TI>(b, st) : fill_tiles<false, false, TI>(b, st);
}

bool launch_fill(const PipelineBufs &b, bool want_R, bool wide, hipStream_t st)
{
    if (b.nat == 0) return false;
    return b.ti_bytes == 8 ? fill_dispatch<int64_t>(b, want_R, wide, st) : fill_dispatch<int32_t>(b, want_R, wide, st);
}

}  // namespace nlh
