// Pair search of the sort-based neighbour list: count (with hit masks) and
// fill, the MI355X counterparts of count_neighbours_kernel! and
// fill_pairs_kernel! (src/gpu_kernels.jl:132-180) around the inner loop
// _for_each_neighbor_pair (src/gpu_kernels.jl:58-101).
//
// Work unit: a brick of up to 4x4x4 home cells per 256-thread workgroup.
// The brick and its one-cell halo are staged in LDS in one load phase (every
// global load of the phase in flight together), then each wave takes home
// cells and tests candidates from LDS only:
//   - candidates of a home cell are its 27 neighbour cells flattened in the
//     reference order (dz, dy, dx; gpu_kernels.jl:69-71) -- nine contiguous
//     LDS rows -- and cut into 64-candidate chunks, one candidate per lane;
//   - the home atoms stream through the wave as uniform values; each
//     (home atom, chunk) test is one ballot, so a home atom's neighbours keep
//     exactly the reference traversal order;
//   - the count pass stores each ballot (kMaskChunks chunks per atom); the
//     fill pass expands stored ballots with one lane per hit, so consecutive
//     lanes write an atom's run of (i, j, S) contiguously.
// Bricks whose neighbour reach exceeds one cell, whose halo does not fit
// LDS, or whose windings do not pack take a per-wave path that reads global
// memory (one home cell per wave), with the same enumeration and output.
#include "device_util.h"

namespace nlh {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int sgpr(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    const uint32_t lo = __shfl((uint32_t)v, src, kWave), hi = __shfl((uint32_t)(v >> 32), src, kWave);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double shfl_d(double v, int src)
{
    return __longlong_as_double((long long)shfl64((uint64_t)__double_as_longlong(v), src));
}

// Position of the r-th (0-based) set bit of m.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r)
{
    int pos = 0;
    uint32_t w = (uint32_t)m;
    int c = __popc(w);
    if (r >= c) {
        r -= c;
        w = (uint32_t)(m >> 32);
        pos = 32;
    }
    c = __popc(w & 0xffffu);
    if (r >= c) { r -= c; w >>= 16; pos += 16; }
    c = __popc(w & 0xffu);
    if (r >= c) { r -= c; w >>= 8; pos += 8; }
    c = __popc(w & 0xfu);
    if (r >= c) { r -= c; w >>= 4; pos += 4; }
    c = __popc(w & 0x3u);
    if (r >= c) { r -= c; w >>= 2; pos += 2; }
    if (r >= (int)(w & 1u)) pos += 1;
    return pos;
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

struct PairsArgs {
    const SysDev *__restrict__ sys;
    const int32_t *__restrict__ sys_cell_base;
    int nsys;
    int32_t M;
    int64_t nat;                // mask stride
    const int32_t *__restrict__ cell_start;
    const Rec *__restrict__ rec;
    const int32_t *__restrict__ wbig;
};

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

// Distance tests of home atoms hg .. hg+nh-1 (sorted indices) against the
// held chunks g in [g0, nch): mine[g] gets home atom a's hit ballot in lane a
// (gpu_kernels.jl:79-93, same evaluation order).
__device__ __forceinline__ void test_chunks(const PairsArgs &A, const Cand (&c)[kChunksPerPass], int g0, int nch,
                                            int32_t hg, int nh, const double *__restrict__ Cm, double rc2,
                                            uint64_t (&mine)[kChunksPerPass])
{
    uint64_t valid[kChunksPerPass], zs[kChunksPerPass];
    double cx[kChunksPerPass], cy[kChunksPerPass], cz[kChunksPerPass];
#pragma unroll
    for (int g = 0; g < kChunksPerPass; ++g) {
        cx[g] = cy[g] = cz[g] = 0.0;
        valid[g] = zs[g] = 0;
        if (g >= g0 && g < nch) {
            valid[g] = __ballot(c[g].valid);
            zs[g] = __ballot(c[g].zero_shift);
        }
    }
    uint32_t wcache = 0xffffffffu;
    Rec hn = A.rec[hg];
    for (int a = 0; a < nh; ++a) {
        const Rec h = hn;
        if (a + 1 < nh) hn = A.rec[hg + a + 1];
        if (h.w != wcache || (h.w & kWindBig)) {   // w_i changed: refresh C' * S per lane
            int32_t wix, wiy, wiz;
            rec_winding(h, A.wbig, hg + a, wix, wiy, wiz);
#pragma unroll
            for (int g = 0; g < kChunksPerPass; ++g)
                if (g >= g0 && g < nch) cell_shift(Cm, c[g].tx + wix, c[g].ty + wiy, c[g].tz + wiz, cx[g], cy[g], cz[g]);
            wcache = h.w;
        }
#pragma unroll
        for (int g = 0; g < kChunksPerPass; ++g) {
            if (g >= g0 && g < nch) {
                const double Rx = (c[g].x - h.x) + cx[g];
                const double Ry = (c[g].y - h.y) + cy[g];
                const double Rz = (c[g].z - h.z) + cz[g];
                const double r2 = (Rx * Rx + Ry * Ry) + Rz * Rz;
                const uint64_t in = __ballot(r2 < rc2);
                const uint64_t self = __ballot(c[g].j == h.j) & zs[g];   // _is_self_interaction
                const uint64_t mask = in & valid[g] & ~self;
                mine[g] = lane_id() == a ? mask : mine[g];
            }
        }
    }
}

__device__ __noinline__ void count_cell(const PairsArgs &A, const SysDev &sd, int cell, int32_t *__restrict__ pair_count,
                           uint64_t *__restrict__ masks)
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
        int cnt = 0;
        int q = 0;
        for (int sb = 0; sb < H.nseg; sb += kWave) {
            int excl, start_g;
            uint32_t shp;
            const int total = segment_batch(A, H, sb, excl, start_g, shp);
            for (int k0 = 0; k0 < total; k0 += kChunksPerPass * kWave) {
                const int nch = sgpr(min(kChunksPerPass, (total - k0 + kWave - 1) / kWave));
                Cand c[kChunksPerPass];
                uint64_t mine[kChunksPerPass];
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) {
                    mine[g] = 0;
                    if (g < nch) c[g] = load_candidate(A, k0 + g * kWave + lane, total, excl, start_g, shp);
                }
                test_chunks(A, c, 0, nch, hg, nh, Cm, rc2, mine);
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) {
                    if (g < nch) {
                        cnt += __popcll(mine[g]);
                        if (q + g < kMaskChunks && lane < nh) masks[(int64_t)(q + g) * A.nat + hg + lane] = mine[g];
                    }
                }
                q += nch;
            }
        }
        if (lane < nh) pair_count[H.abase + hj_lane] = cnt;
    }
}

// Expand one chunk's hit masks with one lane per hit.  mine: lane a's mask
// (home atom a); cursor: lane a's write position; candidate data per lane.
template <bool WANT_R, typename TI>
__device__ __forceinline__ void expand_chunk(uint64_t mine, int64_t &cursor, const Cand &c, int32_t hj_l, int32_t wx_l,
                                             int32_t wy_l, int32_t wz_l, double hx_l, double hy_l, double hz_l,
                                             const double *__restrict__ Cm, TI *__restrict__ i_out,
                                             TI *__restrict__ j_out, TI *__restrict__ S_out, double *__restrict__ R_out)
{
    const int lane = lane_id();
    const int pc = __popcll(mine);
    const int incl = wave_incl_scan(pc);
    const int tot = sgpr(__shfl(incl, kWave - 1, kWave));
    const int exh = incl - pc;
    for (int h0 = 0; h0 < tot; h0 += kWave) {
        const int L = h0 + lane;
        int a = 0;   // home atom: largest a with exh[a] <= L
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const int e2 = __shfl(exh, a + step, kWave);
            if (e2 <= L) a += step;
        }
        const int r = L - __shfl(exh, a, kWave);
        const int l = nth_set_bit(shfl64(mine, a), r);
        const int32_t cj = __shfl(c.j, l, kWave);
        const int32_t Sx = __shfl(c.tx, l, kWave) + __shfl(wx_l, a, kWave);
        const int32_t Sy = __shfl(c.ty, l, kWave) + __shfl(wy_l, a, kWave);
        const int32_t Sz = __shfl(c.tz, l, kWave) + __shfl(wz_l, a, kWave);
        const int32_t hj = __shfl(hj_l, a, kWave);
        const int64_t pos = (int64_t)shfl64((uint64_t)cursor, a) + r;
        double Rx = 0, Ry = 0, Rz = 0;
        if (WANT_R) {
            double cx, cy, cz;
            cell_shift(Cm, Sx, Sy, Sz, cx, cy, cz);
            Rx = (shfl_d(c.x, l) - shfl_d(hx_l, a)) + cx;
            Ry = (shfl_d(c.y, l) - shfl_d(hy_l, a)) + cy;
            Rz = (shfl_d(c.z, l) - shfl_d(hz_l, a)) + cz;
        }
        if (L < tot) {
            i_out[pos] = (TI)(hj + 1);
            j_out[pos] = (TI)(cj + 1);
            TI *S = S_out + 3 * pos;
            S[0] = (TI)Sx;
            S[1] = (TI)Sy;
            S[2] = (TI)Sz;
            if (WANT_R) {
                double *R = R_out + 3 * pos;
                R[0] = Rx;
                R[1] = Ry;
                R[2] = Rz;
            }
        }
    }
    cursor += pc;
}

template <bool WANT_R, typename TI>
__device__ __noinline__ void fill_cell(const PairsArgs &A, const SysDev &sd, int cell, const uint64_t *__restrict__ masks,
                          const int64_t *__restrict__ pair_off, TI *__restrict__ i_out, TI *__restrict__ j_out,
                          TI *__restrict__ S_out, double *__restrict__ R_out)
{
    HomeCell H;
    home_cell(A, sd, cell, H);
    if (H.hs >= H.he) return;
    const int lane = lane_id();
    const double *__restrict__ Cm = sd.C;
    const double rc2 = sd.rc2;
    for (int32_t hg = H.hs; hg < H.he; hg += kWave) {
        const int nh = sgpr(min(kWave, H.he - hg));
        Rec hl;
        if (lane < nh) hl = A.rec[hg + lane];
        else { hl.x = hl.y = hl.z = 0.0; hl.j = 0; hl.w = 0; }
        int32_t wx = 0, wy = 0, wz = 0;
        if (lane < nh) rec_winding(hl, A.wbig, hg + lane, wx, wy, wz);
        int64_t cursor = lane < nh ? pair_off[H.abase + hl.j] : 0;
        int q = 0;
        for (int sb = 0; sb < H.nseg; sb += kWave) {
            int excl, start_g;
            uint32_t shp;
            const int total = segment_batch(A, H, sb, excl, start_g, shp);
            for (int k0 = 0; k0 < total; k0 += kWave, ++q) {
                Cand c[kChunksPerPass];
                c[0] = load_candidate(A, k0 + lane, total, excl, start_g, shp);
                uint64_t mine[kChunksPerPass] = {0, 0, 0, 0};
                if (q < kMaskChunks) {
                    if (lane < nh) mine[0] = masks[(int64_t)q * A.nat + hg + lane];
                } else {
                    test_chunks(A, c, 0, 1, hg, nh, Cm, rc2, mine);
                }
                expand_chunk<WANT_R, TI>(mine[0], cursor, c[0], hl.j, wx, wy, wz, hl.x, hl.y, hl.z, Cm, i_out, j_out,
                                         S_out, R_out);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// brick path
// ---------------------------------------------------------------------------

constexpr int kBrickThreads = 256;
constexpr int kBrickWaves = kBrickThreads / kWave;
constexpr int kBrickCap = 2048;   // halo atoms held in LDS
constexpr int kHaloMax = 512;     // halo cells
constexpr int kBrickUnroll = kBrickCap / kBrickThreads;
constexpr uint32_t kZeroShiftBit = 1u << 30;

struct Brick {
    const SysDev *sd;
    int b0[3];     // 1-based coordinates of the first home cell
    int nh[3];     // home cells per axis (clipped at the grid edge)
    int hd[3];     // halo extent per axis (nh + 2)
    int nhalo, nhome;
    bool fast;     // one-cell reach: the LDS path applies
};

__device__ __forceinline__ void brick_locate(const PairsArgs &A, const int32_t *__restrict__ sys_brick_base, Brick &B)
{
    const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int s = A.nsys == 1 ? 0 : find_sys(sys_brick_base, A.nsys, (int32_t)b);
    const SysDev &sd = A.sys[s];
    B.sd = &sd;
    const int lb = b - sys_brick_base[s];
    const int nb0 = (sd.n[0] + sd.brick[0] - 1) / sd.brick[0];
    const int nb1 = (sd.n[1] + sd.brick[1] - 1) / sd.brick[1];
    const int ib[3] = {lb % nb0, (lb / nb0) % nb1, lb / (nb0 * nb1)};
    for (int k = 0; k < 3; ++k) {
        B.b0[k] = ib[k] * sd.brick[k] + 1;
        B.nh[k] = min(sd.brick[k], sd.n[k] - B.b0[k] + 1);
        B.hd[k] = B.nh[k] + 2;
    }
    B.nhalo = B.hd[0] * B.hd[1] * B.hd[2];
    B.nhome = B.nh[0] * B.nh[1] * B.nh[2];
    B.fast = sd.nxyz[0] == 1 && sd.nxyz[1] == 1 && sd.nxyz[2] == 1 && B.nhalo <= kHaloMax;
}

__device__ __forceinline__ int brick_cell_id(const Brick &B, int hc)
{
    const SysDev &sd = *B.sd;
    const int ix = hc % B.nh[0], iy = (hc / B.nh[0]) % B.nh[1], iz = hc / (B.nh[0] * B.nh[1]);
    return sd.cell_base + (B.b0[0] - 1 + ix) + (B.b0[1] - 1 + iy) * sd.n[0] + (B.b0[2] - 1 + iz) * sd.n[0] * sd.n[1];
}

template <bool POS>
struct BrickLDS {
    double x[POS ? kBrickCap : 1], y[POS ? kBrickCap : 1], z[POS ? kBrickCap : 1];
    int32_t j[kBrickCap];
    uint32_t t[kBrickCap];          // pack(s_halo - w_j) | (s_halo == 0) << 30
    int32_t hoff[kHaloMax + 1];     // LDS offset of each halo cell's atoms
    int32_t hstart[kHaloMax];       // first sorted index of each halo cell
    uint32_t hshift[kHaloMax];      // packed cell shift of each halo cell
    int32_t total, flag;
};

// Stage the brick and its halo in LDS.  Returns false (for every thread of
// the block) when the brick must take the per-wave path.
template <bool POS>
__device__ bool brick_load(const PairsArgs &A, const Brick &B, BrickLDS<POS> &L)
{
    const SysDev &sd = *B.sd;
    const int tid = threadIdx.x;
    const bool p[3] = {(sd.pbc & 1) != 0, (sd.pbc & 2) != 0, (sd.pbc & 4) != 0};
    // halo cell table: thread tid owns halo cells 2 tid and 2 tid + 1
    int cnt[2] = {0, 0};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int h = 2 * tid + u;
        if (h < B.nhalo) {
            const int ux = h % B.hd[0], uy = (h / B.hd[0]) % B.hd[1], uz = h / (B.hd[0] * B.hd[1]);
            const int a[3] = {B.b0[0] - 1 + ux, B.b0[1] - 1 + uy, B.b0[2] - 1 + uz};
            bool inb = true;
            for (int k = 0; k < 3; ++k) inb = inb && (p[k] || (a[k] >= 1 && a[k] <= sd.n[k]));
            int st = 0;
            uint32_t sh = 0;
            if (inb) {
                int32_t c[3], s[3];
                for (int k = 0; k < 3; ++k) wrap_and_shift(a[k], sd.n[k], p[k], c[k], s[k]);
                const int gc = sd.cell_base + (c[0] - 1) + (c[1] - 1) * sd.n[0] + (c[2] - 1) * sd.n[0] * sd.n[1];
                st = A.cell_start[gc];
                cnt[u] = A.cell_start[gc + 1] - st;
                sh = pack_wind(s[0], s[1], s[2]);
            }
            L.hstart[h] = st;
            L.hshift[h] = sh;
        }
    }
    int64_t tot;
    const int ex = (int)block_excl_scan(cnt[0] + cnt[1], tot);
    if (2 * tid < B.nhalo) L.hoff[2 * tid] = ex;
    if (2 * tid + 1 < B.nhalo) L.hoff[2 * tid + 1] = ex + cnt[0];
    if (tid == 0) {
        L.hoff[B.nhalo] = (int32_t)tot;
        L.total = (int32_t)tot;
        L.flag = 0;
    }
    __syncthreads();
    const int total = L.total;
    if (total > kBrickCap) return false;
    // halo atoms: all loads of a thread in flight together
    int pg[kBrickUnroll];
    uint32_t hs[kBrickUnroll];
#pragma unroll
    for (int u = 0; u < kBrickUnroll; ++u) {
        const int idx = tid + u * kBrickThreads;
        pg[u] = 0;
        hs[u] = 0;
        if (idx < total) {
            int lo = 0, hi = B.nhalo - 1;   // largest h with hoff[h] <= idx
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (L.hoff[mid] <= idx) lo = mid;
                else hi = mid - 1;
            }
            pg[u] = L.hstart[lo] + (idx - L.hoff[lo]);
            hs[u] = L.hshift[lo];
        }
    }
    Rec r[kBrickUnroll];
#pragma unroll
    for (int u = 0; u < kBrickUnroll; ++u)
        if (tid + u * kBrickThreads < total) r[u] = A.rec[pg[u]];
#pragma unroll
    for (int u = 0; u < kBrickUnroll; ++u) {
        const int idx = tid + u * kBrickThreads;
        if (idx < total) {
            int32_t wx, wy, wz, sx, sy, sz;
            rec_winding(r[u], A.wbig, pg[u], wx, wy, wz);
            unpack_wind(hs[u], sx, sy, sz);
            const uint32_t t = pack_wind(sx - wx, sy - wy, sz - wz);
            if (t & kWindBig) L.flag = 1;
            L.t[idx] = t | ((sx | sy | sz) == 0 ? kZeroShiftBit : 0u);
            L.j[idx] = r[u].j;
            if (POS) {
                L.x[idx] = r[u].x;
                L.y[idx] = r[u].y;
                L.z[idx] = r[u].z;
            }
        }
    }
    __syncthreads();
    return L.flag == 0;
}

// Per home cell of a brick: its LDS atom range, first sorted index and the
// nine candidate rows (lane r < 9 holds row r).  Returns the candidate count.
struct HomeRows {
    int h0, h1;      // LDS range of the home atoms
    int pbase;       // sorted index of the first home atom
    int rstart, rex; // per lane: LDS start and exclusive candidate prefix of row `lane`
    int total;
};

template <bool POS>
__device__ __forceinline__ HomeRows brick_home(const Brick &B, const BrickLDS<POS> &L, int hc)
{
    HomeRows R;
    const int ix = hc % B.nh[0], iy = (hc / B.nh[0]) % B.nh[1], iz = hc / (B.nh[0] * B.nh[1]);
    const int hh = (ix + 1) + B.hd[0] * ((iy + 1) + B.hd[1] * (iz + 1));
    R.h0 = L.hoff[hh];
    R.h1 = L.hoff[hh + 1];
    R.pbase = L.hstart[hh];
    const int lane = lane_id();
    int len = 0, st = 0;
    if (lane < 9) {   // row (dz, dy) = (lane / 3 - 1, lane % 3 - 1); dx = -1..1 contiguous
        const int dz = lane / 3 - 1, dy = lane % 3 - 1;
        const int hr = ix + B.hd[0] * ((iy + 1 + dy) + B.hd[1] * (iz + 1 + dz));
        st = L.hoff[hr];
        len = L.hoff[hr + 3] - st;
    }
    const int incl = wave_incl_scan(len);
    R.rex = incl - len;
    R.rstart = st;
    R.total = sgpr(__shfl(incl, kWave - 1, kWave));
    return R;
}

// LDS index of candidate k of a home cell (k < total).
__device__ __forceinline__ int row_index(const HomeRows &R, int k)
{
    int m = 0;
#pragma unroll
    for (int step = 8; step >= 1; step >>= 1) {
        const int e2 = __shfl(R.rex, m + step, kWave);
        if (e2 <= k) m += step;
    }
    return __shfl(R.rstart, m, kWave) + (k - __shfl(R.rex, m, kWave));
}

__global__ __launch_bounds__(kBrickThreads) void count_kernel(
    const SysDev *__restrict__ sys, const int32_t *__restrict__ sys_cell_base, const int32_t *__restrict__ sys_brick_base,
    int nsys, int32_t M, int64_t nat, const int32_t *__restrict__ cell_start, const Rec *__restrict__ rec,
    const int32_t *__restrict__ wbig, int32_t *__restrict__ pair_count, uint64_t *__restrict__ masks)
{
    __shared__ BrickLDS<true> L;
    __shared__ double Csh[9];
    const PairsArgs A{sys, sys_cell_base, nsys, M, nat, cell_start, rec, wbig};
    Brick B;
    brick_locate(A, sys_brick_base, B);
    if (threadIdx.x < 9) Csh[threadIdx.x] = B.sd->C[threadIdx.x];   // visible after brick_load's barrier
    const int wv = sgpr((int)(threadIdx.x >> 6));
    const bool lds = B.fast && brick_load<true>(A, B, L);
    if (!lds) {
        for (int hc = wv; hc < B.nhome; hc += kBrickWaves) count_cell(A, *B.sd, brick_cell_id(B, hc), pair_count, masks);
        return;
    }
    const SysDev &sd = *B.sd;
    const double rc2 = sd.rc2;
    const int lane = lane_id();
    constexpr double kFar = 1.0e200;   // invalid candidate: r^2 overflows to +inf, never a hit
    for (int hc = wv; hc < B.nhome; hc += kBrickWaves) {
        const HomeRows R = brick_home(B, L, hc);
        const int nha = R.h1 - R.h0;
        for (int hg = 0; hg < nha; hg += kWave) {
            const int nh = sgpr(min(kWave, nha - hg));
            const int32_t hj_lane = lane < nh ? L.j[R.h0 + hg + lane] : 0;
            int cnt = 0;
            int q = 0;
            for (int k0 = 0; k0 < R.total; k0 += kChunksPerPass * kWave) {
                const int nch = sgpr(min(kChunksPerPass, (R.total - k0 + kWave - 1) / kWave));
                double cxp[kChunksPerPass], cyp[kChunksPerPass], czp[kChunksPerPass];
                // self key: j where the candidate's cell shift is zero, else -3
                // (never a home index), so (key == j_i) is _is_self_interaction
                int32_t key[kChunksPerPass], ctx[kChunksPerPass], cty[kChunksPerPass], ctz[kChunksPerPass];
                uint64_t mine[kChunksPerPass];
                double sx[kChunksPerPass], sy[kChunksPerPass], sz[kChunksPerPass];
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) {
                    mine[g] = 0;
                    cxp[g] = cyp[g] = czp[g] = kFar;
                    sx[g] = sy[g] = sz[g] = 0.0;
                    key[g] = -3;
                    ctx[g] = cty[g] = ctz[g] = 0;
                    if (g < nch) {
                        const int k = k0 + g * kWave + lane;
                        const bool v = k < R.total;
                        const int li = row_index(R, v ? k : 0);
                        if (v) {
                            cxp[g] = L.x[li];
                            cyp[g] = L.y[li];
                            czp[g] = L.z[li];
                            const uint32_t t = L.t[li];
                            key[g] = (t & kZeroShiftBit) ? L.j[li] : -3;
                            unpack_wind(t & ~kZeroShiftBit, ctx[g], cty[g], ctz[g]);
                        }
                    }
                }
                uint32_t wcache = 0xffffffffu;
                for (int a = 0; a < nh; ++a) {
                    const int ha = R.h0 + hg + a;   // uniform LDS address
                    const double hx = L.x[ha], hy = L.y[ha], hz = L.z[ha];
                    const int32_t hj = L.j[ha];
                    const uint32_t ht = (uint32_t)sgpr((int)(L.t[ha] & ~kZeroShiftBit));   // t = -w_i (home shift 0)
                    if (ht != wcache) {
                        int32_t tx, ty, tz;
                        unpack_wind(ht, tx, ty, tz);
#pragma unroll
                        for (int g = 0; g < kChunksPerPass; ++g)
                            if (g < nch) cell_shift(Csh, ctx[g] - tx, cty[g] - ty, ctz[g] - tz, sx[g], sy[g], sz[g]);
                        wcache = ht;
                    }
#pragma unroll
                    for (int g = 0; g < kChunksPerPass; ++g) {
                        if (g < nch) {
                            // R = xj - xi + C' * S ; r^2 = dot(R, R)  (gpu_kernels.jl:89-92)
                            const double Rx = (cxp[g] - hx) + sx[g];
                            const double Ry = (cyp[g] - hy) + sy[g];
                            const double Rz = (czp[g] - hz) + sz[g];
                            const double r2 = (Rx * Rx + Ry * Ry) + Rz * Rz;
                            const uint64_t mask = __ballot((r2 < rc2) & (key[g] != hj));
                            mine[g] = lane == a ? mask : mine[g];
                        }
                    }
                }
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) {
                    if (g < nch) {
                        cnt += __popcll(mine[g]);
                        if (q + g < kMaskChunks && lane < nh)
                            masks[(int64_t)(q + g) * nat + R.pbase + hg + lane] = mine[g];
                    }
                }
                q += nch;
            }
            if (lane < nh) pair_count[sd.atom_base + hj_lane] = cnt;
        }
    }
}

template <bool WANT_R, typename TI>
__global__ __launch_bounds__(kBrickThreads) void fill_kernel(
    const SysDev *__restrict__ sys, const int32_t *__restrict__ sys_cell_base, const int32_t *__restrict__ sys_brick_base,
    int nsys, int32_t M, int64_t nat, const int32_t *__restrict__ cell_start, const Rec *__restrict__ rec,
    const int32_t *__restrict__ wbig, const uint64_t *__restrict__ masks, const int64_t *__restrict__ pair_off,
    const Status *__restrict__ status, TI *__restrict__ i_out, TI *__restrict__ j_out, TI *__restrict__ S_out,
    double *__restrict__ R_out)
{
    __shared__ BrickLDS<WANT_R> L;
    if (status->overflow) return;
    const PairsArgs A{sys, sys_cell_base, nsys, M, nat, cell_start, rec, wbig};
    Brick B;
    brick_locate(A, sys_brick_base, B);
    const int wv = sgpr((int)(threadIdx.x >> 6));
    const bool lds = B.fast && brick_load<WANT_R>(A, B, L);
    if (!lds) {
        for (int hc = wv; hc < B.nhome; hc += kBrickWaves)
            fill_cell<WANT_R, TI>(A, *B.sd, brick_cell_id(B, hc), masks, pair_off, i_out, j_out, S_out, R_out);
        return;
    }
    const SysDev &sd = *B.sd;
    const double *__restrict__ Cm = sd.C;
    const int lane = lane_id();
    for (int hc = wv; hc < B.nhome; hc += kBrickWaves) {
        const HomeRows R = brick_home(B, L, hc);
        const int nha = R.h1 - R.h0;
        if (nha == 0) continue;
        if ((R.total + kWave - 1) / kWave > kMaskChunks) {   // more chunks than stored masks
            fill_cell<WANT_R, TI>(A, sd, brick_cell_id(B, hc), masks, pair_off, i_out, j_out, S_out, R_out);
            continue;
        }
        for (int hg = 0; hg < nha; hg += kWave) {
            const int nh = sgpr(min(kWave, nha - hg));
            const int hl = R.h0 + hg + (lane < nh ? lane : 0);
            const int32_t hj_l = L.j[hl];
            int32_t wx = 0, wy = 0, wz = 0;   // w_i = -(t of the home atom)
            {
                int32_t tx, ty, tz;
                unpack_wind(L.t[hl] & ~kZeroShiftBit, tx, ty, tz);
                wx = -tx; wy = -ty; wz = -tz;
            }
            double hx = 0, hy = 0, hz = 0;
            if (WANT_R) {
                hx = L.x[hl];
                hy = L.y[hl];
                hz = L.z[hl];
            }
            int64_t cursor = lane < nh ? pair_off[sd.atom_base + hj_l] : 0;
            uint64_t mk[kMaskChunks];
            const int nq = (R.total + kWave - 1) / kWave;
#pragma unroll
            for (int q = 0; q < kMaskChunks; ++q)
                mk[q] = (q < nq && lane < nh) ? masks[(int64_t)q * nat + R.pbase + hg + lane] : 0;
            for (int q0 = 0; q0 < nq; q0 += kChunksPerPass) {
                // one pass: chunks q0 .. q0+3, lanes = hits
                uint64_t mine[kChunksPerPass];
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) {
                    mine[g] = 0;
#pragma unroll
                    for (int q = 0; q < kMaskChunks; ++q)
                        if (q == q0 + g) mine[g] = mk[q];
                }
                int pc = 0;
#pragma unroll
                for (int g = 0; g < kChunksPerPass; ++g) pc += __popcll(mine[g]);
                const int incl = wave_incl_scan(pc);
                const int tot = sgpr(__shfl(incl, kWave - 1, kWave));
                const int exh = incl - pc;
                for (int e0 = 0; e0 < tot; e0 += kWave) {
                    const int Lh = e0 + lane;
                    int a = 0;   // home atom: largest a with exh[a] <= Lh
#pragma unroll
                    for (int step = 32; step >= 1; step >>= 1) {
                        const int e2 = __shfl(exh, a + step, kWave);
                        if (e2 <= Lh) a += step;
                    }
                    const int rank = Lh - __shfl(exh, a, kWave);
                    int r = rank, gsel = -1;
                    uint64_t m = 0;
#pragma unroll
                    for (int g = 0; g < kChunksPerPass; ++g) {
                        const uint64_t mg = shfl64(mine[g], a);
                        const int pg = __popcll(mg);
                        const bool here = (gsel < 0) & (r < pg);
                        m = here ? mg : m;
                        gsel = here ? g : gsel;
                        r = (gsel < 0) ? r - pg : r;
                    }
                    const int k = (q0 + gsel) * kWave + nth_set_bit(m, r);
                    const int32_t hj = __shfl(hj_l, a, kWave);
                    const int32_t wix = __shfl(wx, a, kWave), wiy = __shfl(wy, a, kWave), wiz = __shfl(wz, a, kWave);
                    const int64_t pos = (int64_t)shfl64((uint64_t)cursor, a) + rank;
                    double hxa = 0, hya = 0, hza = 0;
                    if (WANT_R) {
                        hxa = shfl_d(hx, a);
                        hya = shfl_d(hy, a);
                        hza = shfl_d(hz, a);
                    }
                    const int li = row_index(R, Lh < tot ? k : 0);
                    if (Lh < tot) {
                        int32_t tx, ty, tz;
                        unpack_wind(L.t[li] & ~kZeroShiftBit, tx, ty, tz);
                        const int32_t Sx = tx + wix, Sy = ty + wiy, Sz = tz + wiz;
                        i_out[pos] = (TI)(hj + 1);
                        j_out[pos] = (TI)(L.j[li] + 1);
                        TI *S = S_out + 3 * pos;
                        S[0] = (TI)Sx;
                        S[1] = (TI)Sy;
                        S[2] = (TI)Sz;
                        if (WANT_R) {   // the count pass's expression, bit for bit
                            double cx, cy, cz;
                            cell_shift(Cm, Sx, Sy, Sz, cx, cy, cz);
                            double *Rp = R_out + 3 * pos;
                            Rp[0] = (L.x[li] - hxa) + cx;
                            Rp[1] = (L.y[li] - hya) + cy;
                            Rp[2] = (L.z[li] - hza) + cz;
                        }
                    }
                }
                cursor += pc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_count(const PipelineBufs &b, hipStream_t st)
{
    if (b.nat == 0 || b.nbricks == 0) return;
    hipLaunchKernelGGL(count_kernel, dim3((unsigned)b.nbricks), dim3(kBrickThreads), 0, st, b.sys, b.sys_cell_base,
                       b.sys_brick_base, b.nsys, b.M, b.nat, b.cell_start, b.rec, b.wbig, b.pair_count, b.masks);
}

template <bool R, typename TI>
static void fill_launch(const PipelineBufs &b, hipStream_t st)
{
    hipLaunchKernelGGL((fill_kernel<R, TI>), dim3((unsigned)b.nbricks), dim3(kBrickThreads), 0, st, b.sys,
                       b.sys_cell_base, b.sys_brick_base, b.nsys, b.M, b.nat, b.cell_start, b.rec, b.wbig, b.masks,
                       b.pair_off, b.status, (TI *)b.i_out, (TI *)b.j_out, (TI *)b.S_out, b.R_out);
}

void launch_fill(const PipelineBufs &b, bool want_R, hipStream_t st)
{
    if (b.nat == 0 || b.nbricks == 0) return;
    if (b.ti_bytes == 8) {
        if (want_R) fill_launch<true, int64_t>(b, st);
        else fill_launch<false, int64_t>(b, st);
    } else {
        if (want_R) fill_launch<true, int32_t>(b, st);
        else fill_launch<false, int32_t>(b, st);
    }
}

}  // namespace nlh
