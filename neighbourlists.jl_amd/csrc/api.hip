// C ABI of libnlhip.so: handles, device buffers, host geometry and the
// orchestration of the kernel pipeline (kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nlhip.h"
#include "../../include/nlhip_device.h"
#include "nlhip_internal.h"

namespace {

thread_local std::string g_last_error;
thread_local int g_device = 0;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? NLH_ENOMEM : NLH_EDEVICE;
        throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define HIPC(x) hip_check((x), #x)

// Device-wide calls (allocation, free, device or null-stream synchronisation,
// stream and event creation) are not allowed while any stream of the process
// is being captured into a graph.  They take this lock shared; a capture
// (a handful of kernel launches) takes it exclusively.
std::shared_mutex g_capture_mutex;
#define HIPD(x)                                                        \
    do {                                                               \
        std::shared_lock<std::shared_mutex> lk_(g_capture_mutex);      \
        hip_check((x), #x);                                            \
    } while (0)
template <typename F>
void device_wide(F &&f)
{
    std::shared_lock<std::shared_mutex> lk(g_capture_mutex);
    f();
}

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return NLH_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return NLH_ENOMEM;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return NLH_EDEVICE;
    }
}

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int grows = 0;   // reallocations after the first allocation
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
    // returns true when (re)allocated
    // `st` is the owning handle's stream, the only stream that uses the buffer.
    // Blocks come from the device's stream-ordered pool: the old block is
    // freed after the owner's work on it and the new one is allocated in the
    // same stream, so a growth waits for this stream only -- a plain free
    // would wait for every stream on the device (with several handles
    // building at once, e.g. the serving workers, each growth stalled all of
    // them).
    bool ensure(hipStream_t st, size_t bytes, bool zero = false)
    {
        if (bytes == 0) bytes = 16;
        if (bytes <= cap) return false;
        // a quarter of headroom at first; a buffer that has to grow again
        // serves builds of varying size (a serving worker, a recycled
        // resident), and doubles from then on
        if (p) ++grows;
        const size_t nb = std::max(grows ? 2 * bytes : bytes + bytes / 4, (size_t)256);
        device_wide([&] {
            if (p) HIPD(hipFreeAsync(p, st));
            p = nullptr;
            cap = 0;
            HIPD(hipMallocAsync(&p, nb, st));
        });
        cap = nb;
        if (zero) HIPC(hipMemsetAsync(p, 0, nb, st));
        HIPC(hipStreamSynchronize(st));   // the block is ready for the host and any stream
        return true;
    }
    void release(hipStream_t st)
    {
        if (p) device_wide([&] {
            (void)hipFreeAsync(p, st);
            (void)hipStreamSynchronize(st);
        });
        p = nullptr;
        cap = 0;
        grows = 0;
    }
};

}  // namespace

struct nlh_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf X, tables, key, slot, cell_count, bmat, recb, cell_start, tmp, rec, wbig, perm, cell_id,
        cell_offsets, pair_count, slots, slotsR, pair_off, tiles, lookback, sysmap, sys_pair_base, first, i, j, S, R, status, Xsorted, nb_j, nb_R,
        nb_S, nb_count, nr_count, nr_off, nr_j, nr_R, nr_S;
    // lazy queries over a range of atoms (nlh_neighbours_range): the last
    // range's results stay on the device, keyed by the build they came from
    uint64_t gen = 0;                     // builds on this handle so far
    bool nr_valid = false;
    uint64_t nr_gen = 0;
    int32_t nr_sys = 0;
    int64_t nr_first = 0, nr_n = 0;
    std::vector<int64_t> nr_off_host;     // n + 1 offsets of the last range
    // system tables, one device buffer: SysDev[nsys], atom bases int64[nsys + 1],
    // cell bases int32[nsys + 1], brick bases int32[nsys + 1]
    nlh::SysDev *sys_d = nullptr;
    int64_t *atom_base_d = nullptr;
    int32_t *cell_base_d = nullptr, *brick_base_d = nullptr;
    // fused build of small systems: units of consecutive systems (first
    // system of each, then the end) and the work items in launch order
    // (unit, then part | parts << 16), in the tables
    std::vector<int32_t> units_host, order_host;
    int32_t nunits = 0, nitems = 0;
    int32_t *units_d = nullptr, *order_d = nullptr;
    bool small_ok = true;   // cleared when a fused build met a winding beyond +-511
    std::vector<char> tables_host;        // bytes of the last upload
    nlh::Status *status_host = nullptr;   // pinned, mapped: the fill kernel writes it
    nlh::Status *status_host_dev = nullptr;   // its device address
    bool status_pub_pending = false;      // the last launched fill publishes the status
    bool inflight = false;                // a replay returned while its fill still writes pairs
    bool offsets_stale = false;           // pair_offsets_host is fetched when first asked for
    bool graph_publishes = false;         // the captured pipeline ends with a fill
    void *staging = nullptr;              // pinned upload staging for system tables
    size_t staging_cap = 0;
    void *xstage = nullptr;               // pinned upload staging for host positions
    size_t xstage_cap = 0;
    hipEvent_t xstage_ev = nullptr;       // recorded after the last upload out of xstage
    bool xstage_pending = false;
    // state of the last build
    int nsys = 0;
    int64_t nat = 0;
    int32_t M = 0;
    int ti_bytes = 4;
    bool built = false, materialized = false, have_R = false;
    int64_t npairs = 0;
    int64_t pair_capacity = 0;
    bool pair_lookback_clean = false;   // the pair scan's descriptors are zero
    int32_t slot_cap = 0;       // neighbour slots per atom
    bool wide = false;          // 16-byte slot entries (cell shifts beyond +-511)
    const double *X_dev = nullptr;
    std::vector<nlh::SysDev> sys_host;
    std::vector<int64_t> atom_base_host;
    std::vector<int32_t> cell_base_host;
    std::vector<int32_t> brick_base_host;
    int64_t nbricks = 0;
    std::vector<nlh_geometry> geoms;
    std::vector<int64_t> pair_offsets_host;
    // per-phase timing (nlh_set_profiling): events on the handle's stream
    bool profile = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float phase_ms[5] = {0, 0, 0, 0, 0};
    bool ev_build = false, ev_pairs = false;
    // captured batch pipeline (nlh_batch_rebuild) and the buffers it was captured on
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    bool graph_want_R = false;
    bool graph_wide = false;             // slot-entry width the captured kernels were built for
    nlh::PipelineBufs graph_bufs;
    bool last_single = false;            // the previous call was a full single build ...
    uint32_t last_flags = 0;             // ... with these flags
    nlh::PipelineBufs last_bufs;         // ... on these buffers and sizes
    int64_t graph_launches = 0;
    // launch configuration (nlh_set_tuning / nlh_autotune)
    nlh_tuning tune = {920, 256, 256, 0, 0, 0};
    int cus = 256;   // compute units of the device

    // count-pass workgroup: 4 waves per brick unless tuned (four 39 KB
    // bricks share a CU; 8-wave bricks of 76 KB fit two)
    int count_block() const { return tune.count_block ? tune.count_block : 256; }

    nlh::PipelineBufs bufs() const
    {
        nlh::PipelineBufs b;
        std::memset(&b, 0, sizeof b);   // padding too: captured configurations are compared bytewise
        b.X = X_dev;
        b.sys = sys_d;
        b.sys_atom_base = atom_base_d;
        b.sys_cell_base = cell_base_d;
        b.sys_brick_base = brick_base_d;
        b.nbricks = nbricks;
        b.nsys = nsys;
        b.nat = nat;
        b.M = M;
        b.sort = nlh::sort_plan(M, nat, tune.cell_sort);
        b.key = key.as<int32_t>();
        b.slot = slot.as<int32_t>();
        b.cell_count = cell_count.as<int32_t>();
        b.bmat = bmat.as<int32_t>();
        // the bucketed sort's scratch: the slot rows once they are allocated
        // (rewritten by every count pass), else a buffer of its own
        {
            char *sc = slots.cap >= (size_t)std::max<int64_t>(nat, 1) * nlh::kBucketBytes ? (char *)slots.p
                                                                                          : (char *)recb.p;
            b.bke = reinterpret_cast<int2 *>(sc);
        }
        b.cell_start = cell_start.as<int32_t>();
        b.tmp = tmp.as<int32_t>();
        b.rec = rec.as<nlh::Rec>();
        b.wbig = wbig.as<int32_t>();
        b.perm_out = perm.p;
        b.cell_id_out = cell_id.p;
        b.cell_offsets_out = cell_offsets.p;
        b.pair_count = pair_count.as<int32_t>();
        b.slots = slots.p;
        b.slotsR = slotsR.as<double>();
        b.slot_cap = slot_cap;
        b.pair_off = pair_off.as<int64_t>();
        b.tile_atom = tiles.as<int32_t>();
        b.tile_cap = (int64_t)(tiles.cap / 8);
        b.lookback = lookback.as<uint64_t>();
        b.atom_sys = nsys > 1 ? sysmap.as<int32_t>() : nullptr;
        b.sys_pair_base = sys_pair_base.as<int64_t>();
        b.first_out = first.p;
        b.i_out = i.p;
        b.j_out = j.p;
        b.S_out = S.p;
        b.R_out = R.as<double>();
        b.pair_capacity = pair_capacity;
        b.status = status.as<nlh::Status>();
        b.status_pub = status_host_dev;
        b.ti_bytes = ti_bytes;
        b.fill_block = tune.fill_block;
        b.fill_pairs = tune.fill_pairs;
        b.count_block = count_block();
        b.nunits = nunits;
        b.nitems = nitems;
        b.units = units_d;
        b.unit_order = order_d;
        return b;
    }
};

namespace {

// ---------------------------------------------------------------------------
// host geometry: analyze_cell (src/cell_list.jl:152-170) with StaticArrays'
// 3x3 evaluation order; this file is compiled with -ffp-contract=off.
// ---------------------------------------------------------------------------

void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
double dot3(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

void inv3(const double a[9], double out[9])
{
    double x0[3] = {a[0], a[1], a[2]};
    const double x1[3] = {a[3], a[4], a[5]}, x2[3] = {a[6], a[7], a[8]};
    double y0[3], y1[3], y2[3];
    cross3(x1, x2, y0);
    const double d = dot3(x0, y0);
    for (int k = 0; k < 3; ++k) {
        x0[k] = x0[k] / d;
        y0[k] = y0[k] / d;
    }
    cross3(x2, x0, y1);
    cross3(x0, x1, y2);
    for (int c = 0; c < 3; ++c) {
        out[0 + 3 * c] = y0[c];
        out[1 + 3 * c] = y1[c];
        out[2 + 3 * c] = y2[c];
    }
}

double det3(const double a[9])
{
    const double x0[3] = {a[0], a[1], a[2]}, x1[3] = {a[3], a[4], a[5]}, x2[3] = {a[6], a[7], a[8]};
    double y[3];
    cross3(x1, x2, y);
    return dot3(x0, y);
}

// lengths (src/cell_list.jl:94-95): rows of C are the lattice vectors
void lengths3(const double c[9], double out[3])
{
    double r[3][3];
    for (int m = 0; m < 3; ++m)
        for (int k = 0; k < 3; ++k) r[m][k] = c[m + 3 * k];
    const double d = det3(c);
    double t[3];
    auto nrm = [](const double v[3]) { return std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); };
    cross3(r[1], r[2], t);
    out[0] = d / nrm(t);
    cross3(r[2], r[0], t);
    out[1] = d / nrm(t);
    cross3(r[0], r[1], t);
    out[2] = d / nrm(t);
}

// Neighbour-cell reach limits: (2 nxyz + 1)^3 cells are visited per atom.
constexpr int64_t kMaxNxyz = 64;
constexpr int64_t kMaxSegments = (int64_t)1 << 21;

void validate_geometry(const nlh_geometry &g)
{
    if (g.ti_bytes != 4 && g.ti_bytes != 8) throw Error(NLH_EINVAL, "ti_bytes must be 4 (Int32) or 8 (Int64)");
    if (!(g.cutoff > 0.0) || !std::isfinite(g.cutoff)) throw Error(NLH_EINVAL, "cutoff must be positive and finite");
    for (int k = 0; k < 3; ++k) {
        if (g.ncells[k] < 1) throw Error(NLH_EINVAL, "ncells must be >= 1");
        if (g.nxyz[k] < 0 || g.nxyz[k] > kMaxNxyz)
            throw Error(NLH_EINVAL, "neighbour-cell reach (nxyz) out of the supported range [0, 64]");
    }
    if ((2 * g.nxyz[0] + 1) * (2 * g.nxyz[1] + 1) * (2 * g.nxyz[2] + 1) > kMaxSegments)
        throw Error(NLH_EINVAL, "too many neighbour cells per atom (cutoff far larger than the cell)");
    if (g.ncells_total != g.ncells[0] * g.ncells[1] * g.ncells[2]) throw Error(NLH_EINVAL, "ncells_total mismatch");
    const int64_t tmax = g.ti_bytes == 4 ? INT32_MAX : INT64_MAX;
    if (g.ncells_total > tmax)
        throw Error(NLH_EOVERFLOW, "Ratio of simulation cell size to cutoff is very large. Use a larger integer "
                                   "type (e.g. Int64), larger cutoff, or smaller simulation cell.");
    if (g.ncells_total > INT32_MAX - 1)
        throw Error(NLH_EOVERFLOW, "more than 2^31-1 cells per build is not supported on the device");
}

void geometry_init(const double cell[9], double cutoff, const uint8_t pbc[3], int32_t ti_bytes, nlh_geometry &g)
{
    std::memset(&g, 0, sizeof g);
    if (ti_bytes != 4 && ti_bytes != 8) throw Error(NLH_EINVAL, "ti_bytes must be 4 (Int32) or 8 (Int64)");
    if (!(cutoff > 0.0) || !std::isfinite(cutoff)) throw Error(NLH_EINVAL, "cutoff must be positive and finite");
    for (int q = 0; q < 9; ++q)
        if (!std::isfinite(cell[q])) throw Error(NLH_EINVAL, "cell has non-finite entries");
    std::memcpy(g.cell, cell, sizeof g.cell);
    g.cutoff = cutoff;
    g.cutoff_sq = cutoff * cutoff;   // cutoff^2 (gpu_kernels.jl:317)
    g.ti_bytes = ti_bytes;
    for (int k = 0; k < 3; ++k) g.pbc[k] = pbc[k] ? 1 : 0;
    g.zero_volume = std::fabs(det3(cell)) < 1e-12;
    inv3(cell, g.inv_cell);
    double lens[3];
    lengths3(cell, lens);
    const double timax = ti_bytes == 4 ? 2147483647.0 : 9.2233720368547758e18;
    for (int k = 0; k < 3; ++k) {
        g.lens[k] = std::fabs(lens[k]);
        const double t = std::floor(g.lens[k] / cutoff);   // floor.(TI, lens / cutoff)
        if (std::isnan(t)) throw Error(NLH_EINVAL, "degenerate cell (face height is NaN)");
        if (t > timax) throw Error(NLH_EOVERFLOW, "ncells does not fit the index type");
        const int64_t n = (int64_t)t;
        g.ncells[k] = n < 1 ? 1 : n;
    }
    const long double prod = (long double)g.ncells[0] * g.ncells[1] * g.ncells[2];
    const long double tmax = ti_bytes == 4 ? 2147483647.0L : 9223372036854775807.0L;
    if (prod > tmax)
        throw Error(NLH_EOVERFLOW, "Ratio of simulation cell size to cutoff is very large. Use a larger integer "
                                   "type (e.g. Int64), larger cutoff, or smaller simulation cell.");
    g.ncells_total = g.ncells[0] * g.ncells[1] * g.ncells[2];
    for (int k = 0; k < 3; ++k) {
        // ceil.(TI, cutoff * (ncells ./ abs.(lens)))  (gpu_kernels.jl:315-316)
        const double v = std::ceil(cutoff * ((double)g.ncells[k] / g.lens[k]));
        if (!std::isfinite(v) || v > timax) throw Error(NLH_EINVAL, "neighbour-cell reach is not finite (zero-volume cell?)");
        g.nxyz[k] = (int64_t)v;
    }
}

nlh::SysDev make_sys(const nlh_geometry &g, int64_t nat, int64_t atom_base, int32_t cell_base, int64_t first_base)
{
    nlh::SysDev s{};
    std::memcpy(s.C, g.cell, sizeof s.C);
    std::memcpy(s.inv, g.inv_cell, sizeof s.inv);
    s.rc2 = g.cutoff_sq;
    for (int k = 0; k < 3; ++k) {
        s.n[k] = (int32_t)g.ncells[k];
        s.nxyz[k] = (int32_t)g.nxyz[k];
    }
    s.pbc = (g.pbc[0] ? 1 : 0) | (g.pbc[1] ? 2 : 0) | (g.pbc[2] ? 4 : 0);
    s.nat = (int32_t)nat;
    s.atom_base = atom_base;
    s.cell_base = cell_base;
    s.M = (int32_t)g.ncells_total;
    s.first_base = first_base;
    return s;
}

void ensure_staging(nlh_handle *h, size_t bytes)
{
    if (bytes <= h->staging_cap) return;
    HIPC(hipStreamSynchronize(h->stream));
    if (h->staging) HIPD(hipHostFree(h->staging));
    h->staging = nullptr;
    const size_t nb = std::max(bytes * 2, (size_t)4096);
    HIPD(hipHostMalloc(&h->staging, nb, hipHostMallocDefault));
    h->staging_cap = nb;
}

void settle(nlh_handle *h);

void drop_graph(nlh_handle *h)
{
    settle(h);   // an executable graph is destroyed only once its replay is done
    device_wide([&] {
        if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
        if (h->graph) (void)hipGraphDestroy(h->graph);
    });
    h->graph_exec = nullptr;
    h->graph = nullptr;
}

// Pair-search brick of a system: the largest of a fixed ladder of shapes
// whose halo atoms, expected at this system's mean atom density per cell,
// stay within the budget (clipped to the cell grid).
void choose_brick(const nlh_geometry &g, int64_t nat, int32_t budget, int32_t halo_max, int32_t out[3],
                  int64_t min_bricks = 0)
{
    static const int32_t ladder[][3] = {{4, 4, 4}, {4, 4, 2}, {4, 2, 2}, {2, 2, 2}, {2, 2, 1}, {2, 1, 1}, {1, 1, 1}};
    const double per_cell = g.ncells_total > 0 ? (double)nat / (double)g.ncells_total : 0.0;
    bool found = false;
    for (const auto &b : ladder) {
        int32_t d[3];
        double halo = 1.0;
        int64_t nb = 1;
        for (int k = 0; k < 3; ++k) {
            d[k] = (int32_t)std::min<int64_t>(b[k], g.ncells[k]);
            halo *= (double)(d[k] + 2);
            nb *= (g.ncells[k] + d[k] - 1) / d[k];
        }
        if (halo * per_cell <= (double)budget && halo <= (double)halo_max) {
            // the first shape that fits; a smaller one while the grid is too
            // coarse to fill the device, down to one home cell per wave
            std::memcpy(out, d, sizeof d);
            found = true;
            if (nb >= min_bricks || (int64_t)d[0] * d[1] * d[2] <= 8) return;
        }
    }
    if (!found) out[0] = out[1] = out[2] = 1;
}

// Brick classes of a system whose brick shape is set: per class its first
// brick in launch order and its shape (see SysDev), largest bricks first, so
// that the bricks that take longest start first and the short ones fill the
// end of the launch.  Returns the system's bricks.
int64_t brick_classes(nlh::SysDev &sd)
{
    struct Cls {
        int64_t cells, bricks;
        int32_t mask;
    };
    Cls cl[8];
    int ncl = 0;
    for (int k = 0; k < 3; ++k) sd.bfull[k] = sd.n[k] / sd.brick[k];
    for (int32_t m = 0; m < 8; ++m) {
        int64_t cells = 1, bricks = 1;
        for (int k = 0; k < 3; ++k) {
            const int32_t rem = sd.n[k] - sd.bfull[k] * sd.brick[k];
            cells *= (m >> k & 1) ? rem : sd.brick[k];
            bricks *= (m >> k & 1) ? (rem > 0 ? 1 : 0) : sd.bfull[k];
        }
        if (bricks > 0) cl[ncl++] = {cells, bricks, m};
    }
    std::stable_sort(cl, cl + ncl, [](const Cls &a, const Cls &b) { return a.cells > b.cells; });
    int64_t first = 0;
    for (int c = 0; c < 8; ++c) {
        if (first > INT32_MAX - 1) throw Error(NLH_EOVERFLOW, "too many pair-search bricks");
        sd.cls_first[c] = (int32_t)first;
        sd.cls_mask[c] = c < ncl ? cl[c].mask : 0;
        if (c < ncl) first += cl[c].bricks;
    }
    if (first > INT32_MAX - 1) throw Error(NLH_EOVERFLOW, "too many pair-search bricks");
    sd.cls_first[8] = (int32_t)first;
    return first;
}

// Repeated single builds replay the captured pipeline below this size.
constexpr int64_t kReplayMaxAtoms = 12288;

// A brick's LDS holds 4 BT halo atoms; the expected halo may fill 90% of it
// (the rest absorbs the density's fluctuation: a brick that overflows is
// counted from global memory instead).
constexpr double kHaloFill = 0.9;

// Pair-search bricks of every system for the current brick budget: brick
// shape per system and the first brick of each system (host tables; the
// caller uploads them).
void assign_bricks(nlh_handle *h)
{
    int64_t bricks = 0;
    const int nsys = (int)h->sys_host.size();
    h->brick_base_host.resize(nsys + 1);
    for (int s = 0; s < nsys; ++s) {
        nlh::SysDev &sd = h->sys_host[s];
        // a single system takes the largest brick that still fills the
        // device once: four 256-thread bricks (39 KB of LDS each) or two
        // 512-thread bricks fit a CU, and the largest-first launch order
        // fills the end of the launch with the short partial bricks
        // the LDS of a BT-thread brick holds 4 BT halo atoms and BT halo cells
        const int bt = h->count_block();
        const int64_t min_bricks = nsys == 1 ? (int64_t)h->cus * (bt == 256 ? 4 : 2) : 0;
        choose_brick(h->geoms[s], sd.nat, std::min(h->tune.brick_atoms, (int32_t)(kHaloFill * 4 * bt)), bt,
                     sd.brick, min_bricks);
        sd.brick_base = (int32_t)bricks;
        h->brick_base_host[s] = (int32_t)bricks;
        bricks += brick_classes(sd);
        if (bricks > INT32_MAX - 1) throw Error(NLH_EOVERFLOW, "too many pair-search bricks in one batch");
    }
    h->brick_base_host[nsys] = (int32_t)bricks;
    h->nbricks = bricks;
}

// Fused build of small systems: a batch of at least kSmallMinSystems systems,
// each within a unit's limits (atoms, cells, neighbour-cell reach), is cut
// into units of consecutive systems, at most kUnitAtoms atoms, kUnitCells
// cells and kUnitSystems systems each.  A unit whose estimated count work is
// well above an even share of the batch is split into parts (each sorts the
// whole unit and counts every parts-th home cell), and the work items are
// launched largest first.  Otherwise (or when turned off, or after a winding
// beyond +-511 made a fused build fall back) no units: the general pipeline.
constexpr int kSmallMinSystems = 64;
constexpr int kMaxUnitParts = 16;       // parts of one unit at most (each at least 4 home cells)
constexpr double kItemRounds = 2.0;     // work items per CU the split aims at

// Estimated count work of a system in distance tests of 64 candidates: every
// atom against the atoms of its (2 nxyz + 1)^3 neighbour cells, the cells
// holding the system's mean number of atoms.
double small_count_work(const nlh::SysDev &sd)
{
    if (sd.nat <= 0) return 0.0;
    const double nseg = (2.0 * sd.nxyz[0] + 1) * (2.0 * sd.nxyz[1] + 1) * (2.0 * sd.nxyz[2] + 1);
    return (double)sd.nat * std::ceil(nseg * sd.nat / std::max(sd.M, 1) / 64.0);
}

void plan_units(nlh_handle *h)
{
    h->units_host.clear();
    h->order_host.clear();
    h->nunits = 0;
    h->nitems = 0;
    const int nsys = (int)h->sys_host.size();
    if (!h->small_ok || h->wide || h->tune.fused_small != 0 || nsys < kSmallMinSystems) return;
    int64_t total = 0;
    for (const nlh::SysDev &sd : h->sys_host) {
        total += sd.nat;
        if (sd.nat > nlh::kUnitAtoms || sd.M > nlh::kUnitCells) return;
        for (int k = 0; k < 3; ++k)
            if (sd.nxyz[k] > nlh::kUnitMaxReach || sd.n[k] > nlh::kUnitMaxAxisCells) return;
    }
    if (total == 0) return;   // no count stage follows: the general build writes the cell list
    std::vector<double> work;
    std::vector<int64_t> atoms_u, cells_u;
    int s = 0;
    while (s < nsys) {
        const int first = s;
        int64_t atoms = 0, cells = 0;
        double w = 0.0;
        while (s < nsys && s - first < nlh::kUnitSystems && atoms + h->sys_host[s].nat <= nlh::kUnitAtoms &&
               cells + h->sys_host[s].M <= nlh::kUnitCells) {
            atoms += h->sys_host[s].nat;
            cells += h->sys_host[s].M;
            w += small_count_work(h->sys_host[s]);
            ++s;
        }
        h->units_host.push_back(first);
        work.push_back(w);
        atoms_u.push_back(atoms);
        cells_u.push_back(cells);
    }
    h->units_host.push_back(nsys);
    h->nunits = (int32_t)work.size();
    // A workgroup's time, fitted on MI355X (us): 14.4 + 0.0115 per atom of
    // the unit (its sort, done by every part) + 0.0072 per estimated test +
    // 0.154 per cell (the last two shared by the parts).
    std::vector<double> fixed(h->nunits), shared(h->nunits);
    double time_total = 0.0;
    for (int u = 0; u < h->nunits; ++u) {
        fixed[u] = 14.4 + 0.0115 * atoms_u[u];
        shared[u] = 0.0072 * work[u] + 0.154 * cells_u[u];
        time_total += fixed[u] + shared[u];
    }
    const double share = std::max(1.0, time_total / (std::max(h->cus, 1) * kItemRounds));
    struct Item {
        double key;
        int32_t u, part, parts;
    };
    std::vector<Item> items;
    for (int u = 0; u < h->nunits; ++u) {
        const int cap = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxUnitParts, cells_u[u] / 4));
        const int parts = (int)std::max<double>(1.0, std::min<double>(cap, std::llround(shared[u] / share)));
        const double key = fixed[u] + shared[u] / parts;
        for (int p = 0; p < parts; ++p) items.push_back({key, u, p, parts});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.key > b.key; });
    h->nitems = (int32_t)items.size();
    h->order_host.resize(2 * items.size());
    for (size_t t = 0; t < items.size(); ++t) {
        h->order_host[2 * t] = items[t].u;
        h->order_host[2 * t + 1] = items[t].part | items[t].parts << 16;
    }
}

// Upload the system tables (sys, atom/cell/brick bases) as one copy through
// pinned staging, unless the device already holds exactly these bytes (a
// rebuild of the same systems).  Waits for earlier work on the stream before
// reusing the staging buffer.
void upload_tables(nlh_handle *h)
{
    const int nsys = (int)h->sys_host.size();
    const size_t sys_bytes = nsys * sizeof(nlh::SysDev);
    const size_t ab_bytes = (nsys + 1) * sizeof(int64_t);
    const size_t cb_bytes = (nsys + 1) * sizeof(int32_t);
    const size_t un_bytes = h->units_host.size() * sizeof(int32_t), or_bytes = h->order_host.size() * sizeof(int32_t);
    const size_t o_ab = (sys_bytes + 15) / 16 * 16, o_cb = o_ab + (ab_bytes + 15) / 16 * 16,
                 o_bb = o_cb + (cb_bytes + 15) / 16 * 16, o_un = o_bb + (cb_bytes + 15) / 16 * 16,
                 o_or = o_un + (un_bytes + 15) / 16 * 16, total = o_or + or_bytes;
    std::vector<char> img(total, 0);
    std::memcpy(img.data(), h->sys_host.data(), sys_bytes);
    std::memcpy(img.data() + o_ab, h->atom_base_host.data(), ab_bytes);
    std::memcpy(img.data() + o_cb, h->cell_base_host.data(), cb_bytes);
    std::memcpy(img.data() + o_bb, h->brick_base_host.data(), cb_bytes);
    if (un_bytes) std::memcpy(img.data() + o_un, h->units_host.data(), un_bytes);
    if (or_bytes) std::memcpy(img.data() + o_or, h->order_host.data(), or_bytes);
    const bool moved = h->tables.ensure(h->stream, total);
    char *base = static_cast<char *>(h->tables.p);
    h->sys_d = reinterpret_cast<nlh::SysDev *>(base);
    h->atom_base_d = reinterpret_cast<int64_t *>(base + o_ab);
    h->cell_base_d = reinterpret_cast<int32_t *>(base + o_cb);
    h->brick_base_d = reinterpret_cast<int32_t *>(base + o_bb);
    h->units_d = h->nunits ? reinterpret_cast<int32_t *>(base + o_un) : nullptr;
    h->order_d = h->nunits ? reinterpret_cast<int32_t *>(base + o_or) : nullptr;
    if (!moved && img == h->tables_host) return;
    HIPC(hipStreamSynchronize(h->stream));
    ensure_staging(h, total);
    std::memcpy(h->staging, img.data(), total);
    HIPC(hipMemcpyAsync(h->tables.p, h->staging, total, hipMemcpyHostToDevice, h->stream));
    h->tables_host.swap(img);
}

// Set up the system tables of a batch (a single build is nsys == 1) and
// size every build-side buffer.  Uploads go through pinned staging.
void setup_batch(nlh_handle *h, int32_t nsys, const int64_t *atom_offsets, const nlh_geometry *geoms)
{
    if (nsys < 1) throw Error(NLH_EINVAL, "nsys must be >= 1");
    const int ti = geoms[0].ti_bytes;
    h->sys_host.resize(nsys);
    h->atom_base_host.resize(nsys + 1);
    h->cell_base_host.resize(nsys + 1);
    h->brick_base_host.resize(nsys + 1);
    h->geoms.assign(geoms, geoms + nsys);
    int64_t cells = 0;
    if (atom_offsets[0] != 0) throw Error(NLH_EINVAL, "atom_offsets[0] must be 0");
    for (int s = 0; s < nsys; ++s) {
        validate_geometry(geoms[s]);
        if (geoms[s].ti_bytes != ti) throw Error(NLH_EINVAL, "all systems of a batch must share ti_bytes");
        const int64_t na = atom_offsets[s + 1] - atom_offsets[s];
        if (na < 0) throw Error(NLH_EINVAL, "atom_offsets must be non-decreasing");
        if (ti == 4 && na >= INT32_MAX) throw Error(NLH_EOVERFLOW, "too many atoms for Int32 indices");
        h->sys_host[s] = make_sys(geoms[s], na, atom_offsets[s], (int32_t)cells, atom_offsets[s] + s);
        h->atom_base_host[s] = atom_offsets[s];
        h->cell_base_host[s] = (int32_t)cells;
        cells += geoms[s].ncells_total;
        if (cells > INT32_MAX - 1) throw Error(NLH_EOVERFLOW, "more than 2^31-1 cells in one batch");
    }
    const int64_t nat = atom_offsets[nsys];
    if (nat > INT32_MAX - 1) throw Error(NLH_EOVERFLOW, "more than 2^31-1 atoms in one build");
    h->atom_base_host[nsys] = nat;
    h->cell_base_host[nsys] = (int32_t)cells;
    h->nsys = nsys;   // before the bricks: the brick workgroup size depends on it
    assign_bricks(h);
    plan_units(h);
    h->nat = nat;
    h->M = (int32_t)cells;
    h->ti_bytes = ti;

    const size_t n1 = (size_t)std::max<int64_t>(nat, 1);
    const size_t M1 = (size_t)cells + 1;
    h->key.ensure(h->stream, n1 * 4);
    h->slot.ensure(h->stream, n1 * 4);
    h->cell_count.ensure(h->stream, M1 * 4, /*zero=*/true);
    h->cell_start.ensure(h->stream, M1 * 4);
    h->tmp.ensure(h->stream, n1 * 4);
    // the bucketed sort's buffers whatever the tuning says (it may change
    // before the next build)
    const nlh::SortPlan sp = nlh::sort_plan(cells, nat, 0);
    if (sp.bw >= 0) {
        h->bmat.ensure(h->stream, (size_t)sp.n_sort * 4);
        if (h->slots.cap < n1 * nlh::kBucketBytes) h->recb.ensure(h->stream, n1 * nlh::kBucketBytes);
    }
    h->rec.ensure(h->stream, n1 * sizeof(nlh::Rec));
    h->wbig.ensure(h->stream, n1 * 12);
    h->perm.ensure(h->stream, n1 * ti);
    h->cell_id.ensure(h->stream, n1 * ti);
    h->cell_offsets.ensure(h->stream, (M1 + nsys) * ti);
    h->lookback.ensure(h->stream,
                       (size_t)nlh::lookback_entries(std::max<int64_t>((int64_t)M1, sp.n_sort), (int64_t)n1) *
                           sizeof(uint64_t));
    if (nsys > 1) h->sysmap.ensure(h->stream, (n1 + M1) * 4);   // system of each atom, then of each cell
    h->status.ensure(h->stream, sizeof(nlh::Status));

    upload_tables(h);
}

// Host positions go through the handle's pinned staging buffer: a copy
// from pageable memory is staged by the runtime itself, and such copies from
// several threads at once were measured to wait on each other for
// milliseconds.  The buffer is reused once the previous upload out of it is
// done (its event).
void stage_positions(nlh_handle *h, const double *X, size_t bytes)
{
    if (h->xstage_pending) {
        HIPC(hipEventSynchronize(h->xstage_ev));
        h->xstage_pending = false;
    }
    if (bytes > h->xstage_cap) {
        if (h->xstage) HIPD(hipHostFree(h->xstage));
        h->xstage = nullptr;
        h->xstage_cap = 0;
        const size_t nb = std::max(bytes + bytes / 4, (size_t)4096);
        HIPD(hipHostMalloc(&h->xstage, nb, hipHostMallocDefault));
        h->xstage_cap = nb;
    }
    if (!h->xstage_ev) HIPD(hipEventCreateWithFlags(&h->xstage_ev, hipEventDisableTiming));
    std::memcpy(h->xstage, X, bytes);
    HIPC(hipMemcpyAsync(h->X.p, h->xstage, bytes, hipMemcpyHostToDevice, h->stream));
    HIPC(hipEventRecord(h->xstage_ev, h->stream));
    h->xstage_pending = true;
}

void upload_positions(nlh_handle *h, const double *X, uint32_t flags)
{
    if (h->nat > 0 && X == nullptr) throw Error(NLH_EINVAL, "X is NULL");
    if (flags & NLH_X_DEVICE) {
        h->X_dev = X;
    } else {
        h->X.ensure(h->stream, (size_t)std::max<int64_t>(h->nat, 1) * 24);
        if (h->nat > 0) stage_positions(h, X, (size_t)h->nat * 24);
        h->X_dev = h->X.as<double>();
    }
}

// Clears the status and every scan descriptor.
void reset_status(nlh_handle *h)
{
    nlh::launch_clear_status(h->bufs(), h->stream);
    h->pair_lookback_clean = true;
    h->status_pub_pending = false;
}

// The pair scan, on descriptors that start from zero: cleared with the
// status, or here when the scan runs again without a status clear (a second
// materialisation, or the repeat after growing the outputs).
void pair_scan(nlh_handle *h, const nlh::PipelineBufs &b)
{
    if (!h->pair_lookback_clean) nlh::launch_clear_pair_lookback(b, h->stream);
    nlh::launch_scan_pairs(b, h->stream);
    h->pair_lookback_clean = false;
}

void mark(nlh_handle *h, int k)
{
    if (h->profile) HIPC(hipEventRecord(h->ev[k], h->stream));
}

void build_cells(nlh_handle *h)
{
    ++h->gen;
    mark(h, 0);
    nlh::launch_build(h->bufs(), h->stream);
    mark(h, 1);
    h->ev_build = h->profile;
    h->ev_pairs = false;
    HIPC(hipGetLastError());
    h->built = true;
    h->materialized = false;
}

int64_t estimate_pairs(const nlh_handle *h)
{
    // rough density estimate; a wrong guess costs one extra scan + fill
    long double est = 0;
    for (int s = 0; s < h->nsys; ++s) {
        const nlh_geometry &g = h->geoms[s];
        const double vol = std::fabs(det3(g.cell));
        const double nat = (double)h->sys_host[s].nat;
        if (nat <= 0) continue;
        const double rho = vol > 0 ? nat / vol : 0.0;
        const double ball = 4.18879020478639 * g.cutoff * g.cutoff * g.cutoff;
        est += (long double)nat * (rho * ball * 1.25 + 4.0);
    }
    return (int64_t)std::min<long double>(est, 4.0e11L) + 1024;
}

// Slot rows hold a multiple of 16 entries, so every row starts on a 128-byte
// line (8-byte entries) and a row of up to 32 4-byte entries fills one line.
int32_t round_slot_cap(double c)
{
    return (int32_t)((c + 15) / 16) * 16;
}

// Expected neighbours per atom, largest over the batch (a slot-size guess;
// an underestimate costs one repeat of the count and fill passes).
int32_t estimate_slot_cap(const nlh_handle *h)
{
    double mx = 0;
    for (int s = 0; s < h->nsys; ++s) {
        const nlh_geometry &g = h->geoms[s];
        const double vol = std::fabs(det3(g.cell));
        const double nat = (double)h->sys_host[s].nat;
        if (nat <= 0 || !(vol > 0)) continue;
        mx = std::max(mx, nat / vol * 4.18879020478639 * g.cutoff * g.cutoff * g.cutoff);
    }
    const double c = std::min(mx * 1.5 + 16.0, 1.0e6);
    return round_slot_cap(c);
}

void ensure_slots(nlh_handle *h, int32_t cap, bool want_R)
{
    h->slot_cap = std::max(h->slot_cap, cap);
    const size_t n1 = (size_t)std::max<int64_t>(h->nat, 1) * (size_t)h->slot_cap;
    h->slots.ensure(h->stream, n1 * (h->wide ? 16 : 8));
    if (want_R) h->slotsR.ensure(h->stream, n1 * 24);
}

void ensure_pair_outputs(nlh_handle *h, int64_t cap, bool want_R)
{
    const int ti = h->ti_bytes;
    h->i.ensure(h->stream, (size_t)cap * ti);
    h->j.ensure(h->stream, (size_t)cap * ti);
    h->S.ensure(h->stream, (size_t)cap * 3 * ti);
    if (want_R) h->R.ensure(h->stream, (size_t)cap * 24);
    int64_t c = (int64_t)(h->i.cap / ti);
    c = std::min<int64_t>(c, (int64_t)(h->j.cap / ti));
    c = std::min<int64_t>(c, (int64_t)(h->S.cap / (3 * ti)));
    if (want_R) c = std::min<int64_t>(c, (int64_t)(h->R.cap / 24));
    h->pair_capacity = c;
    // fill tiles over the capacity at the smallest tile size (128 pairs)
    h->tiles.ensure(h->stream, (size_t)(c / nlh::kMinFillPairs + 2) * 8);
}

void ensure_pair_buffers(nlh_handle *h)
{
    const size_t n1 = (size_t)std::max<int64_t>(h->nat, 1);
    h->pair_count.ensure(h->stream, n1 * 4);
    h->pair_off.ensure(h->stream, (n1 + 1) * 8);
    h->first.ensure(h->stream, (n1 + h->nsys) * h->ti_bytes);
    h->sys_pair_base.ensure(h->stream, (h->nsys + 1) * 8);
}

void finish_status(nlh_handle *h)
{
    // a fill publishes the status itself (the pair stages end with it);
    // otherwise it is copied back
    if (!h->status_pub_pending)
        HIPC(hipMemcpyAsync(h->status_host, h->status.p, sizeof(nlh::Status), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->status_pub_pending = false;
    h->inflight = false;
    if (h->status_host->scan_error) {
        h->built = false;
        throw Error(NLH_EDEVICE, "internal: a scan block waited too long for its predecessors");
    }
}

// The status of a replayed pipeline, read as soon as its fill has published
// it: the fill does so as it starts, so the host goes on (checks, the
// caller's next request, the next launch) while the fill writes the pairs,
// and the pipeline's last stage overlaps the host's turnaround.  The handle
// is then `inflight`: settle() waits for the fill before anything hands out
// or replaces what it writes.  Stops early when the stream has drained (the
// flag must then be set) or reports an error.
void await_published(nlh_handle *h)
{
    const int32_t *pub = &h->status_host->published;
    for (unsigned spin = 1;; ++spin) {
        if (__atomic_load_n(pub, __ATOMIC_ACQUIRE)) {
            h->inflight = true;
            break;
        }
        if ((spin & 255) == 0) {
            const hipError_t q = hipStreamQuery(h->stream);
            if (q == hipSuccess) {
                if (!__atomic_load_n(pub, __ATOMIC_ACQUIRE))
                    throw Error(NLH_EDEVICE, "internal: the pipeline ended without publishing its status");
                break;
            }
            if (q != hipErrorNotReady) HIPC(q);
        }
        __builtin_ia32_pause();
    }
    h->status_pub_pending = false;
    if (h->status_host->scan_error) {
        h->built = false;
        settle(h);
        throw Error(NLH_EDEVICE, "internal: a scan block waited too long for its predecessors");
    }
}

// Wait for a fill that was still writing when its call returned.
void settle(nlh_handle *h)
{
    if (!h->inflight) return;
    HIPC(hipStreamSynchronize(h->stream));
    h->inflight = false;
}

// The status of the stages just launched: early from the fill when it
// publishes (not while phases are being timed: their events need the whole
// stream), else after the stream.
void wait_status(nlh_handle *h)
{
    if (h->status_pub_pending && !h->profile)
        await_published(h);
    else
        finish_status(h);
}

void check_positions(nlh_handle *h)
{
    if (h->status_host->bad_position) {
        h->built = false;
        throw Error(NLH_EINVAL, "positions must be finite and within 1e9 cells of the box");
    }
}

void finish_pairs(nlh_handle *h, bool want_R);

// Host-side bounds check of everything the count and fill kernels write:
// slots of slot_cap entries per atom (and their R), pair outputs of
// pair_capacity entries (the fill skips itself beyond that), per-atom
// counts and offsets.
void check_pair_buffers(const nlh_handle *h, bool want_R)
{
    const size_t n1 = (size_t)std::max<int64_t>(h->nat, 1);
    const size_t ent = (size_t)h->slot_cap * n1;
    const size_t ti = (size_t)h->ti_bytes, cap = (size_t)std::max<int64_t>(h->pair_capacity, 0);
    bool ok = h->slot_cap > 0 && h->slots.p && h->slots.cap >= ent * (h->wide ? 16 : 8);
    ok = ok && h->pair_count.cap >= n1 * 4 && h->pair_off.cap >= (n1 + 1) * 8;
    ok = ok && h->first.cap >= (n1 + h->nsys) * ti && h->sys_pair_base.cap >= (size_t)(h->nsys + 1) * 8;
    ok = ok && h->i.cap >= cap * ti && h->j.cap >= cap * ti && h->S.cap >= cap * 3 * ti;
    ok = ok && h->tiles.cap >= (cap / nlh::kMinFillPairs + 2) * 8;
    if (want_R) ok = ok && h->slotsR.p && h->slotsR.cap >= ent * 24 && h->R.p && h->R.cap >= cap * 24;
    if (!ok) throw Error(NLH_EDEVICE, "internal: pair-stage buffers are smaller than the launch needs");
}

// Pair stages.  Returns after the wait, with npairs known.
void materialize(nlh_handle *h, uint32_t flags)
{
    if (!h->built) throw Error(NLH_ESTATE, "materialize called before a successful build_cell_list");
    const bool want_R = (flags & NLH_WANT_R) != 0;
    ensure_pair_buffers(h);
    if (h->nat == 0) {   // gpu_kernels.jl:303-312: first = [1]
        finish_status(h);
        check_positions(h);
        h->npairs = 0;
        if (h->nsys == 1) {
            const int64_t one = 1;
            if (h->ti_bytes == 8) HIPD(hipMemcpy(h->first.p, &one, 8, hipMemcpyHostToDevice));
            else {
                const int32_t o = 1;
                HIPD(hipMemcpy(h->first.p, &o, 4, hipMemcpyHostToDevice));
            }
        }
        h->materialized = true;
        h->have_R = want_R;
        h->pair_offsets_host.assign(h->nsys + 1, 0);
        h->offsets_stale = false;
        return;
    }
    ensure_pair_outputs(h, std::max(h->pair_capacity, estimate_pairs(h)), want_R);
    ensure_slots(h, std::max(h->slot_cap, estimate_slot_cap(h)), want_R);
    check_pair_buffers(h, want_R);
    nlh::PipelineBufs b = h->bufs();
    mark(h, 1);
    nlh::launch_count(b, want_R, h->wide, h->stream);
    mark(h, 2);
    pair_scan(h, b);
    mark(h, 3);
    h->status_host->published = 0;   // any earlier fill has published (or ended)
    h->status_pub_pending = nlh::launch_fill(b, want_R, h->wide, h->stream);
    mark(h, 4);
    h->ev_pairs = h->profile;
    HIPC(hipGetLastError());
    wait_status(h);
    check_positions(h);
    if (h->nunits > 0 && (h->status_host->small_fallback || h->status_host->wide)) {
        // a winding beyond +-511 in a fused build: the general pipeline redoes it
        settle(h);
        h->small_ok = false;
        plan_units(h);
        upload_tables(h);
        reset_status(h);
        build_cells(h);
        materialize(h, flags);
        return;
    }
    const nlh::Status &st = *h->status_host;
    if (st.npairs > h->pair_capacity || st.max_count > h->slot_cap || (st.wide && !h->wide)) {
        // grow the outputs / slots (or widen the slot entries) and repeat count + fill
        // (the counts, offsets and totals of the first pass stay valid)
        if (st.wide) h->wide = true;
        ensure_pair_outputs(h, std::max<int64_t>(st.npairs, 1), want_R);
        h->slots.release(h->stream);
        ensure_slots(h, round_slot_cap(std::max(st.max_count, 1)), want_R);
        check_pair_buffers(h, want_R);
        b = h->bufs();
        nlh::launch_count(b, want_R, h->wide, h->stream);
        pair_scan(h, b);   // fill tiles over the grown capacity
        h->status_host->published = 0;
        h->status_pub_pending = nlh::launch_fill(b, want_R, h->wide, h->stream);
        HIPC(hipGetLastError());
        wait_status(h);
        if (h->status_host->npairs > h->pair_capacity || h->status_host->max_count > h->slot_cap)
            throw Error(NLH_EDEVICE, "pair capacity still exceeded after growth");
    }
    finish_pairs(h, want_R);
}

// Bookkeeping after a successful pair pass (status already on the host).
void finish_pairs(nlh_handle *h, bool want_R)
{
    h->npairs = h->status_host->npairs;
    h->have_R = want_R;
    // per-system pair offsets
    h->pair_offsets_host.resize(h->nsys + 1);
    h->offsets_stale = false;
    if (h->nsys == 1) {
        h->pair_offsets_host[0] = 0;
        h->pair_offsets_host[1] = h->npairs;
    } else if (h->inflight && !(h->ti_bytes == 4 && h->npairs >= INT32_MAX)) {
        // a replay still filling: fetched by nlh_batch_pair_offsets once the
        // stream is done (no system can exceed Int32 when the total does not)
        h->offsets_stale = true;
    } else {
        settle(h);
        HIPD(hipMemcpy(h->pair_offsets_host.data(), h->sys_pair_base.p, (h->nsys + 1) * 8, hipMemcpyDeviceToHost));
    }
    if (h->ti_bytes == 4 && !h->offsets_stale)
        for (int s = 0; s < h->nsys; ++s)
            if (h->pair_offsets_host[s + 1] - h->pair_offsets_host[s] >= INT32_MAX)
                throw Error(NLH_EOVERFLOW, "pair count exceeds typemax(Int32); use int_type=Int64");
    h->materialized = true;
}

// The whole rebuild of the current batch -- status clear, binning, sort,
// offsets, count, pair scan, fill -- captured once as a HIP graph on the
// handle's stream.  Buffers, capacities and positions are baked in; the
// capture is redone when any of them changes.  No memset nodes: the status
// is cleared by a kernel.
void capture_graph(nlh_handle *h, bool want_R)
{
    drop_graph(h);
    const nlh::PipelineBufs b = h->bufs();
    hipGraph_t g = nullptr;
    hipError_t le = hipSuccess, ee = hipSuccess;
    {
        // kernel launches only between begin and end (nothing that takes the lock)
        std::unique_lock<std::shared_mutex> lk(g_capture_mutex);
        HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        nlh::launch_clear_status(b, h->stream);
        nlh::launch_build(b, h->stream);
        nlh::launch_count(b, want_R, h->wide, h->stream);
        nlh::launch_scan_pairs(b, h->stream);
        h->graph_publishes = nlh::launch_fill(b, want_R, h->wide, h->stream);
        le = hipGetLastError();
        ee = hipStreamEndCapture(h->stream, &g);
    }
    HIPC(le);
    HIPC(ee);
    h->graph = g;
    HIPD(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
    h->graph_bufs = b;
    h->graph_want_R = want_R;
    h->graph_wide = h->wide;
}

// Rebuild through the captured graph.  Returns false when the pair outputs
// or slots were too small (the fill skipped itself); the caller then runs
// the eager pipeline, which grows them.
bool rebuild_graph(nlh_handle *h, bool want_R)
{
    // size every pair-stage buffer for this request (R may be asked for now
    // and not at the build that sized them)
    ensure_pair_buffers(h);
    ensure_pair_outputs(h, std::max(h->pair_capacity, estimate_pairs(h)), want_R);
    ensure_slots(h, std::max(h->slot_cap, estimate_slot_cap(h)), want_R);
    check_pair_buffers(h, want_R);
    const nlh::PipelineBufs b = h->bufs();
    // the slot-entry width is baked into the captured count and fill: a build
    // that widened the entries (a shift beyond +-511) recaptures even when
    // the regrown slot buffer came back at the same address
    if (!h->graph_exec || h->graph_want_R != want_R || h->graph_wide != h->wide ||
        std::memcmp(&b, &h->graph_bufs, sizeof b) != 0)
        capture_graph(h, want_R);
    if (h->graph_publishes) h->status_host->published = 0;   // the previous fill has published (or ended)
    ++h->gen;
    HIPC(hipGraphLaunch(h->graph_exec, h->stream));
    h->pair_lookback_clean = false;   // the replay's pair scan used them
    h->status_pub_pending = h->graph_publishes;
    ++h->graph_launches;
    h->built = true;
    h->materialized = false;
    h->ev_build = h->ev_pairs = false;
    wait_status(h);
    check_positions(h);
    if (h->nunits > 0 && (h->status_host->small_fallback || h->status_host->wide)) {
        settle(h);
        h->small_ok = false;   // the general pipeline redoes the build (and is captured next time)
        plan_units(h);
        upload_tables(h);
        return false;
    }
    const nlh::Status &st = *h->status_host;
    if (st.npairs > h->pair_capacity || st.max_count > h->slot_cap || (st.wide && !h->graph_wide)) return false;
    finish_pairs(h, want_R);
    return true;
}

// Full rebuild of the current batch on the eager path.
void rebuild_eager(nlh_handle *h, uint32_t flags)
{
    reset_status(h);
    build_cells(h);
    materialize(h, flags);
}

void validate_tuning(const nlh_tuning &t)
{
    if (t.brick_atoms < 64 || t.brick_atoms > 2048)
        throw Error(NLH_EINVAL, "brick_atoms must be in [64, 2048] (the LDS holds 2048 halo atoms)");
    if (t.fill_block < 64 || t.fill_block > 1024 || t.fill_block % 64 != 0)
        throw Error(NLH_EINVAL, "fill_block must be a multiple of 64 in [64, 1024]");
    if (t.fill_pairs != 128 && t.fill_pairs != 256 && t.fill_pairs != 512)
        throw Error(NLH_EINVAL, "fill_pairs must be 128, 256 or 512");
    if (t.count_block != 0 && t.count_block != 256 && t.count_block != 512)
        throw Error(NLH_EINVAL, "count_block must be 0 (automatic), 256 or 512");
    if (t.cell_sort != 0 && t.cell_sort != 1)
        throw Error(NLH_EINVAL, "cell_sort must be 0 (bucketed where it applies) or 1 (atomic histogram)");
    if (t.fused_small != 0 && t.fused_small != 1)
        throw Error(NLH_EINVAL, "fused_small must be 0 (fused build of small systems where it applies) or 1 (off)");
}

void apply_tuning(nlh_handle *h, const nlh_tuning &t)
{
    validate_tuning(t);
    const bool bricks = t.brick_atoms != h->tune.brick_atoms || t.count_block != h->tune.count_block ||
                        t.fused_small != h->tune.fused_small;
    h->tune = t;
    if (bricks && !h->sys_host.empty()) {
        assign_bricks(h);
        plan_units(h);
        upload_tables(h);
    }
    drop_graph(h);
}

}  // namespace

// ---------------------------------------------------------------------------
// neighbours of one atom (lazy API), via the device-side header
// ---------------------------------------------------------------------------

template <typename TI>
__global__ void neighbours_one_kernel(nlh_celllist_dev<TI> c, int64_t i, int64_t cap, TI *j, double *R, TI *S,
                                      int64_t *count)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t n = 0;
    nlh_for_each_neighbour(c, i, [&](TI jj, const double *RR, const TI *SS) {
        if (n < cap) {
            j[n] = jj;
            for (int k = 0; k < 3; ++k) {
                R[3 * n + k] = RR[k];
                S[3 * n + k] = SS[k];
            }
        }
        ++n;
    });
    *count = n;
}

// neighbours of a range of atoms: one thread per atom, a counting pass and
// a filling pass at host-scanned offsets, both through the device-side
// header's traversal (the reference's order, R and S bit for bit)
template <typename TI>
__global__ __launch_bounds__(256) void neighbours_count_kernel(nlh_celllist_dev<TI> c, int64_t i_first, int64_t n,
                                                               int64_t *__restrict__ count)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    int64_t k = 0;
    nlh_for_each_neighbour(c, i_first + a, [&](TI, const double *, const TI *) { ++k; });
    count[a] = k;
}

template <typename TI>
__global__ __launch_bounds__(256) void neighbours_fill_kernel(nlh_celllist_dev<TI> c, int64_t i_first, int64_t n,
                                                              const int64_t *__restrict__ off, TI *__restrict__ j,
                                                              double *__restrict__ R, TI *__restrict__ S)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    int64_t q = off[a];
    nlh_for_each_neighbour(c, i_first + a, [&](TI jj, const double *RR, const TI *SS) {
        j[q] = jj;
        for (int k = 0; k < 3; ++k) {
            R[3 * q + k] = RR[k];
            S[3 * q + k] = SS[k];
        }
        ++q;
    });
}

namespace {

// Cell-list view of system `sys` of the last build (positions, perm and
// cell offsets of that system; perm and cell_offsets are system-local).
nlh_celllist_view system_view(const nlh_handle *h, int32_t sys)
{
    const nlh::SysDev &sd = h->sys_host[sys];
    const int ti = h->ti_bytes;
    nlh_celllist_view v{};
    v.nat = sd.nat;
    v.ncells_total = sd.M;
    v.X_orig = h->X_dev + 3 * sd.atom_base;
    v.perm = static_cast<char *>(h->perm.p) + (size_t)sd.atom_base * ti;
    v.cell_offsets = static_cast<char *>(h->cell_offsets.p) + ((size_t)sd.cell_base + sys) * ti;
    v.ti_bytes = ti;
    return v;
}

template <typename TI>
void neighbours_range_launch(nlh_handle *h, int32_t sys, int64_t i_first, int64_t n)
{
    const auto c = nlh_make_celllist_dev<TI>(system_view(h, sys), h->geoms[sys]);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    h->nr_count.ensure(h->stream, (size_t)n * 8);
    hipLaunchKernelGGL(neighbours_count_kernel<TI>, dim3(blocks), dim3(256), 0, h->stream, c, i_first, n,
                       h->nr_count.as<int64_t>());
    HIPC(hipGetLastError());
    // exclusive scan of the counts on the host (one pass over n values)
    h->nr_off_host.assign((size_t)n + 1, 0);
    HIPC(hipMemcpyAsync(h->nr_off_host.data() + 1, h->nr_count.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int64_t a = 0; a < n; ++a) h->nr_off_host[a + 1] += h->nr_off_host[a];
    const int64_t total = h->nr_off_host[n];
    const size_t t1 = (size_t)std::max<int64_t>(total, 1), ti = sizeof(TI);
    h->nr_off.ensure(h->stream, ((size_t)n + 1) * 8);
    h->nr_j.ensure(h->stream, t1 * ti);
    h->nr_R.ensure(h->stream, t1 * 24);
    h->nr_S.ensure(h->stream, t1 * 3 * ti);
    HIPC(hipMemcpyAsync(h->nr_off.p, h->nr_off_host.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice, h->stream));
    if (total > 0)
        hipLaunchKernelGGL(neighbours_fill_kernel<TI>, dim3(blocks), dim3(256), 0, h->stream, c, i_first, n,
                           h->nr_off.as<int64_t>(), h->nr_j.as<TI>(), h->nr_R.as<double>(), h->nr_S.as<TI>());
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
}

}  // namespace

// ---------------------------------------------------------------------------
// The build that allocates nothing (nlh_neighbour_list_noalloc): every
// buffer beyond the caller's inputs and outputs is carved out of one
// file-scope device array.  Each device the library runs on holds its own
// copy of the array, at its own address, until the process ends.
// ---------------------------------------------------------------------------

constexpr size_t kScratchBytes = (size_t)128 << 20;   // 128 MiB per device
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;                 // set-up, and one no-alloc build at a time per process
void *g_scratch_ptr[kMaxDevices] = {};      // device address of `scratch` on each device
int g_scratch_cus[kMaxDevices] = {};

// The device address of `scratch` on `dev` (current).  The array's name in
// host code is a host-side stand-in, not a device address: only this
// pointer may reach kernels and copies.  Call with g_scratch_mutex held.
void *scratch_ptr(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) throw Error(NLH_EINVAL, "device index beyond the supported range");
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPD(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(NLH_EDEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        int cus = 0;
        HIPC(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        g_scratch_cus[dev] = cus > 0 ? cus : 256;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Bump allocation of 256-byte aligned pieces; with base == nullptr it only
// measures.
struct Carve {
    char *base;
    size_t used = 0;
    template <typename T>
    T *take(size_t bytes)
    {
        const size_t o = (used + 255) & ~(size_t)255;
        used = o + std::max<size_t>(bytes, 8);
        return reinterpret_cast<T *>(base ? base + o : nullptr);
    }
};

struct NoAllocPlan {
    nlh::SysDev sd;
    int64_t atom_base[2];
    int32_t cell_base[2];
    int32_t brick_base[2];
    int64_t nbricks;
    nlh::SortPlan sort;
};

// Lay out every buffer of a single-system build in `base` (or only measure
// when base is null).  The slot rows come last, so that growing them leaves
// the cell list where the count pass reads it.  Returns the bytes used.
size_t carve_noalloc(char *base, const NoAllocPlan &P, int ti, int32_t cap, bool wide, bool want_R, int64_t pair_capacity,
                     const double *X, void *i_out, void *j_out, void *S_out, void *first_out, double *R_out,
                     nlh::PipelineBufs &b)
{
    const size_t n1 = (size_t)std::max<int64_t>(P.sd.nat, 1);
    const size_t M1 = (size_t)P.sd.M + 1;
    Carve c{base};
    std::memset(&b, 0, sizeof b);
    b.X = X;
    b.sys = c.take<nlh::SysDev>(sizeof(nlh::SysDev));
    b.sys_atom_base = c.take<int64_t>(sizeof P.atom_base);
    b.sys_cell_base = c.take<int32_t>(sizeof P.cell_base);
    b.sys_brick_base = c.take<int32_t>(sizeof P.brick_base);
    b.nsys = 1;
    b.atom_sys = nullptr;
    b.nbricks = P.nbricks;
    b.nat = P.sd.nat;
    b.M = P.sd.M;
    b.sort = P.sort;
    b.status = c.take<nlh::Status>(sizeof(nlh::Status));
    b.status_pub = c.take<nlh::Status>(sizeof(nlh::Status));
    b.key = c.take<int32_t>(n1 * 4);
    b.slot = c.take<int32_t>(n1 * 4);
    b.tmp = c.take<int32_t>(n1 * 4);
    b.cell_count = c.take<int32_t>(M1 * 4 + 8);   // zeroed in whole 8-byte words
    b.cell_start = c.take<int32_t>(M1 * 4);
    b.bmat = c.take<int32_t>((size_t)std::max<int64_t>(P.sort.bw >= 0 ? P.sort.n_sort : 0, 1) * 4);
    b.rec = c.take<nlh::Rec>(n1 * sizeof(nlh::Rec));
    b.wbig = c.take<int32_t>(n1 * 12);
    b.perm_out = c.take<void>(n1 * ti);
    b.cell_id_out = c.take<void>(n1 * ti);
    b.cell_offsets_out = c.take<void>(M1 * ti);
    b.lookback = c.take<uint64_t>(
        (size_t)nlh::lookback_entries(std::max<int64_t>((int64_t)M1, P.sort.n_sort), (int64_t)n1) * 8);
    b.pair_count = c.take<int32_t>(n1 * 4);
    b.pair_off = c.take<int64_t>((n1 + 1) * 8);
    b.tile_cap = pair_capacity / nlh::kMinFillPairs + 2;
    b.tile_atom = c.take<int32_t>((size_t)b.tile_cap * 8);
    b.sys_pair_base = c.take<int64_t>(16);
    // the tail: slot rows (the bucketed sort's (cell, index) scratch lives
    // at their start until the count pass rewrites them), then their R
    const size_t rows = n1 * (size_t)cap;
    b.slots = c.take<void>(std::max(rows * (wide ? 16 : 8), n1 * nlh::kBucketBytes));
    b.bke = reinterpret_cast<int2 *>(b.slots);
    b.slot_cap = cap;
    b.slotsR = want_R ? c.take<double>(rows * 24) : nullptr;
    b.first_out = first_out;
    b.i_out = i_out;
    b.j_out = j_out;
    b.S_out = S_out;
    b.R_out = R_out;
    b.pair_capacity = pair_capacity;
    b.ti_bytes = ti;
    b.fill_block = 256;
    b.fill_pairs = 256;
    b.count_block = 256;
    return c.used;
}

// Expected neighbours per atom of one system, as a slot-row length.
int32_t slot_cap_for(const nlh_geometry &g, int64_t nat)
{
    const double vol = std::fabs(det3(g.cell));
    double mx = 0;
    if (nat > 0 && vol > 0) mx = nat / vol * 4.18879020478639 * g.cutoff * g.cutoff * g.cutoff;
    return round_slot_cap(std::min(mx * 1.5 + 16.0, 1.0e6));
}

// Wait for the stream and read the device status back.
nlh::Status read_status(const nlh::PipelineBufs &b, hipStream_t st)
{
    nlh::Status s;
    HIPC(hipMemcpyAsync(&s, b.status, sizeof s, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (s.scan_error) throw Error(NLH_EDEVICE, "internal: a scan block waited too long for its predecessors");
    if (s.bad_position) throw Error(NLH_EINVAL, "positions must be finite and within 1e9 cells of the box");
    return s;
}

[[noreturn]] void scratch_too_small(size_t need)
{
    throw Error(NLH_ENOMEM, "this build needs " + std::to_string(need) + " bytes of scratch; the no-alloc entry point's "
                            "static scratch holds " + std::to_string(kScratchBytes) +
                            " bytes per device (nlh_neighbour_list has no such limit)");
}

void neighbour_list_noalloc(hipStream_t st, const double *X, int64_t nat, const nlh_geometry &g, uint32_t flags,
                            int64_t pair_capacity, void *i_out, void *j_out, void *S_out, void *first_out,
                            double *R_out, int64_t *npairs)
{
    validate_geometry(g);
    const bool want_R = (flags & NLH_WANT_R) != 0;
    if (nat < 0) throw Error(NLH_EINVAL, "nat < 0");
    if (nat > INT32_MAX - 1 || (g.ti_bytes == 4 && nat >= INT32_MAX))
        throw Error(NLH_EOVERFLOW, "too many atoms for one build");
    if (pair_capacity < 0) throw Error(NLH_EINVAL, "pair_capacity < 0");
    if (!first_out) throw Error(NLH_EINVAL, "first is NULL");
    if (nat > 0 && !X) throw Error(NLH_EINVAL, "X is NULL");
    if (pair_capacity > 0 && (!i_out || !j_out || !S_out || (want_R && !R_out)))
        throw Error(NLH_EINVAL, "a pair output is NULL");
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    char *base = static_cast<char *>(scratch_ptr(dev));
    const int ti = g.ti_bytes;
    if (nat == 0) {   // gpu_kernels.jl:303-312: first = [1]
        const int64_t one64 = 1;
        const int32_t one32 = 1;
        HIPC(hipMemcpyAsync(first_out, ti == 8 ? (const void *)&one64 : (const void *)&one32, ti,
                            hipMemcpyHostToDevice, st));
        HIPC(hipStreamSynchronize(st));
        *npairs = 0;
        return;
    }
    // the system table (host side of the upload; the stream is waited for
    // before it goes out of scope)
    NoAllocPlan P{};
    P.sd = make_sys(g, nat, 0, 0, 0);
    const int cus = g_scratch_cus[dev];
    choose_brick(g, nat, std::min<int32_t>(920, (int32_t)(kHaloFill * 4 * 256)), 256, P.sd.brick,
                 (int64_t)cus * 4);
    P.sd.brick_base = 0;
    P.nbricks = brick_classes(P.sd);
    P.atom_base[0] = 0;
    P.atom_base[1] = nat;
    P.cell_base[0] = 0;
    P.cell_base[1] = P.sd.M;
    P.brick_base[0] = 0;
    P.brick_base[1] = (int32_t)P.nbricks;
    P.sort = nlh::sort_plan(P.sd.M, nat, 0);
    int32_t cap = slot_cap_for(g, nat);
    bool wide = false;
    nlh::PipelineBufs b;
    size_t need = carve_noalloc(nullptr, P, ti, cap, wide, want_R, pair_capacity, X, i_out, j_out, S_out, first_out,
                                R_out, b);
    if (need > kScratchBytes) scratch_too_small(need);
    carve_noalloc(base, P, ti, cap, wide, want_R, pair_capacity, X, i_out, j_out, S_out, first_out, R_out, b);
    HIPC(hipMemcpyAsync((void *)b.sys, &P.sd, sizeof P.sd, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync((void *)b.sys_atom_base, P.atom_base, sizeof P.atom_base, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync((void *)b.sys_cell_base, P.cell_base, sizeof P.cell_base, hipMemcpyHostToDevice, st));
    HIPC(hipMemcpyAsync((void *)b.sys_brick_base, P.brick_base, sizeof P.brick_base, hipMemcpyHostToDevice, st));
    // the array keeps the previous call's contents: the status and the scan
    // descriptors start at zero, and so does the atomic sort's histogram
    nlh::launch_clear_status(b, st);
    if (b.sort.bw < 0) nlh::launch_zero_words(reinterpret_cast<uint64_t *>(b.cell_count), ((int64_t)b.M + 2) / 2, st);
    nlh::launch_build(b, st);
    nlh::launch_count(b, want_R, wide, st);
    nlh::launch_scan_pairs(b, st);
    nlh::launch_fill(b, want_R, wide, st);
    HIPC(hipGetLastError());
    nlh::Status s = read_status(b, st);
    if ((s.wide && !wide) || s.max_count > cap) {
        // longer or wider slot rows, then count, scan and fill again (the
        // cell list stays where it is: the slot rows are the layout's tail)
        wide = wide || s.wide;
        cap = std::max(cap, round_slot_cap(std::max(s.max_count, 1)));
        need = carve_noalloc(nullptr, P, ti, cap, wide, want_R, pair_capacity, X, i_out, j_out, S_out, first_out,
                             R_out, b);
        if (need > kScratchBytes) scratch_too_small(need);
        carve_noalloc(base, P, ti, cap, wide, want_R, pair_capacity, X, i_out, j_out, S_out, first_out, R_out, b);
        nlh::launch_clear_pair_lookback(b, st);
        nlh::launch_count(b, want_R, wide, st);
        nlh::launch_scan_pairs(b, st);
        nlh::launch_fill(b, want_R, wide, st);
        HIPC(hipGetLastError());
        s = read_status(b, st);
        if ((s.wide && !wide) || s.max_count > cap)
            throw Error(NLH_EDEVICE, "internal: slot rows still too short after growth");
    }
    *npairs = s.npairs;
    if (ti == 4 && s.npairs >= INT32_MAX)
        throw Error(NLH_EOVERFLOW, "pair count exceeds typemax(Int32); use int_type=Int64");
    if (s.npairs > pair_capacity)
        throw Error(NLH_ENOMEM, "the pair outputs hold " + std::to_string(pair_capacity) + " pairs, the list has " +
                                    std::to_string(s.npairs) + " (*npairs): call again with larger outputs");
}

}  // namespace

extern "C" {

int nlh_device_malloc(uint64_t bytes, void **ptr)
{
    return guarded([&] {
        if (!ptr) throw Error(NLH_EINVAL, "ptr is NULL");
        HIPC(hipSetDevice(g_device));
        *ptr = nullptr;
        HIPD(hipMalloc(ptr, bytes ? bytes : 16));
    });
}

int nlh_device_free(void *ptr)
{
    return guarded([&] {
        if (ptr) HIPD(hipFree(ptr));
    });
}

int nlh_memcpy(void *dst, const void *src, uint64_t bytes)
{
    return guarded([&] {
        if (bytes == 0) return;
        if (!dst || !src) throw Error(NLH_EINVAL, "NULL pointer");
        HIPD(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    });
}

int nlh_device_synchronize(void)
{
    return guarded([&] {
        HIPC(hipSetDevice(g_device));
        HIPD(hipDeviceSynchronize());
    });
}

int nlh_version(int32_t *major, int32_t *minor)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    return NLH_OK;
}

const char *nlh_last_error(void) { return g_last_error.c_str(); }

// Handle buffers come from the device's stream-ordered pool; keep freed
// blocks in it for the next growth rather than returning them at every sync.
static void keep_pool_blocks(int device)
{
    hipMemPool_t pool = nullptr;
    HIPC(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = UINT64_MAX;
    HIPC(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
}

int nlh_init(int device)
{
    return guarded([&] {
        int n = 0;
        HIPC(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw Error(NLH_EINVAL, "no such device");
        HIPC(hipSetDevice(device));
        g_device = device;
        keep_pool_blocks(device);
        // the no-alloc build's scratch: its device address on this device
        std::lock_guard<std::mutex> lk(g_scratch_mutex);
        (void)scratch_ptr(device);
    });
}

int nlh_create(nlh_handle **out)
{
    return guarded([&] {
        if (!out) throw Error(NLH_EINVAL, "out is NULL");
        int n = 0;
        HIPC(hipGetDeviceCount(&n));
        if (n == 0) throw Error(NLH_EDEVICE, "no HIP device");
        auto *h = new nlh_handle();
        h->device = g_device;
        try {
            HIPC(hipSetDevice(h->device));
            keep_pool_blocks(h->device);
            HIPC(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device));
            HIPD(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            HIPD(hipHostMalloc((void **)&h->status_host, sizeof(nlh::Status),
                               hipHostMallocMapped | hipHostMallocCoherent));
            std::memset(h->status_host, 0, sizeof(nlh::Status));
            HIPD(hipHostGetDevicePointer((void **)&h->status_host_dev, h->status_host, 0));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int nlh_destroy(nlh_handle *h)
{
    if (!h) return NLH_OK;
    return guarded([&] {
        (void)hipSetDevice(h->device);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        drop_graph(h);
        device_wide([&] {
            for (hipEvent_t &e : h->ev)
                if (e) (void)hipEventDestroy(e);
        });
        DevBuf *bufs[] = {&h->X,        &h->tables,    &h->key,    &h->slot,
                          &h->cell_count, &h->bmat,     &h->recb,       &h->cell_start, &h->tmp,    &h->rec,
                          &h->wbig,     &h->perm,       &h->cell_id,  &h->cell_offsets, &h->pair_count, &h->slots, &h->slotsR, &h->pair_off,   &h->tiles, &h->lookback, &h->sysmap,
                          &h->sys_pair_base, &h->first, &h->i,          &h->j,            &h->S,      &h->R,
                          &h->status,   &h->Xsorted,   &h->nb_j,       &h->nb_R,         &h->nb_S,   &h->nb_count,
                          &h->nr_count, &h->nr_off,    &h->nr_j,       &h->nr_R,         &h->nr_S};
        for (DevBuf *b : bufs) b->release(h->stream);
        device_wide([&] {
            if (h->status_host) (void)hipHostFree(h->status_host);
            if (h->staging) (void)hipHostFree(h->staging);
            if (h->xstage) (void)hipHostFree(h->xstage);
            if (h->xstage_ev) (void)hipEventDestroy(h->xstage_ev);
            if (h->stream) (void)hipStreamDestroy(h->stream);
        });
        delete h;
    });
}

int nlh_synchronize(nlh_handle *h)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        HIPC(hipSetDevice(h->device));
        finish_status(h);
        check_positions(h);
    });
}

int nlh_get_stream(nlh_handle *h, void **stream)
{
    return guarded([&] {
        if (!h || !stream) throw Error(NLH_EINVAL, "NULL argument");
        *stream = (void *)h->stream;
    });
}

int nlh_geometry_init(const double cell[9], double cutoff, const uint8_t pbc[3], int32_t ti_bytes, nlh_geometry *out)
{
    return guarded([&] {
        if (!cell || !pbc || !out) throw Error(NLH_EINVAL, "NULL argument");
        geometry_init(cell, cutoff, pbc, ti_bytes, *out);
    });
}

int nlh_build_cell_list(nlh_handle *h, const double *X, int64_t nat, const nlh_geometry *geom, uint32_t flags)
{
    return guarded([&] {
        if (!h || !geom) throw Error(NLH_EINVAL, "NULL argument");
        if (nat < 0) throw Error(NLH_EINVAL, "nat < 0");
        HIPC(hipSetDevice(h->device));
        settle(h);
        h->last_single = false;   // not a single full build: no repeat to replay
        h->built = false;
        drop_graph(h);
        const int64_t offs[2] = {0, nat};
        setup_batch(h, 1, offs, geom);
        upload_positions(h, X, flags);
        reset_status(h);
        build_cells(h);
        if (!(flags & NLH_NO_SYNC)) {
            finish_status(h);
            check_positions(h);
        }
    });
}

int nlh_materialize(nlh_handle *h, uint32_t flags, int64_t *npairs)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        HIPC(hipSetDevice(h->device));
        settle(h);
        materialize(h, flags);
        if (npairs) *npairs = h->npairs;
    });
}

int nlh_neighbour_list(nlh_handle *h, const double *X, int64_t nat, const nlh_geometry *geom, uint32_t flags,
                       int64_t *npairs)
{
    return guarded([&] {
        if (!h || !geom) throw Error(NLH_EINVAL, "NULL argument");
        if (nat < 0) throw Error(NLH_EINVAL, "nat < 0");
        HIPC(hipSetDevice(h->device));
        h->built = false;
        const int64_t offs[2] = {0, nat};
        setup_batch(h, 1, offs, geom);
        upload_positions(h, X, flags);
        // a repeat of the previous call's shape replays the captured pipeline
        // (captured on the first repeat); otherwise, or when a replay
        // outgrew a capacity, the stages run one by one
        const nlh::PipelineBufs b = h->bufs();
        const bool repeat = h->last_single && h->last_flags == flags && std::memcmp(&b, &h->last_bufs, sizeof b) == 0;
        // small builds are launch-bound and replay the graph; a larger one
        // launches its stages directly: returning at the fill's start, the
        // host has the next build's first kernels queued before the fill
        // ends, sooner than a graph launch gets going (per build, direct
        // against replayed: 100k atoms 90 / 98 us, 16k 58 / 63 us, 10k 64 /
        // 61 us)
        const bool graph = repeat && !(flags & NLH_NO_GRAPH) && !h->profile && h->materialized && h->nat > 0 &&
                           h->nat < kReplayMaxAtoms;
        h->last_single = false;
        if (!(graph && rebuild_graph(h, (flags & NLH_WANT_R) != 0))) {
            reset_status(h);
            build_cells(h);
            materialize(h, flags);
        }
        h->last_single = true;
        h->last_flags = flags;
        h->last_bufs = h->bufs();
        if (npairs) *npairs = h->npairs;
    });
}

int nlh_pairlist_device(nlh_handle *h, nlh_pairlist_view *out)
{
    return guarded([&] {
        if (!h || !out) throw Error(NLH_EINVAL, "NULL argument");
        if (!h->materialized) throw Error(NLH_ESTATE, "no materialised pair list");
        HIPC(hipSetDevice(h->device));
        settle(h);   // every output of the last build is complete
        out->nat = h->nat;
        out->npairs = h->npairs;
        out->i = h->i.p;
        out->j = h->j.p;
        out->S = h->S.p;
        out->first = h->first.p;
        out->R = h->have_R ? h->R.as<double>() : nullptr;
        out->ti_bytes = h->ti_bytes;
    });
}

int nlh_celllist_device(nlh_handle *h, nlh_celllist_view *out)
{
    return guarded([&] {
        if (!h || !out) throw Error(NLH_EINVAL, "NULL argument");
        if (!h->built) throw Error(NLH_ESTATE, "no cell list built");
        if (h->nsys != 1) throw Error(NLH_ESTATE, "cell-list views are per single-system build");
        HIPC(hipSetDevice(h->device));
        settle(h);   // every output of the last build is complete
        out->nat = h->nat;
        out->ncells_total = h->M;
        out->X_orig = h->X_dev;
        out->perm = h->perm.p;
        out->cell_id = h->cell_id.p;
        out->cell_offsets = h->cell_offsets.p;
        out->ti_bytes = h->ti_bytes;
    });
}

int nlh_copy_pairlist(nlh_handle *h, void *i, void *j, void *S, void *first, double *R)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        if (!h->materialized) throw Error(NLH_ESTATE, "no materialised pair list");
        HIPC(hipSetDevice(h->device));
        settle(h);
        HIPC(hipStreamSynchronize(h->stream));
        const size_t P = (size_t)h->npairs, ti = (size_t)h->ti_bytes;
        if (P > 0) {
            if (i) HIPD(hipMemcpy(i, h->i.p, P * ti, hipMemcpyDefault));
            if (j) HIPD(hipMemcpy(j, h->j.p, P * ti, hipMemcpyDefault));
            if (S) HIPD(hipMemcpy(S, h->S.p, 3 * P * ti, hipMemcpyDefault));
            if (R) {
                if (!h->have_R) throw Error(NLH_ESTATE, "R was not requested (NLH_WANT_R)");
                HIPD(hipMemcpy(R, h->R.p, 3 * P * 8, hipMemcpyDefault));
            }
        }
        if (first) HIPD(hipMemcpy(first, h->first.p, ((size_t)h->nat + h->nsys) * ti, hipMemcpyDefault));
    });
}

int nlh_copy_celllist(nlh_handle *h, double *X_sorted, void *perm, void *cell_id, void *cell_offsets)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        if (!h->built) throw Error(NLH_ESTATE, "no cell list built");
        HIPC(hipSetDevice(h->device));
        settle(h);
        const size_t n = (size_t)h->nat, ti = (size_t)h->ti_bytes;
        if (X_sorted && n > 0) {
            h->Xsorted.ensure(h->stream, n * 24);
            nlh::launch_sorted_positions(h->bufs(), h->Xsorted.as<double>(), h->stream);
            HIPC(hipGetLastError());
        }
        HIPC(hipStreamSynchronize(h->stream));
        if (X_sorted && n > 0) HIPD(hipMemcpy(X_sorted, h->Xsorted.p, n * 24, hipMemcpyDefault));
        if (perm && n > 0) HIPD(hipMemcpy(perm, h->perm.p, n * ti, hipMemcpyDefault));
        if (cell_id && n > 0) HIPD(hipMemcpy(cell_id, h->cell_id.p, n * ti, hipMemcpyDefault));
        if (cell_offsets)
            HIPD(hipMemcpy(cell_offsets, h->cell_offsets.p, ((size_t)h->M + h->nsys) * ti, hipMemcpyDefault));
    });
}

int nlh_neighbours(nlh_handle *h, int32_t sys, int64_t i, int64_t cap, void *j, double *R, void *S, int64_t *count)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        if (!h->built) throw Error(NLH_ESTATE, "no cell list built");
        if (sys < 0 || sys >= h->nsys) throw Error(NLH_EINVAL, "no such system");
        const nlh::SysDev &sd = h->sys_host[sys];
        if (i < 1 || i > sd.nat) throw Error(NLH_EINVAL, "atom index out of range (1-based)");
        if (cap < 0) throw Error(NLH_EINVAL, "cap < 0");
        HIPC(hipSetDevice(h->device));
        settle(h);
        const int ti = h->ti_bytes;
        const size_t c1 = (size_t)std::max<int64_t>(cap, 1);
        h->nb_j.ensure(h->stream, c1 * ti);
        h->nb_R.ensure(h->stream, c1 * 24);
        h->nb_S.ensure(h->stream, c1 * 3 * ti);
        h->nb_count.ensure(h->stream, 8);
        const nlh_celllist_view v = system_view(h, sys);
        if (ti == 8) {
            auto c = nlh_make_celllist_dev<int64_t>(v, h->geoms[sys]);
            hipLaunchKernelGGL(neighbours_one_kernel<int64_t>, dim3(1), dim3(64), 0, h->stream, c, i, cap,
                               h->nb_j.as<int64_t>(), h->nb_R.as<double>(), h->nb_S.as<int64_t>(),
                               h->nb_count.as<int64_t>());
        } else {
            auto c = nlh_make_celllist_dev<int32_t>(v, h->geoms[sys]);
            hipLaunchKernelGGL(neighbours_one_kernel<int32_t>, dim3(1), dim3(64), 0, h->stream, c, i, cap,
                               h->nb_j.as<int32_t>(), h->nb_R.as<double>(), h->nb_S.as<int32_t>(),
                               h->nb_count.as<int64_t>());
        }
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(h->stream));
        int64_t n = 0;
        HIPD(hipMemcpy(&n, h->nb_count.p, 8, hipMemcpyDeviceToHost));
        const size_t m = (size_t)std::min(n, cap);
        if (m > 0) {
            if (j) HIPD(hipMemcpy(j, h->nb_j.p, m * ti, hipMemcpyDeviceToHost));
            if (R) HIPD(hipMemcpy(R, h->nb_R.p, m * 24, hipMemcpyDeviceToHost));
            if (S) HIPD(hipMemcpy(S, h->nb_S.p, m * 3 * ti, hipMemcpyDeviceToHost));
        }
        if (count) *count = n;
    });
}

int nlh_neighbours_range(nlh_handle *h, int32_t sys, int64_t i_first, int64_t n, int64_t cap, int64_t *offsets,
                         void *j, double *R, void *S)
{
    return guarded([&] {
        if (!h || !offsets) throw Error(NLH_EINVAL, "NULL argument");
        if (!h->built) throw Error(NLH_ESTATE, "no cell list built");
        if (sys < 0 || sys >= h->nsys) throw Error(NLH_EINVAL, "no such system");
        const nlh::SysDev &sd = h->sys_host[sys];
        if (n < 0 || i_first < 1 || i_first + n - 1 > sd.nat)
            throw Error(NLH_EINVAL, "atom range out of bounds (1-based, within the system)");
        if (cap < 0) throw Error(NLH_EINVAL, "cap < 0");
        HIPC(hipSetDevice(h->device));
        settle(h);
        const bool cached = h->nr_valid && h->nr_gen == h->gen && h->nr_sys == sys && h->nr_first == i_first &&
                            h->nr_n == n;
        if (!cached) {
            h->nr_valid = false;
            if (n > 0) {
                if (h->ti_bytes == 8) neighbours_range_launch<int64_t>(h, sys, i_first, n);
                else neighbours_range_launch<int32_t>(h, sys, i_first, n);
            } else {
                h->nr_off_host.assign(1, 0);
            }
            h->nr_valid = true;
            h->nr_gen = h->gen;
            h->nr_sys = sys;
            h->nr_first = i_first;
            h->nr_n = n;
        }
        std::memcpy(offsets, h->nr_off_host.data(), ((size_t)n + 1) * sizeof(int64_t));
        const int64_t total = h->nr_off_host[n];
        if (total > 0 && total <= cap) {
            const size_t ti = (size_t)h->ti_bytes;
            if (j) HIPD(hipMemcpy(j, h->nr_j.p, (size_t)total * ti, hipMemcpyDeviceToHost));
            if (R) HIPD(hipMemcpy(R, h->nr_R.p, (size_t)total * 24, hipMemcpyDeviceToHost));
            if (S) HIPD(hipMemcpy(S, h->nr_S.p, (size_t)total * 3 * ti, hipMemcpyDeviceToHost));
        }
    });
}

int nlh_neighbour_list_noalloc(void *stream, const double *X, int64_t nat, const nlh_geometry *geom, uint32_t flags,
                               int64_t pair_capacity, void *i, void *j, void *S, void *first, double *R,
                               int64_t *npairs)
{
    return guarded([&] {
        if (!geom || !npairs) throw Error(NLH_EINVAL, "NULL argument");
        *npairs = 0;
        neighbour_list_noalloc(static_cast<hipStream_t>(stream), X, nat, *geom, flags, pair_capacity, i, j, S, first,
                               R, npairs);
    });
}

int nlh_set_profiling(nlh_handle *h, int enable)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        HIPC(hipSetDevice(h->device));
        settle(h);
        if (enable && !h->ev[0])
            for (hipEvent_t &e : h->ev) HIPD(hipEventCreate(&e));
        h->profile = enable != 0;
    });
}

int nlh_graph_launches(nlh_handle *h, int64_t *n)
{
    return guarded([&] {
        if (!h || !n) throw Error(NLH_EINVAL, "NULL argument");
        *n = h->graph_launches;
    });
}

int nlh_phase_times(nlh_handle *h, float ms[5])
{
    return guarded([&] {
        if (!h || !ms) throw Error(NLH_EINVAL, "NULL argument");
        HIPC(hipSetDevice(h->device));
        for (int k = 0; k < 5; ++k) ms[k] = 0.0f;
        if (h->ev_build) HIPC(hipEventElapsedTime(&ms[0], h->ev[0], h->ev[1]));
        if (h->ev_pairs) {
            HIPC(hipEventElapsedTime(&ms[1], h->ev[1], h->ev[2]));
            HIPC(hipEventElapsedTime(&ms[2], h->ev[2], h->ev[3]));
            HIPC(hipEventElapsedTime(&ms[3], h->ev[3], h->ev[4]));
        }
        if (h->ev_build && h->ev_pairs) HIPC(hipEventElapsedTime(&ms[4], h->ev[0], h->ev[4]));
    });
}

int nlh_batch_neighbour_list(nlh_handle *h, int32_t nsys, const double *X_batch, const int64_t *atom_offsets,
                             const nlh_geometry *geoms, uint32_t flags, int64_t *total_pairs)
{
    return guarded([&] {
        if (!h || !atom_offsets || !geoms) throw Error(NLH_EINVAL, "NULL argument");
        HIPC(hipSetDevice(h->device));
        settle(h);
        h->last_single = false;   // not a single full build: no repeat to replay
        h->built = false;
        drop_graph(h);
        setup_batch(h, nsys, atom_offsets, geoms);
        upload_positions(h, X_batch, flags);
        reset_status(h);
        build_cells(h);
        materialize(h, flags);
        if (total_pairs) *total_pairs = h->npairs;
    });
}

int nlh_batch_pair_offsets(nlh_handle *h, int64_t *pair_offsets)
{
    return guarded([&] {
        if (!h || !pair_offsets) throw Error(NLH_EINVAL, "NULL argument");
        if (!h->materialized) throw Error(NLH_ESTATE, "no materialised pair list");
        if (h->offsets_stale) {
            HIPC(hipSetDevice(h->device));
            settle(h);
            HIPD(hipMemcpy(h->pair_offsets_host.data(), h->sys_pair_base.p, (h->nsys + 1) * 8,
                           hipMemcpyDeviceToHost));
            h->offsets_stale = false;
        }
        std::memcpy(pair_offsets, h->pair_offsets_host.data(), (h->nsys + 1) * sizeof(int64_t));
    });
}

int nlh_batch_rebuild(nlh_handle *h, const double *X_batch, uint32_t flags, int64_t *total_pairs)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        if (h->sys_host.empty()) throw Error(NLH_ESTATE, "no previous batch to rebuild");
        HIPC(hipSetDevice(h->device));
        h->last_single = false;   // not a single full build: no repeat to replay
        upload_positions(h, X_batch, flags);
        const bool want_R = (flags & NLH_WANT_R) != 0;
        const bool graph = !(flags & NLH_NO_GRAPH) && !h->profile && h->materialized && h->nat > 0;
        if (!(graph && rebuild_graph(h, want_R))) rebuild_eager(h, flags);
        if (total_pairs) *total_pairs = h->npairs;
    });
}

int nlh_set_tuning(nlh_handle *h, const nlh_tuning *t)
{
    return guarded([&] {
        if (!h || !t) throw Error(NLH_EINVAL, "NULL argument");
        HIPC(hipSetDevice(h->device));
        settle(h);
        apply_tuning(h, *t);
    });
}

int nlh_get_tuning(nlh_handle *h, nlh_tuning *t)
{
    return guarded([&] {
        if (!h || !t) throw Error(NLH_EINVAL, "NULL argument");
        *t = h->tune;
    });
}

int nlh_autotune(nlh_handle *h, const double *X_batch, uint32_t flags, int32_t reps, nlh_tuning *best)
{
    return guarded([&] {
        if (!h) throw Error(NLH_EINVAL, "handle is NULL");
        if (h->sys_host.empty()) throw Error(NLH_ESTATE, "autotune needs a previous build on this handle");
        if (reps < 1 || reps > 100) throw Error(NLH_EINVAL, "reps must be in [1, 100]");
        HIPC(hipSetDevice(h->device));
        settle(h);
        upload_positions(h, X_batch, flags);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPD(hipEventCreate(&e0));
        HIPD(hipEventCreate(&e1));
        const nlh_tuning start = h->tune;
        nlh_tuning win = start;
        float win_ms = 0.0f;
        bool have = false;
        try {
            static const int32_t bricks[] = {460, 920, 1840};
            static const int32_t fills[] = {128, 256, 512};
            static const int32_t tiles[] = {256, 512};
            static const int32_t blocks[] = {256, 512};
            for (int32_t cb : blocks)
            for (int32_t ba : bricks)
                for (int32_t fb : fills)
                    for (int32_t fa : tiles) {
                    if (cb == 256 && ba > 920) continue;   // above what a 256-thread brick holds
                    apply_tuning(h, nlh_tuning{ba, fb, fa, cb, start.cell_sort, start.fused_small});
                    rebuild_eager(h, flags & ~NLH_NO_SYNC);   // warm-up (and buffer growth)
                    float tot = 0.0f;
                    for (int r = 0; r < reps; ++r) {
                        HIPC(hipEventRecord(e0, h->stream));
                        rebuild_eager(h, flags & ~NLH_NO_SYNC);
                        HIPC(hipEventRecord(e1, h->stream));
                        HIPC(hipEventSynchronize(e1));
                        float ms = 0.0f;
                        HIPC(hipEventElapsedTime(&ms, e0, e1));
                        tot += ms;
                    }
                    if (!have || tot < win_ms) {
                        win = h->tune;
                        win_ms = tot;
                        have = true;
                    }
                }
            apply_tuning(h, win);
            rebuild_eager(h, flags & ~NLH_NO_SYNC);   // results of the winning configuration
        } catch (...) {
            device_wide([&] {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
            });
            throw;
        }
        device_wide([&] {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        });
        if (best) *best = win;
    });
}

}  // extern "C"
