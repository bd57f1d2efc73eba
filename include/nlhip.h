/*
 * nlhip -- MI355X-native (gfx950) sort-based neighbour lists.
 *
 * C ABI of libnlhip.so.  This is the drop-in boundary for the sort-based
 * hot path of NeighbourLists.jl (v0.6.2): a host binding (the Julia `ccall`
 * shim, or the Python `nlhip` package) calls these entry points in place of
 * the reference's backend-dispatched build and materialisation methods.
 *
 * Layouts are the reference's:
 *   - positions X: nat x 3 float64, AoS (Vector{SVector{3,Float64}} memory);
 *   - cell: 3x3 float64 column-major (SMatrix memory order), rows are the
 *     lattice vectors (src/cell_list.jl:627);
 *   - pbc: 3 bytes (Julia Bool);
 *   - every index array (i, j, S, first, perm, cell_id, cell_offsets) holds
 *     1-based values of the integer type TI (4 = Int32, 8 = Int64 bytes);
 *   - S is nP x 3 TI, AoS (Vector{SVector{3,TI}} memory);
 *   - PairList has no R field (src/types.jl:34-43); R is optional output.
 *
 * Conventions: every call returns an nlh_status; no C++ exception crosses
 * the boundary; nlh_last_error() returns a thread-local message for the
 * last failing call on the calling thread.  A handle owns one HIP stream
 * and its device buffers; separate handles may be used concurrently from
 * separate host threads.  Input buffers are only read during the call.
 * Device results stay owned by the handle until the next build on it or
 * nlh_destroy.
 */
#ifndef NLHIP_H
#define NLHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NLH_OK = 0,
    NLH_EINVAL = 1,     /* invalid argument (bad cell, cutoff, positions:
                           a coordinate whose cell index is not finite or
                           lies beyond 1e9 cells; the reference's
                           unsafe_trunc is undefined there,
                           src/cell_list.jl:66-74)                        */
    NLH_EOVERFLOW = 2,  /* cell count or pair count exceeds the index type */
    NLH_ENOMEM = 3,     /* device allocation failed / budget exceeded      */
    NLH_EDEVICE = 4,    /* HIP runtime error                               */
    NLH_ESTATE = 5      /* call order: e.g. materialise before a build     */
} nlh_status;

/* build / materialise flags */
#define NLH_X_DEVICE 0x1u  /* X (and X_batch) is a device pointer             */
#define NLH_WANT_R 0x2u    /* also write R = X[j] - X[i] + C' S per pair      */
#define NLH_NO_SYNC 0x4u   /* do not wait for the stream (nlh_synchronize)    */
#define NLH_NO_GRAPH 0x8u  /* nlh_batch_rebuild, nlh_neighbour_list: run the
                              stages one by one (no graph replay)            */

/* Geometry of one system: the outputs of analyze_cell plus the neighbour
 * reach and cutoff^2 (src/cell_list.jl:152-170, 655-659;
 * src/gpu_kernels.jl:315-317).  nlh_geometry_init fills it; a caller may
 * instead fill inv_cell / ncells / nxyz from its own arithmetic (e.g. the
 * Julia shim passing StaticArrays' inv) so binning uses exactly its values. */
typedef struct {
    double cell[9];      /* column-major, rows = lattice vectors          */
    double inv_cell[9];  /* column-major inverse                          */
    double lens[3];      /* abs.(lengths(cell)), face-to-face heights     */
    double cutoff;
    double cutoff_sq;    /* cutoff^2                                      */
    int64_t ncells[3];
    int64_t nxyz[3];     /* ceil.(cutoff * (ncells ./ lens))              */
    int64_t ncells_total;
    uint8_t pbc[3];
    uint8_t zero_volume; /* |det(cell)| < 1e-12 (the reference warns)     */
    int32_t ti_bytes;    /* 4 (Int32) or 8 (Int64)                        */
} nlh_geometry;

/* Device-pointer views of a handle's results (valid until the next build). */
typedef struct {
    int64_t nat, npairs;
    void *i, *j, *S, *first; /* TI arrays (1-based)                       */
    double *R;               /* NULL unless NLH_WANT_R                    */
    int32_t ti_bytes;
} nlh_pairlist_view;

typedef struct {
    int64_t nat, ncells_total;
    const double *X_orig;    /* input positions on the device             */
    void *perm, *cell_id;    /* TI[nat], 1-based, sorted order            */
    void *cell_offsets;      /* TI[ncells_total + 1], 1-based CSR         */
    int32_t ti_bytes;
} nlh_celllist_view;

typedef struct nlh_handle nlh_handle;

/* Select the device used by handles created afterwards on this thread and
 * initialise the runtime (one-time).  Replaces the backend choice
 * get_array_backend (src/NeighbourLists.jl:14-18). */
int nlh_init(int device);

int nlh_create(nlh_handle **out);
int nlh_destroy(nlh_handle *h);
const char *nlh_last_error(void);
int nlh_synchronize(nlh_handle *h);
/* The handle's hipStream_t, for callers that order their own work on it. */
int nlh_get_stream(nlh_handle *h, void **stream);

/* analyze_cell + overflow check + nxyz (src/cell_list.jl:152-170, 655-659;
 * src/gpu_kernels.jl:315-317).  NLH_EOVERFLOW when prod(ncells) exceeds
 * typemax(TI) (the reference's error at cell_list.jl:655-659). */
int nlh_geometry_init(const double cell[9], double cutoff, const uint8_t pbc[3], int32_t ti_bytes,
                      nlh_geometry *out);

/* Sort-based cell list: replaces _build_sorted_celllist
 * (src/cell_list.jl:647-679) -- _compute_cell_ids (:684 / gpu_kernels.jl:378),
 * _get_sortperm (:706 / :711) and _compute_cell_offsets (:724 /
 * gpu_kernels.jl:386).  X is nat x 3 float64 (host, or device with
 * NLH_X_DEVICE; a device X must stay valid while the handle uses it). */
int nlh_build_cell_list(nlh_handle *h, const double *X, int64_t nat, const nlh_geometry *geom, uint32_t flags);

/* Two-pass pair materialisation: replaces materialize_pairlist
 * (src/gpu_kernels.jl:299-364).  On return *npairs is the pair count. */
int nlh_materialize(nlh_handle *h, uint32_t flags, int64_t *npairs);

/* neighbour_list(X, cutoff, cell, pbc; lazy=false) in one call
 * (src/cell_list.jl:897-916): build + materialise with a single wait.  A
 * call that repeats the previous call's shape on this handle (same atom
 * count, cell grid, positions buffer, capacities and flags; positions and
 * geometry values may differ) of fewer than 12,288 atoms replays the
 * pipeline captured as a HIP graph, with identical results; NLH_NO_GRAPH
 * runs the stages one by one.  Every build (and nlh_materialize,
 * nlh_batch_neighbour_list, nlh_batch_rebuild) returns once *npairs is
 * final, while the last stage may still be writing the pairs on the
 * handle's stream: work queued on that stream, any later nlh_* call on the
 * handle, nlh_synchronize and nlh_device_synchronize see the finished pair
 * list. */
int nlh_neighbour_list(nlh_handle *h, const double *X, int64_t nat, const nlh_geometry *geom, uint32_t flags,
                       int64_t *npairs);

/* Results: device views, or copies into caller (host) buffers.  Any
 * destination pointer may be NULL to skip that array.  Sizes: i, j: npairs;
 * S, R: 3*npairs; first: nat+1; perm, cell_id: nat; cell_offsets:
 * ncells_total+1; X_sorted: 3*nat. */
int nlh_pairlist_device(nlh_handle *h, nlh_pairlist_view *out);
int nlh_celllist_device(nlh_handle *h, nlh_celllist_view *out);
int nlh_copy_pairlist(nlh_handle *h, void *i, void *j, void *S, void *first, double *R);
int nlh_copy_celllist(nlh_handle *h, double *X_sorted, void *perm, void *cell_id, void *cell_offsets);

/* Lazy access, one atom: for_each_neighbour / neighbours / count_neighbours
 * on a SortedCellList (src/cell_list.jl:779-833).  i is 1-based (and for a
 * batch, local to system `sys`).  Writes up to cap neighbours to the host
 * buffers j (TI), R (3 x float64), S (3 x TI), any may be NULL; *count is the
 * full neighbour count. */
int nlh_neighbours(nlh_handle *h, int32_t sys, int64_t i, int64_t cap, void *j, double *R, void *S,
                   int64_t *count);

/* Lazy access, a range of atoms: neighbours(clist, i) for every i in
 * i_first .. i_first + n - 1 (1-based, local to system `sys`) in one pass
 * on the device -- what a loop of for_each_neighbour / neighbours calls
 * over the atoms returns (src/cell_list.jl:779-833), in the same order.
 * offsets (host, n + 1 entries) receives the 0-based start of each atom's
 * neighbours, offsets[n] their total; when offsets[n] <= cap the
 * neighbours are written to the host buffers j (TI), R (3 x float64) and
 * S (3 x TI), any of which may be NULL.  The results stay on the device
 * until the next build or range query on the handle: asking again for the
 * same range (e.g. with cap = offsets[n]) only copies them. */
int nlh_neighbours_range(nlh_handle *h, int32_t sys, int64_t i_first, int64_t n, int64_t cap, int64_t *offsets,
                         void *j, double *R, void *S);

/* ---- batched build (many independent systems, one launch per stage) ----
 * Systems s = 0..nsys-1 own atoms [atom_offsets[s], atom_offsets[s+1]) of
 * X_batch; every geometry must share ti_bytes.  The reference has no
 * batching API: each system's result equals a single nlh_neighbour_list on
 * that system.  Pair arrays are concatenated by system; system s owns pairs
 * [pair_offsets[s], pair_offsets[s+1]) with j local and 1-based, and its
 * `first` array is entries [atom_offsets[s] + s, atom_offsets[s+1] + s] of
 * the concatenated first array. */
int nlh_batch_neighbour_list(nlh_handle *h, int32_t nsys, const double *X_batch, const int64_t *atom_offsets,
                             const nlh_geometry *geoms, uint32_t flags, int64_t *total_pairs);
int nlh_batch_pair_offsets(nlh_handle *h, int64_t *pair_offsets /* nsys + 1, host */);
/* Rebuild the last batch (or single system) on new positions of the same
 * systems (same atom counts and geometries).  The first call captures the
 * whole pipeline (status clear, binning, sort, cell offsets, pair count,
 * pair scan, fill) as one HIP graph on the handle's stream; later calls
 * replay it while buffers, capacities, flags and the position pointer stay
 * the same (host positions are copied into the handle's own buffer first).
 * A replay that outgrows the pair or slot capacity is redone on the stage
 * path, which grows them.  NLH_NO_GRAPH forces the stage path.  The
 * reference has no rebuild entry: a rebuild there is a fresh
 * neighbour_list (src/cell_list.jl:897-916). */
int nlh_batch_rebuild(nlh_handle *h, const double *X_batch, uint32_t flags, int64_t *total_pairs);

/* The build of nlh_neighbour_list that allocates nothing.  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * that runs under HIP graph capture elsewhere, tight serving loops,
 * harnesses that forbid hipMalloc).  The caller passes the device positions
 * X (nat x 3 float64) and device outputs i, j (pair_capacity TI), S
 * (3 x pair_capacity TI), first (nat + 1 TI) and, with NLH_WANT_R, R
 * (3 x pair_capacity float64); the results equal nlh_neighbour_list's bit
 * for bit.  Every other buffer of the build lives in one static device
 * array of 128 MiB per device (held from the first nlh_init or call on a
 * device until the process ends): a build that needs more (about 220,000
 * atoms at 0.05 atoms/A^3 and cutoff 5 A without R, 60,000 with R) fails
 * with NLH_ENOMEM and a message saying so.  Work runs on `stream` (a
 * hipStream_t; NULL for the default stream), which is waited for before the
 * call returns.  When the list has more pairs than pair_capacity the call
 * fails with NLH_ENOMEM, writes nothing to i, j, S, R and reports the pair
 * count in *npairs (call again with larger outputs).  Calls share the
 * device's array: they are serialised within this process, and calls from
 * several processes on one GPU that overlap race. */
int nlh_neighbour_list_noalloc(void *stream, const double *X, int64_t nat, const nlh_geometry *geom, uint32_t flags,
                               int64_t pair_capacity, void *i, void *j, void *S, void *first, double *R,
                               int64_t *npairs);

/* Launch configuration of the build.  brick_atoms: largest expected atom
 * count (halo included) of one count-pass brick, which sets the brick shape
 * of each system (64..2048; default 920; at most 90% of the 4 x count_block
 * atoms a brick's LDS holds is used).  fill_block: threads per fill
 * workgroup, a multiple of 64 up to 1024.  fill_pairs: output pairs per
 * fill wave (128, 256 or 512; each wave writes one aligned tile of the
 * pair arrays; default 256).  count_block: threads per count-pass brick workgroup, 256 or
 * 512, or 0 for the default (256).  cell_sort: 0 (default) sorts
 * the atoms by cell through buckets of cells counted in LDS wherever the
 * cell grid allows (up to 8192 buckets of 4096 cells), 1 uses one global
 * atomic histogram over the cells.  fused_small: 0 (default) builds a batch
 * of at least 64 systems, each with at most 2048 atoms, 2048 cells, 256
 * cells per axis and a neighbour-cell reach of at most 3 per axis, by one
 * workgroup per unit of
 * consecutive systems (cell list, counts and slot rows in one launch, all in
 * LDS); 1 always uses the general pipeline.
 * Every configuration gives identical results; no reference counterpart
 * (KernelAbstractions picks workgroup sizes itself). */
typedef struct {
    int32_t brick_atoms;
    int32_t fill_block;
    int32_t fill_pairs;
    int32_t count_block;
    int32_t cell_sort;
    int32_t fused_small;
} nlh_tuning;
int nlh_set_tuning(nlh_handle *h, const nlh_tuning *t);
int nlh_get_tuning(nlh_handle *h, nlh_tuning *t);
/* Time the current batch (set up by a previous build on this handle) under
 * each candidate configuration (count_block in {256, 512} x brick_atoms
 * in {460, 920, 1840} (920 at most for 256) x fill_block in {128, 256, 512}
 * x fill_pairs in {256, 512}, reps timed rebuilds each, HIP events on the
 * handle's stream), keep the fastest and return it in *best.  Leaves the
 * handle holding a build of X_batch under the winning configuration. */
int nlh_autotune(nlh_handle *h, const double *X_batch, uint32_t flags, int32_t reps, nlh_tuning *best);

/* Per-phase timing of the last build on this handle, from HIP events on
 * the handle's stream (enabled by nlh_set_profiling): ms[0] cell list (bin,
 * histogram, scan, placement), ms[1] pair count, ms[2] pair-offset scan
 * and i array, ms[3] pair fill (j, S, R), ms[4] whole pipeline.  The
 * reference has no counterpart (its benchmark times whole calls,
 * scripts/benchmark.jl:124-151). */
int nlh_set_profiling(nlh_handle *h, int enable);
int nlh_phase_times(nlh_handle *h, float ms[5]);

/* Number of builds on this handle that replayed the captured HIP graph
 * (nlh_batch_rebuild, repeated nlh_neighbour_list).  No reference
 * counterpart: a diagnostic of the replay path. */
int nlh_graph_launches(nlh_handle *h, int64_t *n);

/* Device memory of the library's own HIP runtime, for callers that keep
 * positions resident without another GPU runtime (no reference
 * counterpart: Julia arrays come from AMDGPU.jl). */
int nlh_device_malloc(uint64_t bytes, void **ptr);
int nlh_device_free(void *ptr);
int nlh_memcpy(void *dst, const void *src, uint64_t bytes); /* any direction */
int nlh_device_synchronize(void);

/* Library self-description for the host bindings. */
int nlh_version(int32_t *major, int32_t *minor);

#ifdef __cplusplus
}
#endif
#endif /* NLHIP_H */
