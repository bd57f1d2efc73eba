/*
 * CPU restatement of NeighbourLists.jl's sort-based and legacy neighbour-list
 * algorithms.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and
 * the CPU-baseline leg of bench.py as the checker; never linked into the
 * product library.
 *
 * Conventions follow the reference:
 *   - cell is 3x3 column-major (Julia SMatrix memory order), element (m,k)
 *     at cell[m + 3*k]; ROWS are lattice vectors (src/cell_list.jl:627).
 *   - every index written by these functions is 1-based (Julia).
 *   - all integers are int64 here; the values equal the reference's TI
 *     (Int32/Int64) values whenever they fit.
 *   - floating-point evaluation order is the reference's left-to-right,
 *     non-fused order (compiled with -ffp-contract=off).
 */
#ifndef NL_ORACLE_H
#define NL_ORACLE_H
#include <stdint.h>

typedef struct {
    double cell[9];      /* column-major, rows = lattice vectors           */
    double inv[9];       /* inv(cell), column-major                         */
    double lens[3];      /* abs.(lengths(cell))                             */
    double cutoff;
    double cutoff_sq;    /* cutoff^2 = cutoff*cutoff                        */
    int64_t ncells[3];
    int64_t nxyz[3];
    int64_t ncells_total;
    int32_t pbc[3];
    int32_t zero_volume; /* analyze_cell's |det| < 1e-12 warning           */
} orc_geom;

/* status codes */
#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_EOVERFLOW 2

/* analyze_cell + overflow check + nxyz (cell_list.jl:152-170, 655-659;
 * gpu_kernels.jl:315-317).  ti_bits is 32 or 64. */
int orc_analyze(const double cell[9], double cutoff, const int32_t pbc[3],
                int ti_bits, orc_geom *g);

/* StaticArrays-order 3x3 helpers, exposed for tests. */
void orc_inv3(const double a[9], double out[9]);
double orc_det3(const double a[9]);
void orc_lengths(const double c[9], double out[3]);

/* _compute_cell_ids (cell_list.jl:684-701): 1-based linear cell ids, plus
 * the per-atom winding numbers (cell_list.jl:141-147) in wind[3*i+k]. */
void orc_cell_ids(const double *X, int64_t nat, const orc_geom *g,
                  int64_t *ids, int64_t *wind);

/* stable sortperm (cell_list.jl:706-708): perm is 1-based */
void orc_sortperm(const int64_t *ids, int64_t nat, int64_t M, int64_t *perm);

/* _compute_cell_offsets (cell_list.jl:724-748): M+1 entries, 1-based */
void orc_cell_offsets(const int64_t *sorted_ids, int64_t nat, int64_t M,
                      int64_t *offsets);

/* count_neighbours_kernel! (gpu_kernels.jl:132-152) over all atoms,
 * parallel over atoms with nthreads POSIX threads.  Returns total pairs. */
int64_t orc_count(const double *X, int64_t nat, const orc_geom *g,
                  const int64_t *perm, const int64_t *cell_offsets,
                  int64_t *counts, int nthreads);

/* compute_pair_offsets CPU (gpu_kernels.jl:211-219): nat+1 entries */
void orc_pair_offsets(const int64_t *counts, int64_t nat, int64_t *first);

/* fill_pairs_kernel! (gpu_kernels.jl:159-180).  R may be NULL. */
void orc_fill(const double *X, int64_t nat, const orc_geom *g,
              const int64_t *perm, const int64_t *cell_offsets,
              const int64_t *first, int64_t *i_out, int64_t *j_out,
              int64_t *S_out, double *R_out, int nthreads);

/* _for_each_neighbor_pair for one atom i (1-based), lazy API
 * (cell_list.jl:779-833).  Writes up to cap neighbours; returns the count. */
int64_t orc_neighbours_of(const double *X, int64_t nat, const orc_geom *g,
                          const int64_t *perm, const int64_t *cell_offsets,
                          int64_t i, int64_t cap, int64_t *j_out,
                          int64_t *S_out, double *R_out);

/* Full sort-based build (build_cell_list + materialize_pairlist).  Output
 * arrays are malloc'ed and owned by the caller (free with orc_free). */
typedef struct {
    int64_t nat, npairs, M;
    int64_t *perm, *cell_id, *cell_offsets, *wind;
    int64_t *i, *j, *S, *first;
    double *R;
    orc_geom g;
} orc_result;
int orc_neighbour_list(const double *X, int64_t nat, const double cell[9],
                       double cutoff, const int32_t pbc[3], int ti_bits,
                       int want_R, int nthreads, orc_result *out);
void orc_free(orc_result *r);

/* Legacy linked-list PairList (cell_list.jl:369-505): _fix_cell_,
 * _celllist_, _pairlist_, get_first, sort_neigs!.  Returns the possibly
 * modified X / cell in Xfix / cellfix (caller buffers of 3*nat / 9). */
int orc_legacy_pairlist(const double *X, int64_t nat, const double cell[9],
                        double cutoff, const int32_t pbc[3], int ti_bits,
                        int fixcell, double *Xfix, double *cellfix,
                        orc_result *out);

/* Brute force over all (i, j, S) with |S_k| bounded by the face heights;
 * same R / r^2 expression as the sort-based path. */
int orc_bruteforce(const double *X, int64_t nat, const double cell[9],
                   double cutoff, const int32_t pbc[3], orc_result *out);

#endif
