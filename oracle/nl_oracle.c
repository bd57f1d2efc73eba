/*
 * CPU restatement of NeighbourLists.jl (v0.6.2) neighbour-list construction.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): tests/, the smoke check and the
 * CPU-baseline leg of bench.py load this; the product library never does.
 *
 * Every function names the reference lines it restates.  Floating-point
 * expressions keep the reference's evaluation order: StaticArrays expands
 * 3x3 mat-vec products and dot products as left-to-right sums of products,
 * ((a1*b1 + a2*b2) + a3*b3), with no fused multiply-add (Julia only fuses
 * under muladd/@fastmath).  Built with -ffp-contract=off.
 *
 * What pins each function (tests/test_oracle.py): geometry against numpy's
 * inv/det/cross and the floor/ceil closed forms; cell ids against
 * floor(frac*n + 1) with wrap/clamp; the sort and offsets against their
 * invariants; the pair lists against the reference tests' known answers,
 * FCC coordination closed forms, each other (sort-based, legacy, brute
 * force) and scipy's cKDTree.  Parity unpinned at the ulp level against
 * Julia itself: StaticArrays' exact rounding of inv, lengths, inv_cell'*x
 * and C'*S cannot be compared without a Julia runtime; R is pinned to 1e-12
 * by R == X[j] - X[i] + C'*S.
 */
#include "nl_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* StaticArrays 3x3 algebra, in StaticArrays' own evaluation order.   */
/* ------------------------------------------------------------------ */

static inline void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static inline double dot3(const double a[3], const double b[3])
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

static inline double norm3(const double a[3])
{
    return sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
}

/* inv(::SMatrix{3,3}): columns x0,x1,x2; y0 = x1 x x2, d = x0.y0,
 * x0 /= d, y0 /= d, y1 = x2 x x0, y2 = x0 x x1; rows of the inverse are
 * y0, y1, y2. */
void orc_inv3(const double a[9], double out[9])
{
    double x0[3] = {a[0], a[1], a[2]};
    const double x1[3] = {a[3], a[4], a[5]};
    const double x2[3] = {a[6], a[7], a[8]};
    double y0[3], y1[3], y2[3];
    cross3(x1, x2, y0);
    double d = dot3(x0, y0);
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

/* det(::SMatrix{3,3}) = x0 . (x1 x x2) over the columns. */
double orc_det3(const double a[9])
{
    const double x0[3] = {a[0], a[1], a[2]};
    const double x1[3] = {a[3], a[4], a[5]};
    const double x2[3] = {a[6], a[7], a[8]};
    double y[3];
    cross3(x1, x2, y);
    return dot3(x0, y);
}

/* lengths (cell_list.jl:94-95): det(C) ./ (|C[2,:]xC[3,:]|, |C[3,:]xC[1,:]|,
 * |C[1,:]xC[2,:]|) -- rows of C are lattice vectors. */
void orc_lengths(const double c[9], double out[3])
{
    double r[3][3];
    for (int m = 0; m < 3; ++m)
        for (int k = 0; k < 3; ++k) r[m][k] = c[m + 3 * k];
    double d = orc_det3(c);
    double t[3];
    cross3(r[1], r[2], t);
    out[0] = d / norm3(t);
    cross3(r[2], r[0], t);
    out[1] = d / norm3(t);
    cross3(r[0], r[1], t);
    out[2] = d / norm3(t);
}

/* ------------------------------------------------------------------ */
/* analyze_cell (cell_list.jl:152-170), overflow check (:655-659),    */
/* nxyz and cutoff^2 (gpu_kernels.jl:315-317).                        */
/* ------------------------------------------------------------------ */

int orc_analyze(const double cell[9], double cutoff, const int32_t pbc[3],
                int ti_bits, orc_geom *g)
{
    memset(g, 0, sizeof(*g));
    memcpy(g->cell, cell, 9 * sizeof(double));
    if (!(cutoff > 0.0) || !isfinite(cutoff)) return ORC_EINVAL;
    for (int k = 0; k < 9; ++k)
        if (!isfinite(cell[k])) return ORC_EINVAL;
    g->cutoff = cutoff;
    g->cutoff_sq = cutoff * cutoff;
    for (int k = 0; k < 3; ++k) g->pbc[k] = pbc[k] ? 1 : 0;
    double volume = fabs(orc_det3(cell));
    g->zero_volume = volume < 1e-12;
    orc_inv3(cell, g->inv);
    double lens[3];
    orc_lengths(cell, lens);
    const double timax = ti_bits == 32 ? 2147483647.0 : 9223372036854775807.0;
    for (int k = 0; k < 3; ++k) {
        g->lens[k] = fabs(lens[k]);
        double t = floor(g->lens[k] / cutoff); /* floor.(TI, lens / cutoff) */
        if (isnan(t)) return ORC_EINVAL;
        if (t > timax) return ORC_EOVERFLOW;   /* InexactError */
        int64_t n = (int64_t)t;
        g->ncells[k] = n < 1 ? 1 : n;          /* max.(_t, one(TI)) */
    }
    __int128 prod = (__int128)g->ncells[0] * g->ncells[1] * g->ncells[2];
    __int128 tmax = ti_bits == 32 ? (__int128)2147483647 : (__int128)INT64_MAX;
    if (prod > tmax) return ORC_EOVERFLOW;
    g->ncells_total = (int64_t)prod;
    for (int k = 0; k < 3; ++k) {
        /* ceil.(TI, cutoff * (ncells ./ abs.(lens))) */
        double v = ceil(cutoff * ((double)g->ncells[k] / g->lens[k]));
        if (!isfinite(v) || v > timax) return ORC_EINVAL;
        g->nxyz[k] = (int64_t)v;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Cell-index algebra (cell_list.jl:33-147).                           */
/* ------------------------------------------------------------------ */

static inline int64_t floordiv(int64_t a, int64_t b)
{
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

/* wrap_and_shift (cell_list.jl:105-120); the while loops are the floor
 * division below. */
static inline void wrap_and_shift(int64_t i, int64_t n, int p, int64_t *w,
                                  int64_t *s)
{
    if (!p) {
        *w = i < 1 ? 1 : (i > n ? n : i);
        *s = 0;
        return;
    }
    if (i >= 1 && i <= n) { /* neither while loop runs */
        *w = i;
        *s = 0;
        return;
    }
    int64_t q = floordiv(i - 1, n);
    *w = i - q * n;
    *s = q;
}

/* position_to_cell_index (cell_list.jl:66-74): frac = inv_cell' * x,
 * c_k = trunc(floor(frac_k * n_k + 1)). */
static inline void position_to_cell_index(const orc_geom *g, const double *x,
                                          int64_t c[3])
{
    for (int k = 0; k < 3; ++k) {
        const double *iv = g->inv + 3 * k; /* column k of inv = row k of inv' */
        double f = (iv[0] * x[0] + iv[1] * x[1]) + iv[2] * x[2];
        c[k] = (int64_t)floor(f * (double)g->ncells[k] + 1.0);
    }
}

/* position_to_cell_index + bin_wrap_and_shift (cell_list.jl:141-147) */
static inline void cell_and_winding(const orc_geom *g, const double *x,
                                    int64_t c[3], int64_t w[3])
{
    int64_t c0[3];
    position_to_cell_index(g, x, c0);
    for (int k = 0; k < 3; ++k)
        wrap_and_shift(c0[k], g->ncells[k], g->pbc[k], &c[k], &w[k]);
}

static inline int64_t sub2ind(const int64_t n[3], const int64_t c[3])
{
    return c[0] + (c[1] - 1) * n[0] + (c[2] - 1) * n[0] * n[1];
}

/* ------------------------------------------------------------------ */
/* Sort-based build (cell_list.jl:647-748).                            */
/* ------------------------------------------------------------------ */

void orc_cell_ids(const double *X, int64_t nat, const orc_geom *g,
                  int64_t *ids, int64_t *wind)
{
    for (int64_t i = 0; i < nat; ++i) {
        int64_t c[3], w[3];
        cell_and_winding(g, X + 3 * i, c, w);
        ids[i] = sub2ind(g->ncells, c);
        if (wind)
            for (int k = 0; k < 3; ++k) wind[3 * i + k] = w[k];
    }
}

/* Stable counting sort: ties keep ascending original index, which is the
 * permutation Julia's stable sortperm returns. */
void orc_sortperm(const int64_t *ids, int64_t nat, int64_t M, int64_t *perm)
{
    int64_t *cnt = calloc((size_t)M + 2, sizeof(int64_t));
    for (int64_t i = 0; i < nat; ++i) cnt[ids[i]]++;
    int64_t run = 0;
    for (int64_t c = 0; c <= M + 1; ++c) {
        int64_t t = cnt[c];
        cnt[c] = run;
        run += t;
    }
    for (int64_t i = 0; i < nat; ++i) perm[cnt[ids[i]]++] = i + 1;
    free(cnt);
}

void orc_cell_offsets(const int64_t *sorted_ids, int64_t nat, int64_t M,
                      int64_t *offsets)
{
    memset(offsets, 0, (size_t)(M + 1) * sizeof(int64_t));
    if (nat == 0) {
        for (int64_t c = 0; c <= M; ++c) offsets[c] = 1;
        return;
    }
    for (int64_t p = 0; p < nat; ++p) offsets[sorted_ids[p]] += 1; /* cid+1, 1-based */
    offsets[0] = 1;
    for (int64_t c = 1; c <= M; ++c) offsets[c] += offsets[c - 1];
}

/* _for_each_neighbor_pair (gpu_kernels.jl:58-101) for atom i (0-based).
 * If j_out is NULL only counts.  Returns the neighbour count. */
static int64_t neighbour_loop(const double *X, const orc_geom *g,
                              const int64_t *perm, const int64_t *coff,
                              int64_t i, int64_t cap, int64_t *j_out,
                              int64_t *S_out, double *R_out)
{
    const double *xi = X + 3 * i;
    int64_t ci[3], wi[3];
    cell_and_winding(g, xi, ci, wi);
    const int64_t *n = g->ncells;
    const int64_t *nx = g->nxyz;
    const double *C = g->cell;
    int64_t cnt = 0;
    for (int64_t dz = -nx[2]; dz <= nx[2]; ++dz)
        for (int64_t dy = -nx[1]; dy <= nx[1]; ++dy)
            for (int64_t dx = -nx[0]; dx <= nx[0]; ++dx) {
                const int64_t d[3] = {dx, dy, dz};
                int ok = 1; /* _is_cell_in_bounds (gpu_kernels.jl:20-25) */
                for (int k = 0; k < 3; ++k)
                    if (!g->pbc[k] && !(1 <= ci[k] + d[k] && ci[k] + d[k] <= n[k])) ok = 0;
                if (!ok) continue;
                int64_t cj[3], s[3]; /* _get_neighbor_cell (:39-47) */
                for (int k = 0; k < 3; ++k)
                    wrap_and_shift(ci[k] + d[k], n[k], g->pbc[k], &cj[k], &s[k]);
                int64_t lin = sub2ind(n, cj);
                int64_t a = coff[lin - 1], b = coff[lin] - 1;
                for (int64_t idx = a; idx <= b; ++idx) {
                    int64_t j = perm[idx - 1]; /* 1-based original index */
                    if (j == i + 1 && s[0] == 0 && s[1] == 0 && s[2] == 0) continue;
                    const double *xj = X + 3 * (j - 1);
                    int64_t cjj[3], wj[3];
                    cell_and_winding(g, xj, cjj, wj);
                    int64_t S[3];
                    for (int k = 0; k < 3; ++k) S[k] = s[k] + wi[k] - wj[k];
                    double R[3];
                    for (int k = 0; k < 3; ++k) {
                        /* (C' * S)_k = sum_m C[m,k] S_m */
                        double cs = (C[0 + 3 * k] * (double)S[0] + C[1 + 3 * k] * (double)S[1]) +
                                    C[2 + 3 * k] * (double)S[2];
                        R[k] = (xj[k] - xi[k]) + cs;
                    }
                    double r2 = (R[0] * R[0] + R[1] * R[1]) + R[2] * R[2];
                    if (r2 < g->cutoff_sq) {
                        if (j_out && cnt < cap) {
                            j_out[cnt] = j;
                            for (int k = 0; k < 3; ++k) S_out[3 * cnt + k] = S[k];
                            if (R_out)
                                for (int k = 0; k < 3; ++k) R_out[3 * cnt + k] = R[k];
                        }
                        ++cnt;
                    }
                }
            }
    return cnt;
}

typedef struct {
    const double *X;
    const orc_geom *g;
    const int64_t *perm, *coff, *first;
    int64_t lo, hi;
    int64_t *counts;
    int64_t *i_out, *j_out, *S_out;
    double *R_out;
    int64_t total;
} work_t;

static void *count_worker(void *p)
{
    work_t *w = p;
    int64_t tot = 0;
    for (int64_t i = w->lo; i < w->hi; ++i) {
        int64_t c = neighbour_loop(w->X, w->g, w->perm, w->coff, i, 0, NULL, NULL, NULL);
        w->counts[i] = c;
        tot += c;
    }
    w->total = tot;
    return NULL;
}

static void *fill_worker(void *p)
{
    work_t *w = p;
    for (int64_t i = w->lo; i < w->hi; ++i) {
        int64_t o = w->first[i] - 1;
        int64_t cap = w->first[i + 1] - w->first[i];
        int64_t c = neighbour_loop(w->X, w->g, w->perm, w->coff, i, cap, w->j_out + o,
                                   w->S_out + 3 * o, w->R_out ? w->R_out + 3 * o : NULL);
        for (int64_t q = 0; q < c && q < cap; ++q) w->i_out[o + q] = i + 1;
    }
    return NULL;
}

static int64_t run_parallel(void *(*fn)(void *), work_t *tmpl, int64_t nat, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((int64_t)nthreads > nat) nthreads = nat > 0 ? (int)nat : 1;
    work_t w[256];
    pthread_t th[256];
    for (int t = 0; t < nthreads; ++t) {
        w[t] = *tmpl;
        w[t].lo = nat * t / nthreads;
        w[t].hi = nat * (t + 1) / nthreads;
        w[t].total = 0;
    }
    if (nthreads == 1) {
        fn(&w[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, &w[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    int64_t tot = 0;
    for (int t = 0; t < nthreads; ++t) tot += w[t].total;
    return tot;
}

int64_t orc_count(const double *X, int64_t nat, const orc_geom *g,
                  const int64_t *perm, const int64_t *cell_offsets,
                  int64_t *counts, int nthreads)
{
    work_t t = {0};
    t.X = X; t.g = g; t.perm = perm; t.coff = cell_offsets; t.counts = counts;
    return run_parallel(count_worker, &t, nat, nthreads);
}

void orc_pair_offsets(const int64_t *counts, int64_t nat, int64_t *first)
{
    first[0] = 1;
    for (int64_t i = 0; i < nat; ++i) first[i + 1] = first[i] + counts[i];
}

void orc_fill(const double *X, int64_t nat, const orc_geom *g,
              const int64_t *perm, const int64_t *cell_offsets,
              const int64_t *first, int64_t *i_out, int64_t *j_out,
              int64_t *S_out, double *R_out, int nthreads)
{
    work_t t = {0};
    t.X = X; t.g = g; t.perm = perm; t.coff = cell_offsets; t.first = first;
    t.i_out = i_out; t.j_out = j_out; t.S_out = S_out; t.R_out = R_out;
    run_parallel(fill_worker, &t, nat, nthreads);
}

int64_t orc_neighbours_of(const double *X, int64_t nat, const orc_geom *g,
                          const int64_t *perm, const int64_t *cell_offsets,
                          int64_t i, int64_t cap, int64_t *j_out,
                          int64_t *S_out, double *R_out)
{
    if (i < 1 || i > nat) return -1;
    return neighbour_loop(X, g, perm, cell_offsets, i - 1, cap, j_out, S_out, R_out);
}

void orc_free(orc_result *r)
{
    free(r->perm); free(r->cell_id); free(r->cell_offsets); free(r->wind);
    free(r->i); free(r->j); free(r->S); free(r->first); free(r->R);
    memset(r, 0, sizeof(*r));
}

int orc_neighbour_list(const double *X, int64_t nat, const double cell[9],
                       double cutoff, const int32_t pbc[3], int ti_bits,
                       int want_R, int nthreads, orc_result *out)
{
    memset(out, 0, sizeof(*out));
    int st = orc_analyze(cell, cutoff, pbc, ti_bits, &out->g);
    if (st) return st;
    const orc_geom *g = &out->g;
    int64_t M = g->ncells_total;
    out->nat = nat;
    out->M = M;
    size_t n1 = (size_t)(nat > 0 ? nat : 1);
    int64_t *ids = malloc(n1 * sizeof(int64_t));
    out->wind = malloc(3 * n1 * sizeof(int64_t));
    out->perm = malloc(n1 * sizeof(int64_t));
    out->cell_id = malloc(n1 * sizeof(int64_t));
    out->cell_offsets = malloc((size_t)(M + 1) * sizeof(int64_t));
    out->first = malloc((n1 + 1) * sizeof(int64_t));
    orc_cell_ids(X, nat, g, ids, out->wind);
    orc_sortperm(ids, nat, M, out->perm);
    for (int64_t p = 0; p < nat; ++p) out->cell_id[p] = ids[out->perm[p] - 1];
    orc_cell_offsets(out->cell_id, nat, M, out->cell_offsets);
    free(ids);
    int64_t *counts = malloc(n1 * sizeof(int64_t));
    int64_t P = orc_count(X, nat, g, out->perm, out->cell_offsets, counts, nthreads);
    orc_pair_offsets(counts, nat, out->first);
    free(counts);
    if (ti_bits == 32 && P > 2147483646) return ORC_EOVERFLOW;
    out->npairs = P;
    size_t p1 = (size_t)(P > 0 ? P : 1);
    out->i = malloc(p1 * sizeof(int64_t));
    out->j = malloc(p1 * sizeof(int64_t));
    out->S = malloc(3 * p1 * sizeof(int64_t));
    out->R = want_R ? malloc(3 * p1 * sizeof(double)) : NULL;
    orc_fill(X, nat, g, out->perm, out->cell_offsets, out->first, out->i, out->j, out->S,
             out->R, nthreads);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Legacy linked-list path (cell_list.jl:185-505).                     */
/* ------------------------------------------------------------------ */

/* Julia mod(x::Float64, 1.0): r = rem(x, y); r == 0 -> copysign(r, y);
 * sign(r) != sign(y) -> r + y; else r. */
static double jl_mod1(double x)
{
    double r = fmod(x, 1.0);
    if (r == 0.0) return copysign(r, 1.0);
    if ((r > 0) != (1.0 > 0)) return r + 1.0;
    return r;
}

/* _fix_cell_ (cell_list.jl:459-505) */
static void fix_cell(double *X, int64_t nat, double *C, const int32_t pbc[3])
{
    double inv[9];
    orc_inv3(C, inv);
    double min_lam[3] = {0, 0, 0}, max_lam[3] = {1, 1, 1};
    for (int64_t ix = 0; ix < nat; ++ix) {
        double *x = X + 3 * ix;
        double lam[3];
        for (int k = 0; k < 3; ++k) /* invCt * x, invCt = inv(C)' */
            lam[k] = (inv[0 + 3 * k] * x[0] + inv[1 + 3 * k] * x[1]) + inv[2 + 3 * k] * x[2];
        int upd = 0;
        for (int i = 0; i < 3; ++i) {
            if (!(0.0 <= lam[i] && lam[i] < 1.0)) {
                if (pbc[i]) {
                    lam[i] = jl_mod1(lam[i]);
                    upd = 1;
                } else {
                    if (lam[i] < min_lam[i]) min_lam[i] = lam[i];
                    if (lam[i] > max_lam[i]) max_lam[i] = lam[i];
                }
            }
        }
        if (upd) { /* X[ix] = C' * SVec(lam) */
            double nx[3];
            for (int k = 0; k < 3; ++k)
                nx[k] = (C[0 + 3 * k] * lam[0] + C[1 + 3 * k] * lam[1]) + C[2 + 3 * k] * lam[2];
            memcpy(x, nx, sizeof nx);
        }
    }
    double mn = fmin(fmin(min_lam[0], min_lam[1]), min_lam[2]);
    double mx = fmax(fmax(max_lam[0], max_lam[1]), max_lam[2]);
    if (mn < 0.0 || mx > 1.0) {
        double t[3]; /* t = -C' * min_lam */
        for (int k = 0; k < 3; ++k)
            t[k] = ((-C[0 + 3 * k]) * min_lam[0] + (-C[1 + 3 * k]) * min_lam[1]) +
                   (-C[2 + 3 * k]) * min_lam[2];
        for (int64_t n = 0; n < nat; ++n)
            for (int k = 0; k < 3; ++k) X[3 * n + k] += t[k];
        for (int i = 0; i < 3; ++i) {
            max_lam[i] -= min_lam[i];
            min_lam[i] = 0.0;
        }
        for (int i = 0; i < 3; ++i)
            if (max_lam[i] > 1) max_lam[i] *= 1.01;
        double Cn[9];
        for (int m = 0; m < 3; ++m) /* row m scaled by max_lam[m] */
            for (int k = 0; k < 3; ++k) Cn[m + 3 * k] = max_lam[m] * C[m + 3 * k];
        memcpy(C, Cn, sizeof Cn);
    }
}

typedef struct {
    int64_t n, cap;
    int64_t *i, *j, *S;
} pvec;

static void pvec_push(pvec *v, int64_t i, int64_t j, const int64_t S[3])
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->i = realloc(v->i, (size_t)v->cap * sizeof(int64_t));
        v->j = realloc(v->j, (size_t)v->cap * sizeof(int64_t));
        v->S = realloc(v->S, 3 * (size_t)v->cap * sizeof(int64_t));
    }
    v->i[v->n] = i;
    v->j[v->n] = j;
    memcpy(v->S + 3 * v->n, S, 3 * sizeof(int64_t));
    v->n++;
}

static inline int64_t trunc_div(int64_t a, int64_t b) { return a / b; } /* Julia ÷ */

/* bins * v with bins column k = C[k,:] / ns_k (cell_list.jl:263) */
static inline void bins_mul(const double bins[9], const double v[3], double o[3])
{
    for (int m = 0; m < 3; ++m)
        o[m] = (bins[m + 0] * v[0] + bins[m + 3] * v[1]) + bins[m + 6] * v[2];
}

int orc_legacy_pairlist(const double *X0, int64_t nat, const double cell0[9],
                        double cutoff, const int32_t pbc[3], int ti_bits,
                        int fixcell, double *X, double *C, orc_result *out)
{
    memset(out, 0, sizeof(*out));
    memcpy(X, X0, (size_t)nat * 3 * sizeof(double));
    memcpy(C, cell0, 9 * sizeof(double));
    if (fixcell) fix_cell(X, nat, C, pbc);
    int st = orc_analyze(C, cutoff, pbc, ti_bits, &out->g);
    if (st) return st;
    const orc_geom *g = &out->g;
    const int64_t *ns = g->ncells;
    int64_t M = g->ncells_total;
    out->nat = nat;
    out->M = M;
    /* _celllist_ (cell_list.jl:185-230): linked list per bin */
    int64_t *seed = malloc((size_t)M * sizeof(int64_t));
    int64_t *last = malloc((size_t)M * sizeof(int64_t));
    int64_t *next = malloc((size_t)(nat > 0 ? nat : 1) * sizeof(int64_t));
    for (int64_t c = 0; c < M; ++c) seed[c] = -1;
    for (int64_t i = 0; i < nat; ++i) {
        int64_t c0[3], c[3], w;
        position_to_cell_index(g, X + 3 * i, c0);
        for (int k = 0; k < 3; ++k) wrap_and_shift(c0[k], ns[k], pbc[k], &c[k], &w);
        int64_t ci = sub2ind(ns, c) - 1;
        next[i] = -1;
        if (seed[ci] < 0) {
            seed[ci] = i;
            last[ci] = i;
        } else {
            next[last[ci]] = i;
            last[ci] = i;
        }
    }
    /* _pairlist_ (cell_list.jl:233-297) */
    double bins[9];
    for (int k = 0; k < 3; ++k)
        for (int m = 0; m < 3; ++m) bins[m + 3 * k] = C[k + 3 * m] / (double)ns[k];
    const int64_t *nx = g->nxyz;
    pvec v = {0};
    for (int64_t i = 0; i < nat; ++i) { /* _find_neighbours_! (:301-366) */
        const double *xi = X + 3 * i;
        int64_t ci0[3], ci[3];
        position_to_cell_index(g, xi, ci0);
        for (int k = 0; k < 3; ++k) /* bin_trunc.(ci0, pbc, ns) */
            ci[k] = pbc[k] ? ci0[k] : (ci0[k] <= 0 ? 1 : (ci0[k] > ns[k] ? ns[k] : ci0[k]));
        double cm1[3], bx[3], dxi[3];
        for (int k = 0; k < 3; ++k) cm1[k] = (double)(ci[k] - 1);
        bins_mul(bins, cm1, bx);
        for (int k = 0; k < 3; ++k) dxi[k] = xi[k] - bx[k];
        for (int k = 0; k < 3; ++k) { /* bin_wrap_or_trunc.(ci0, pbc, ns) */
            int64_t s;
            wrap_and_shift(ci0[k], ns[k], pbc[k], &ci[k], &s);
        }
        for (int64_t z = -nx[2]; z <= nx[2]; ++z)
            for (int64_t y = -nx[1]; y <= nx[1]; ++y)
                for (int64_t x = -nx[0]; x <= nx[0]; ++x) {
                    const int64_t xyz[3] = {x, y, z};
                    int64_t cj[3];
                    int ok = 1;
                    for (int k = 0; k < 3; ++k) {
                        int64_t t = ci[k] + xyz[k], s;
                        if (pbc[k]) wrap_and_shift(t, ns[k], 1, &cj[k], &s);
                        else cj[k] = t;
                        if (!(1 <= cj[k] && cj[k] <= ns[k])) ok = 0;
                    }
                    if (!ok) continue;
                    int64_t ncj = sub2ind(ns, cj) - 1;
                    double fx[3], off[3];
                    for (int k = 0; k < 3; ++k) fx[k] = (double)xyz[k];
                    bins_mul(bins, fx, off);
                    for (int64_t j = seed[ncj]; j >= 0; j = next[j]) {
                        if (!(i != j || x != 0 || y != 0 || z != 0)) continue;
                        const double *xj = X + 3 * j;
                        int64_t cjj[3];
                        position_to_cell_index(g, xj, cjj);
                        for (int k = 0; k < 3; ++k)
                            cjj[k] = pbc[k] ? cjj[k]
                                            : (cjj[k] <= 0 ? 1 : (cjj[k] > ns[k] ? ns[k] : cjj[k]));
                        double c2[3], bj[3], d[3];
                        for (int k = 0; k < 3; ++k) c2[k] = (double)(cjj[k] - 1);
                        bins_mul(bins, c2, bj);
                        for (int k = 0; k < 3; ++k) d[k] = ((xj[k] - bj[k]) - dxi[k]) + off[k];
                        double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
                        if (r2 < g->cutoff_sq) {
                            int64_t S[3];
                            for (int k = 0; k < 3; ++k) S[k] = trunc_div(ci0[k] - cjj[k] + xyz[k], ns[k]);
                            pvec_push(&v, i + 1, j + 1, S);
                        }
                    }
                }
    }
    free(seed); free(last); free(next);
    int64_t P = v.n;
    out->npairs = P;
    out->i = v.i ? v.i : malloc(sizeof(int64_t));
    out->j = v.j ? v.j : malloc(sizeof(int64_t));
    out->S = v.S ? v.S : malloc(3 * sizeof(int64_t));
    /* get_first (cell_list.jl:401-418) */
    out->first = malloc((size_t)(nat + 1) * sizeof(int64_t));
    {
        int64_t idx = 1, n = 1;
        while (n <= nat && idx <= P) {
            out->first[n - 1] = idx;
            while (out->i[idx - 1] == n) {
                idx++;
                if (idx > P) break;
            }
            n++;
        }
        for (int64_t q = n; q <= nat + 1; ++q) out->first[q - 1] = P + 1;
    }
    /* sort_neigs! (cell_list.jl:427-446): stable sort of j within each site */
    for (int64_t n = 0; n < nat; ++n) {
        int64_t a = out->first[n] - 1, b = out->first[n + 1] - 1;
        /* insertion sort is stable */
        for (int64_t p = a + 1; p < b; ++p) {
            int64_t jj = out->j[p];
            int64_t SS[3] = {out->S[3 * p], out->S[3 * p + 1], out->S[3 * p + 2]};
            int64_t q = p - 1;
            while (q >= a && out->j[q] > jj) {
                out->j[q + 1] = out->j[q];
                memcpy(out->S + 3 * (q + 1), out->S + 3 * q, 3 * sizeof(int64_t));
                --q;
            }
            out->j[q + 1] = jj;
            memcpy(out->S + 3 * (q + 1), SS, 3 * sizeof(int64_t));
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Brute force: every (i, j, S), S_k = 0 on open axes and, on periodic */
/* axes, every integer with |frac_j - frac_i + S_k| h_k possibly < rc. */
/* ------------------------------------------------------------------ */

int orc_bruteforce(const double *X, int64_t nat, const double cell[9],
                   double cutoff, const int32_t pbc[3], orc_result *out)
{
    memset(out, 0, sizeof(*out));
    int st = orc_analyze(cell, cutoff, pbc, 64, &out->g);
    if (st) return st;
    const orc_geom *g = &out->g;
    const double *C = cell;
    out->nat = nat;
    double *frac = malloc((size_t)(nat > 0 ? nat : 1) * 3 * sizeof(double));
    for (int64_t i = 0; i < nat; ++i)
        for (int k = 0; k < 3; ++k)
            frac[3 * i + k] = (g->inv[0 + 3 * k] * X[3 * i] + g->inv[1 + 3 * k] * X[3 * i + 1]) +
                              g->inv[2 + 3 * k] * X[3 * i + 2];
    int64_t K[3];
    for (int k = 0; k < 3; ++k) K[k] = (int64_t)ceil(cutoff / g->lens[k]) + 1;
    pvec v = {0};
    for (int64_t i = 0; i < nat; ++i)
        for (int64_t j = 0; j < nat; ++j) {
            int64_t lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                if (!pbc[k]) {
                    lo[k] = hi[k] = 0;
                } else {
                    double d = frac[3 * i + k] - frac[3 * j + k];
                    lo[k] = (int64_t)floor(d) - K[k];
                    hi[k] = (int64_t)ceil(d) + K[k];
                }
            }
            for (int64_t s0 = lo[0]; s0 <= hi[0]; ++s0)
                for (int64_t s1 = lo[1]; s1 <= hi[1]; ++s1)
                    for (int64_t s2 = lo[2]; s2 <= hi[2]; ++s2) {
                        if (i == j && s0 == 0 && s1 == 0 && s2 == 0) continue;
                        const int64_t S[3] = {s0, s1, s2};
                        double R[3];
                        for (int k = 0; k < 3; ++k) {
                            double cs = (C[0 + 3 * k] * (double)S[0] + C[1 + 3 * k] * (double)S[1]) +
                                        C[2 + 3 * k] * (double)S[2];
                            R[k] = (X[3 * j + k] - X[3 * i + k]) + cs;
                        }
                        double r2 = (R[0] * R[0] + R[1] * R[1]) + R[2] * R[2];
                        if (r2 < g->cutoff_sq) pvec_push(&v, i + 1, j + 1, S);
                    }
        }
    free(frac);
    out->npairs = v.n;
    out->i = v.i ? v.i : malloc(sizeof(int64_t));
    out->j = v.j ? v.j : malloc(sizeof(int64_t));
    out->S = v.S ? v.S : malloc(3 * sizeof(int64_t));
    return ORC_OK;
}
