/*
 * nlhip device-side lazy neighbour access for user HIP kernels.
 *
 * The MI355X counterpart of passing a SortedCellList into a user kernel and
 * calling for_each_neighbour there (NeighbourLists.jl src/adapt.jl:21-34,
 * src/cell_list.jl:779-801, exercised by test/test_gpu.jl:94-122).  The
 * view points at the device arrays a handle owns after nlh_build_cell_list
 * (nlh_celllist_device) and carries the geometry by value.
 *
 *   __global__ void k(nlh_celllist_dev<int32_t> c, int32_t *out) {
 *       int64_t i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // 1-based
 *       if (i > c.nat) return;
 *       int32_t n = 0;
 *       nlh_for_each_neighbour(c, i, [&](int32_t j, const double *R, const int32_t *S) { ++n; });
 *       out[i - 1] = n;
 *   }
 *
 * The loop is the reference's _for_each_neighbor_pair
 * (src/gpu_kernels.jl:58-101): neighbour cells in (dz, dy, dx) order, atoms
 * of a cell in sorted order, S = s_loop + w_i - w_j, R = (x_j - x_i) + C' S,
 * r^2 = (R1^2 + R2^2) + R3^2 < cutoff^2 with the same non-fused evaluation
 * order as the materialised pair list (compile with -ffp-contract=off), so
 * the neighbours and R match nlh_materialize bit for bit.
 */
#ifndef NLHIP_DEVICE_H
#define NLHIP_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nlhip.h"

template <typename TI>
struct nlh_celllist_dev {
    const double *X_orig;    /* nat x 3 AoS, unsorted input positions   */
    const TI *perm;          /* sorted index -> original index, 1-based */
    const TI *cell_offsets;  /* ncells_total + 1, 1-based CSR           */
    double cell[9];          /* column-major, rows = lattice vectors    */
    double inv_cell[9];
    double cutoff_sq;
    int32_t ncells[3];
    int32_t nxyz[3];
    int32_t pbc;             /* bit k: periodic along lattice vector k */
    int64_t nat;
};

/* Host helper: device view from a handle's cell-list view and geometry. */
template <typename TI>
inline nlh_celllist_dev<TI> nlh_make_celllist_dev(const nlh_celllist_view &v, const nlh_geometry &g)
{
    nlh_celllist_dev<TI> c;
    c.X_orig = v.X_orig;
    c.perm = static_cast<const TI *>(v.perm);
    c.cell_offsets = static_cast<const TI *>(v.cell_offsets);
    for (int q = 0; q < 9; ++q) {
        c.cell[q] = g.cell[q];
        c.inv_cell[q] = g.inv_cell[q];
    }
    c.cutoff_sq = g.cutoff_sq;
    for (int k = 0; k < 3; ++k) {
        c.ncells[k] = (int32_t)g.ncells[k];
        c.nxyz[k] = (int32_t)g.nxyz[k];
    }
    c.pbc = (g.pbc[0] ? 1 : 0) | (g.pbc[1] ? 2 : 0) | (g.pbc[2] ? 4 : 0);
    c.nat = v.nat;
    return c;
}

namespace nlh_device_detail {

/* wrap_and_shift (src/cell_list.jl:105-120) */
__device__ inline void wrap_and_shift(int32_t i, int32_t n, bool p, int32_t &w, int32_t &s)
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
    const int32_t a = i - 1;
    int32_t q = a / n;
    if (a % n != 0 && a < 0) --q;
    w = i - q * n;
    s = q;
}

/* position_to_cell_index + bin_wrap_and_shift (src/cell_list.jl:66-74, 141-147).
 * The reference's unsafe_trunc is undefined for values that do not fit the
 * index type; here a cell index beyond +-1e9 (or NaN) becomes cell 1, the
 * same value the library's binning uses -- a build rejects such positions
 * (NLH_EINVAL), so a cell list that built never holds one. */
template <typename TI>
__device__ inline void cell_and_winding(const nlh_celllist_dev<TI> &c, const double *x, int32_t ci[3], int32_t w[3])
{
    for (int k = 0; k < 3; ++k) {
        const double f = (c.inv_cell[0 + 3 * k] * x[0] + c.inv_cell[1 + 3 * k] * x[1]) + c.inv_cell[2 + 3 * k] * x[2];
        const double v = floor(f * (double)c.ncells[k] + 1.0);
        const int32_t c0 = (v > -1.0e9 && v < 1.0e9) ? (int32_t)v : 1;
        wrap_and_shift(c0, c.ncells[k], (c.pbc >> k) & 1, ci[k], w[k]);
    }
}

}  // namespace nlh_device_detail

/* Calls f(TI j, const double R[3], const TI S[3]) for every neighbour j
 * (1-based) of atom i (1-based), in the reference traversal order. */
template <typename TI, typename F>
__device__ inline void nlh_for_each_neighbour(const nlh_celllist_dev<TI> &c, int64_t i, F &&f)
{
    using namespace nlh_device_detail;
    const double *xi = c.X_orig + 3 * (i - 1);
    int32_t ci[3], wi[3];
    cell_and_winding(c, xi, ci, wi);
    const int32_t n0 = c.ncells[0], n1 = c.ncells[1], n2 = c.ncells[2];
    for (int32_t dz = -c.nxyz[2]; dz <= c.nxyz[2]; ++dz)
        for (int32_t dy = -c.nxyz[1]; dy <= c.nxyz[1]; ++dy)
            for (int32_t dx = -c.nxyz[0]; dx <= c.nxyz[0]; ++dx) {
                const int32_t d[3] = {dx, dy, dz};
                bool inb = true;
                for (int k = 0; k < 3; ++k)
                    if (!((c.pbc >> k) & 1) && !(ci[k] + d[k] >= 1 && ci[k] + d[k] <= c.ncells[k])) inb = false;
                if (!inb) continue;
                int32_t cj[3], s[3];
                for (int k = 0; k < 3; ++k) wrap_and_shift(ci[k] + d[k], c.ncells[k], (c.pbc >> k) & 1, cj[k], s[k]);
                const int64_t lin = cj[0] + (int64_t)(cj[1] - 1) * n0 + (int64_t)(cj[2] - 1) * n0 * n1;
                const int64_t a = (int64_t)c.cell_offsets[lin - 1], b = (int64_t)c.cell_offsets[lin] - 1;
                (void)n2;
                for (int64_t idx = a; idx <= b; ++idx) {
                    const TI j = c.perm[idx - 1];
                    if ((int64_t)j == i && s[0] == 0 && s[1] == 0 && s[2] == 0) continue;
                    const double *xj = c.X_orig + 3 * ((int64_t)j - 1);
                    int32_t cjj[3], wj[3];
                    cell_and_winding(c, xj, cjj, wj);
                    TI S[3];
                    double R[3];
                    for (int k = 0; k < 3; ++k) S[k] = (TI)(s[k] + wi[k] - wj[k]);
                    for (int k = 0; k < 3; ++k) {
                        const double cs = (c.cell[0 + 3 * k] * (double)S[0] + c.cell[1 + 3 * k] * (double)S[1]) +
                                          c.cell[2 + 3 * k] * (double)S[2];
                        R[k] = (xj[k] - xi[k]) + cs;
                    }
                    const double r2 = (R[0] * R[0] + R[1] * R[1]) + R[2] * R[2];
                    if (r2 < c.cutoff_sq) f(j, R, S);
                }
            }
}

#endif /* NLHIP_DEVICE_H */
