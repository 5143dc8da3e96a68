/*
 * oracle_grid.c -- CPU restatement of the MiniCell grid build, cell-pair
 * enumeration and exclusion bitmasks of the reference (src/celeste/core/MiniCell.cpp).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 *
 * Compiled with -ffp-contract=off: float expressions round exactly where the
 * reference host build (x86-64 SSE, ISO C++) rounds them.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NAC ORC_N_ATOM_CELL

/* PBC::set_pbc (PBC.cpp:18-41): L is float, L_half = L*0.5, lb = c - L_half, ub = lb + L. */
void orc_pbc_from_box(const double val[12], float L[3], float lb[3], float ub[3]) {
    L[0] = (float)val[0];
    L[1] = (float)val[4];
    L[2] = (float)val[8];
    for (int i = 0; i < 3; i++) {
        float L_half = (float)(L[i] * 0.5);
        lb[i]        = (float)(val[9 + i] - (double)L_half);
        ub[i]        = lb[i] + L[i];
    }
}

/* MiniCell::set_grid_parameters (MiniCell.cpp:1279-1303), set_box_info (:1327-1362)
 * and set_grid_xy (:221-290) for a single box (box_div 1 1 1). */
int orc_grid_setup(orc_grid_t *g, int n_atoms, const double box[12], double nsgrid_cutoff) {
    float ub[3];
    memset(g, 0, sizeof(*g));
    orc_pbc_from_box(box, g->L, g->lb, ub);
    g->cutoff_pair      = (float)nsgrid_cutoff;
    g->cutoff_pair_half = (float)(g->cutoff_pair * 0.5);
    g->cutoff_pair2     = g->cutoff_pair * g->cutoff_pair;
    g->n_uni_z          = (int)(g->L[2] / g->cutoff_pair_half);
    g->L_z_uni          = g->L[2] / (float)g->n_uni_z;

    /* exbox_l == box_l == PBC::L for one box (MiniCell.cpp:1341-1356) */
    float exl[3] = {g->L[0], g->L[1], g->L[2]};
    float volume = exl[0] * exl[1] * exl[2];
    float w      = powf((float)((float)NAC / ((float)n_atoms / volume)), (float)(1.0 / 3.0));
    if (w > g->cutoff_pair_half) w = g->cutoff_pair_half;
    for (int d = 0; d < 2; d++) {
        g->n_cells_xyz[d]     = (int)floorf(exl[d] / w);
        g->L_cell_xyz[d]      = exl[d] / (float)g->n_cells_xyz[d];
        g->n_neighbors_xyz[d] = (int)ceilf(g->cutoff_pair / g->L_cell_xyz[d]);
    }
    g->n_columns = g->n_cells_xyz[0] * g->n_cells_xyz[1];
    /* capacity heuristics: our own, generous (the reference's depend on an
     * uninitialised member, MiniCell.cpp:1386-1398 + :244-250). */
    g->max_n_atoms_column = n_atoms;
    g->max_n_cells        = n_atoms / NAC + g->n_columns + 1;
    g->max_n_cell_pairs   = (int)((double)(2 * g->n_neighbors_xyz[0] + 1) * (2 * g->n_neighbors_xyz[1] + 1)
                                  * (g->n_neighbors_xyz[0] + g->n_neighbors_xyz[1] + 2) * 0.5 * g->max_n_cells
                                  * 2.5);
    if (g->n_neighbors_xyz[0] > g->n_cells_xyz[0] || g->n_neighbors_xyz[1] > g->n_cells_xyz[1]) return -1;
    if (g->n_uni_z < 2) return -2;
    return 0;
}

int orc_cells_alloc(orc_cells_t *c, const orc_grid_t *g, int n_atoms) {
    memset(c, 0, sizeof(*c));
    c->cap_atom_array   = n_atoms + g->n_columns * NAC;
    c->cap_cells        = c->cap_atom_array / NAC;
    c->atomids          = (int *)malloc(sizeof(int) * c->cap_atom_array);
    c->crd              = (double *)calloc((size_t)c->cap_atom_array * 3, sizeof(double));
    c->cell_crd         = (int *)malloc(sizeof(int) * c->cap_cells * 3);
    c->idx_xy_head_cell = (int *)malloc(sizeof(int) * (g->n_columns + 1));
    c->idx_xy_head_atom = (int *)malloc(sizeof(int) * (g->n_columns + 1));
    c->n_atoms_col      = (int *)malloc(sizeof(int) * g->n_columns);
    c->n_cells_z        = (int *)malloc(sizeof(int) * g->n_columns);
    c->column_of_atom   = (int *)malloc(sizeof(int) * n_atoms);
    return (c->atomids && c->crd && c->cell_crd) ? 0 : -1;
}

void orc_cells_free(orc_cells_t *c) {
    free(c->atomids);
    free(c->crd);
    free(c->cell_crd);
    free(c->idx_xy_head_cell);
    free(c->idx_xy_head_atom);
    free(c->n_atoms_col);
    free(c->n_cells_z);
    free(c->column_of_atom);
    memset(c, 0, sizeof(*c));
}

/* MiniCell::get_column_id_from_crd (MiniCell.cpp:1250-1258): periodic wrap of a column index. */
static int column_id(const orc_grid_t *g, int x, int y) {
    while (x < 0) x += g->n_cells_xyz[0];
    while (x >= g->n_cells_xyz[0]) x -= g->n_cells_xyz[0];
    while (y < 0) y += g->n_cells_xyz[1];
    while (y >= g->n_cells_xyz[1]) y -= g->n_cells_xyz[1];
    return x + y * g->n_cells_xyz[0];
}

/* Hoare partition quicksort on the float z key (MiniCell.cpp:636-691, swap :618-635). */
static void qsort_column(double *crd, int *atomids, int l, int r) {
    if (r <= l) return;
    float v = (float)crd[r * 3 + 2];
    int   i = l - 1, j = r + 1;
    for (;;) {
        while (crd[(++i) * 3 + 2] < v) {
        }
        while (crd[(--j) * 3 + 2] > v && j > i) {
        }
        if (i >= j) break;
        int t      = atomids[i];
        atomids[i] = atomids[j];
        atomids[j] = t;
        for (int d = 0; d < 3; d++) {
            float x            = (float)crd[i * 3 + d];
            crd[i * 3 + d]     = crd[j * 3 + d];
            crd[j * 3 + d]     = x;
        }
    }
    qsort_column(crd, atomids, l, i - 1);
    qsort_column(crd, atomids, i, r);
}

int orc_cells_build(orc_cells_t *c, const orc_grid_t *g, int n_atoms, const double *in_crd) {
    const int ncol = g->n_columns;
    const int cap  = c->cap_atom_array;
    int      *col  = (int *)malloc(sizeof(int) * cap); /* idx_atom_cell_xy */
    if (!col) return -1;

    /* set_crds_to_homebox (MiniCell.cpp:1380-1400): grid slots start in atom order. */
    for (int i = 0; i < cap; i++) col[i] = -1;
    for (int i = 0; i < ncol; i++) c->n_atoms_col[i] = 0;
    for (int i = 0; i < n_atoms; i++)
        for (int d = 0; d < 3; d++) c->crd[i * 3 + d] = in_crd[i * 3 + d];

    /* set_atoms_into_grid_xy binning (MiniCell.cpp:316-359) */
    for (int a = 0; a < n_atoms; a++) {
        int cxy[2];
        for (int d = 0; d < 2; d++) {
            float x = (float)(c->crd[a * 3 + d] - (double)g->lb[d]);
            while (x < 0.0f) x += g->L[d];
            while (x >= g->L[d]) x -= g->L[d];
            cxy[d] = (int)floorf(x / g->L_cell_xyz[d]);
        }
        int cid = column_id(g, cxy[0], cxy[1]);
        col[a]  = cid;
        c->column_of_atom[a] = cid;
        c->n_atoms_col[cid]++;
    }

    /* set_idx_xy_head (MiniCell.cpp:785-861) */
    int n_cells = 0, atom_index = 0;
    for (int k = 0; k < ncol; k++) {
        int cx = k % g->n_cells_xyz[0], cy = k / g->n_cells_xyz[0];
        c->idx_xy_head_cell[k] = n_cells;
        c->idx_xy_head_atom[k] = atom_index;
        c->n_cells_z[k]        = (c->n_atoms_col[k] + NAC - 1) / NAC;
        if (c->n_atoms_col[k] == 0) {
            c->idx_xy_head_cell[k] = -1;
            c->idx_xy_head_atom[k] = -1;
        }
        for (int z = 0; z < c->n_cells_z[k]; z++) {
            int cell             = n_cells + z;
            c->cell_crd[cell * 3 + 0] = cx;
            c->cell_crd[cell * 3 + 1] = cy;
            c->cell_crd[cell * 3 + 2] = z;
        }
        n_cells += c->n_cells_z[k];
        atom_index += c->n_cells_z[k] * NAC;
    }
    c->idx_xy_head_cell[ncol] = n_cells;
    c->idx_xy_head_atom[ncol] = atom_index;
    c->n_cells                = n_cells;
    c->n_atom_array           = n_cells * NAC;

    /* atomids from buf (MiniCell.cpp:871-879 + :366-368): identity then -1. */
    for (int i = 0; i < cap; i++) c->atomids[i] = (i < n_atoms) ? i : -1;

    /* reorder_atominfo_for_columns (MiniCell.cpp:404-501): in-place cycle-following
     * permutation that moves each atom to the next free slot of its column, rounding
     * its coordinates through float on the way. */
    int *const head = c->idx_xy_head_atom;
    int *const nac  = c->n_atoms_col;
    for (int k = 0; k < ncol; k++) nac[k] = 0;
    float t1x[3], t2x[3];
    int   t1id, t2id, t1col, t2col;
    int   ag  = 0;
    t2col     = col[ag];
    int to    = head[t2col] + nac[t2col];
    t2id      = c->atomids[ag];
    for (int d = 0; d < 3; d++) t2x[d] = (float)c->crd[ag * 3 + d];
    col[ag] = -1;
    for (int nre = 1; nre < n_atoms; nre++) {
        for (int d = 0; d < 3; d++) {
            t1x[d]            = (float)c->crd[to * 3 + d];
            c->crd[to * 3 + d] = t2x[d];
        }
        t1id           = c->atomids[to];
        c->atomids[to] = t2id;
        t1col          = col[to];
        col[to]        = t2col;
        nac[t2col]++;
        ag = to;
        if (ag >= n_atoms || t1col == -1) {
            int seek = 1;
            for (int k = 0; seek && k < ncol; k++) {
                if (head[k] < 0) continue;
                int k2 = k + 1;
                while (head[k2] < 0) k2++;
                for (int s = head[k] + nac[k]; seek && s < head[k2]; s++) {
                    if (col[s] >= 0) {
                        ag   = s;
                        seek = 0;
                    }
                }
            }
            t1col   = col[ag];
            col[ag] = -1;
            for (int d = 0; d < 3; d++) t1x[d] = (float)c->crd[ag * 3 + d];
            t1id = c->atomids[ag];
        }
        to    = head[t1col] + nac[t1col];
        t2col = t1col;
        for (int d = 0; d < 3; d++) t2x[d] = t1x[d];
        t2id = t1id;
    }
    for (int d = 0; d < 3; d++) c->crd[to * 3 + d] = t2x[d];
    col[to] = t2col;
    nac[t2col]++;
    c->atomids[to] = t2id;

    /* set_dummy_atoms (MiniCell.cpp:503-530) */
    for (int k = 0; k < ncol; k++) {
        if (head[k] < 0) continue;
        int k2 = k + 1;
        while (head[k2] < 0) k2++;
        for (int i = head[k] + nac[k]; i < head[k2]; i++) c->atomids[i] = -1;
    }

    /* reorder_atominfo_in_columns (MiniCell.cpp:531-613): z sort and padding
     * with copies of the column's last real atom. */
    for (int k = 0; k < ncol; k++) {
        qsort_column(c->crd, c->atomids, head[k], head[k] + nac[k] - 1);
        int last = head[k] + nac[k] - 1;
        for (int i = head[k] + nac[k]; i < head[k] + c->n_cells_z[k] * NAC; i++) {
            for (int d = 0; d < 3; d++) c->crd[i * 3 + d] = c->crd[last * 3 + d];
            c->atomids[i] = -1;
        }
    }
    free(col);
    return 0;
}

/* ---- uniform z grid + enumeration ---- */

static int uni_id(const orc_grid_t *g, int x, int y, int z) {
    /* MiniCell::get_uni_id_from_crd (MiniCell.cpp:1596-1598) */
    return x * g->n_uni_z + y * (g->n_uni_z * g->n_cells_xyz[0]) + z;
}

/* MiniCell::check_valid_pair (MiniCell.cpp:1150-1160): keeps one of (c1,c2),(c2,c1). */
static int valid_pair(int c1, int c2) {
    int o1 = (c1 % 2) != 0, o2 = (c2 % 2) != 0;
    if (o1) {
        if ((c2 < c1 && !o2) || (c2 > c1 && o2)) return 0;
    } else {
        if ((c2 < c1 && o2) || (c2 > c1 && !o2)) return 0;
    }
    return 1;
}

/* MiniCell::set_cell_pair_bitmask (MiniCell.cpp:1198-1237): bit a2*8+a1 set = skip. */
static void cell_pair_mask(const orc_cells_t *c, const int *nb15off, int c1, int c2, int32_t mask[2]) {
    mask[0] = mask[1] = 0;
    for (int i1 = 0; i1 < NAC; i1++) {
        int a1   = c1 * NAC + i1;
        int id1  = c->atomids[a1];
        for (int i2 = 0; i2 < NAC; i2++) {
            int a2  = c2 * NAC + i2;
            int id2 = c->atomids[a2];
            int skip = 0;
            if (id1 == -1 || id2 == -1) skip = 1;
            else if (c1 == c2 && a1 >= a2) skip = 1;
            else {
                const int *l = nb15off + (size_t)id1 * ORC_MAX_N_NB15OFF;
                for (int k = 0; k < ORC_MAX_N_NB15OFF && l[k] != -1; k++)
                    if (l[k] == id2) {
                        skip = 1;
                        break;
                    }
            }
            if (skip) {
                int bit = i2 * NAC + i1;
                mask[bit / 32] |= (int32_t)(1u << (bit % 32));
            }
        }
    }
}

int orc_enumerate_cell_pairs(const orc_grid_t *g, const orc_cells_t *c, const int *nb15off, int gap_bug,
                             orc_cellpair_t *out, int cap, int *idx_head) {
    const int n_uni   = g->n_columns * g->n_uni_z;
    const int n_cells = c->n_cells;
    /* +n_uni_z slack: a cell top that rounds onto the upper box face indexes one
     * bin past its column (reproduced, see set_uniform_grid). */
    int *u2c = (int *)malloc(sizeof(int) * 2 * (n_uni + g->n_uni_z));
    int *c2u = (int *)malloc(sizeof(int) * 2 * (n_cells + 1));
    for (int i = 0; i < n_uni + g->n_uni_z; i++) u2c[2 * i] = u2c[2 * i + 1] = -1;

    /* set_uniform_grid (MiniCell.cpp:896-936) */
    for (int cl = 0; cl < n_cells; cl++) {
        int   x = c->cell_crd[cl * 3], y = c->cell_crd[cl * 3 + 1], z = c->cell_crd[cl * 3 + 2];
        float zmin = (float)c->crd[(cl * NAC) * 3 + 2] - g->lb[2];
        float zmax = (float)c->crd[(cl * NAC + NAC - 1) * 3 + 2] - g->lb[2];
        int   u0 = (int)(zmin / g->L_z_uni), u1 = (int)(zmax / g->L_z_uni);
        c2u[2 * cl]     = u0;
        c2u[2 * cl + 1] = u1;
        for (int u = u0; u <= u1; u++) {
            int id = uni_id(g, x, y, u);
            if (u2c[2 * id] < 0 || z < u2c[2 * id]) u2c[2 * id] = z;
            if (z > u2c[2 * id + 1]) u2c[2 * id + 1] = z;
        }
    }

    int np = 0;
    for (int c1 = 0; c1 < n_cells; c1++) {
        idx_head[c1] = np;
        int   x1 = c->cell_crd[c1 * 3], y1 = c->cell_crd[c1 * 3 + 1];
        float z1b = (float)c->crd[(c1 * NAC) * 3 + 2];
        float z1t = (float)c->crd[(c1 * NAC + NAC - 1) * 3 + 2];
        float dd[3];
        int   img[3], c2xy[2];
        for (int dx = -g->n_neighbors_xyz[0]; dx <= g->n_neighbors_xyz[0]; dx++) {
            int rel = x1 + dx;
            img[0]  = 0;
            c2xy[0] = rel;
            if (rel < 0) {
                img[0]  = -1;
                c2xy[0] = g->n_cells_xyz[0] + rel;
            } else if (rel >= g->n_cells_xyz[0]) {
                img[0]  = 1;
                c2xy[0] = rel - g->n_cells_xyz[0];
            }
            if (dx > 0) dd[0] = (float)(dx - 1) * g->L_cell_xyz[0];
            else if (dx < 0) dd[0] = (float)(dx + 1) * g->L_cell_xyz[0];
            else dd[0] = 0.0f;
            dd[0] = dd[0] * dd[0];
            for (int dy = -g->n_neighbors_xyz[1]; dy <= g->n_neighbors_xyz[1]; dy++) {
                int rely = y1 + dy;
                img[1]   = 0;
                c2xy[1]  = rely;
                if (rely < 0) {
                    img[1]  = -1;
                    c2xy[1] = g->n_cells_xyz[1] + rely;
                } else if (rely >= g->n_cells_xyz[1]) {
                    img[1]  = 1;
                    c2xy[1] = rely - g->n_cells_xyz[1];
                }
                if (dy > 0) dd[1] = (float)(dy - 1) * g->L_cell_xyz[1];
                else if (dy < 0) dd[1] = (float)(dy + 1) * g->L_cell_xyz[1];
                else dd[1] = 0.0f;
                dd[1] = dd[1] * dd[1];

                int fz[3], lz[3];
                int tf = c2u[2 * c1] - 2;
                if (tf < 0) {
                    fz[0] = g->n_uni_z + tf;
                    lz[0] = g->n_uni_z - 1;
                    tf    = 0;
                } else {
                    fz[0] = lz[0] = -1;
                }
                fz[1]  = tf;
                int tl = c2u[2 * c1 + 1] + 2;
                if (tl >= g->n_uni_z) {
                    fz[2] = 0;
                    lz[2] = tl - g->n_uni_z;
                    lz[1] = g->n_uni_z - 1;
                } else {
                    fz[2] = lz[2] = -1;
                    lz[1]         = tl;
                }
                int col2  = c2xy[0] + c2xy[1] * g->n_cells_xyz[0];
                int head2 = c->idx_xy_head_cell[col2];
                for (int ii = 0; ii < 3; ii++) {
                    img[2] = ii - 1;
                    if (fz[ii] < 0) continue;
                    int fu = uni_id(g, c2xy[0], c2xy[1], fz[ii]);
                    int lu = uni_id(g, c2xy[0], c2xy[1], lz[ii]);
                    int stg = 0;
                    if (gap_bug) {
                        for (int u = fu; u <= lu; u++) {
                            if (stg == 0 && u2c[2 * u] >= 0) {
                                stg = 1;
                                fu  = u;
                            } else if (stg == 1 && u2c[2 * u] < 0) {
                                lu = u - 1;
                                break;
                            }
                        }
                    } else {
                        int f = -1, l = -1;
                        for (int u = fu; u <= lu; u++)
                            if (u2c[2 * u] >= 0) {
                                if (f < 0) f = u;
                                l = u;
                            }
                        if (f >= 0) {
                            stg = 1;
                            fu  = f;
                            lu  = l;
                        }
                    }
                    if (stg == 0) continue;
                    int zc0 = u2c[2 * fu], zc1 = u2c[2 * lu + 1];
                    if (zc0 > zc1 || head2 < 0) continue;
                    int first = head2 + zc0, last = head2 + zc1;
                    for (int c2 = first; c2 <= last; c2++) {
                        float z2t = (float)c->crd[(c2 * NAC + NAC - 1) * 3 + 2] + (float)img[2] * g->L[2];
                        float z2b = (float)c->crd[(c2 * NAC) * 3 + 2] + (float)img[2] * g->L[2];
                        if (z2t < z1b) {
                            dd[2] = z1b - z2t;
                            dd[2] = dd[2] * dd[2];
                        } else if (z2b > z1t) {
                            dd[2] = z2b - z1t;
                            dd[2] = dd[2] * dd[2];
                        } else {
                            dd[2] = 0.0f;
                        }
                        if (dd[0] + dd[1] + dd[2] > g->cutoff_pair2) continue;
                        if (!valid_pair(c1, c2)) continue;
                        if (np >= cap) {
                            free(u2c);
                            free(c2u);
                            return -1;
                        }
                        int bits = 0;
                        if (img[0] == -1) bits |= 1;
                        else if (img[0] == 1) bits |= 2;
                        if (img[1] == -1) bits |= 4;
                        else if (img[1] == 1) bits |= 8;
                        if (img[2] == -1) bits |= 16;
                        else if (img[2] == 1) bits |= 32;
                        out[np].cell_id1 = c1;
                        out[np].cell_id2 = c2;
                        out[np].image    = bits;
                        cell_pair_mask(c, nb15off, c1, c2, out[np].pair_mask);
                        np++;
                    }
                }
            }
        }
    }
    idx_head[n_cells] = np;
    free(u2c);
    free(c2u);
    return np;
}
