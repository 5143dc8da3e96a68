// og_grid.hip -- device build of the 8-atom-cell grid and the cell-pair list.
//
// Bit-exact integer work: column binning, column order, in-column z order,
// dummy padding, uniform-z bins, cell-pair enumeration order and the 64-bit
// exclusion masks reproduce MiniCell (src/celeste/core/MiniCell.cpp) on the
// same FP64 coordinates.  This file is compiled with -ffp-contract=off and
// correctly rounded float division so every float expression rounds where
// the reference host build rounds it.
#include "og_internal.h"

// ---------------------------------------------------------------- clear --
// Zero per-column counters and the uniform-z bin table of each system.
__global__ void __launch_bounds__(OG_THREADS) k_grid_clear(const og_dev_sys *__restrict__ S,
                                                           const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         t  = bm.y * OG_THREADS + threadIdx.x;
    if (t < s.n_columns) {
        s.n_atoms_col[t] = 0;
        s.col_fill[t]    = 0;
    }
    const int n_bins = (s.n_columns + 1) * s.n_uni_z;
    if (t < n_bins) {
        s.u2c_min[t] = OG_EMPTY_BIN;
        s.u2c_max[t] = -1;
    }
    // overflow flags stay set until the host has read them (check_flags), so a
    // failed rebuild inside a multi-step graph is never hidden by a later one
}

// ------------------------------------------------------------------ bin --
// MiniCell::set_atoms_into_grid_xy binning (MiniCell.cpp:316-359):
//   float x = crd - lb (double subtraction rounded to float), periodic wrap
//   in float, column = floor(x / L_cell) with IEEE float division.
// Wave-aggregated counter add: lanes with the same key share one atomic
// (atoms in input order are mostly spatial neighbours, so a wave touches a
// few columns).  Returns the old counter value plus the lane's rank among
// the lanes of its key, i.e. what a per-lane atomicAdd(ctr + key, 1) could
// have returned.
__device__ __forceinline__ int og_wave_count_add(int *ctr, int key, bool active) {
    const int lane = threadIdx.x & 63;
    uint64_t  todo = __ballot(active);
    int       res  = 0;
    while (todo) {
        const int      leader = __builtin_ctzll(todo);
        const int      lkey   = __builtin_amdgcn_readlane(key, leader);
        const uint64_t m      = __ballot(active && key == lkey) & todo;
        int            base   = 0;
        if (lane == leader) base = atomicAdd(ctr + lkey, __builtin_popcountll(m));
        base = __builtin_amdgcn_readlane(base, leader);
        if ((m >> lane) & 1ull) res = base + __builtin_popcountll(m & ((1ull << lane) - 1ull));
        todo &= ~m;
    }
    return res;
}

__global__ void __launch_bounds__(OG_THREADS) k_grid_bin(const og_dev_sys *__restrict__ S,
                                                         const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         i  = bm.y * OG_THREADS + threadIdx.x;
    const bool        in = i < s.n_atoms;
    int  c[2] = {0, 0};
    bool bad = false;
    if (in) {
        for (int d = 0; d < 2; d++) {
            const double cd = s.crd[3 * i + d];
            if (!(cd == cd) || cd > 1e30 || cd < -1e30) {
                bad = true;
                c[d] = 0;
                continue;
            }
            if (cd > (double)s.ub[d] || cd < (double)s.lb[d]) atomicOr(&s.cnt->flags, OG_F_OUTSIDE_BOX);
            float x = (float)(cd - (double)s.lb[d]);
            if (x < -4.0f * s.L[d] || x >= 5.0f * s.L[d]) {  // far outside: flag instead of looping
                atomicOr(&s.cnt->flags, OG_F_OUTSIDE_BOX);
                x = 0.0f;
            }
            while (x < 0.0f) x += s.L[d];
            while (x >= s.L[d]) x -= s.L[d];
            int k = (int)floorf(x / s.L_cell[d]);
            const int n = d == 0 ? s.nx : s.ny;
            while (k < 0) k += n;
            while (k >= n) k -= n;
            c[d] = k;
        }
        const double z = s.crd[3 * i + 2];
        if (!(z == z) || z > 1e30 || z < -1e30) bad = true;
        if (bad) atomicOr(&s.cnt->flags, OG_F_NAN);
        const int col    = c[0] + c[1] * s.nx;
        s.col_of_atom[i] = col;
        s.zkey[i]        = og_float_key(bad ? 0.0f : (float)z);
    }
    og_wave_count_add(s.n_atoms_col, c[0] + c[1] * s.nx, in);
}

// ----------------------------------------------------------------- scan --
// MiniCell::set_idx_xy_head (MiniCell.cpp:785-861): cells per column,
// exclusive scan to head cell/atom, -1 heads for empty columns.
__global__ void __launch_bounds__(1024) k_grid_scan(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const og_dev_sys &s  = S[map[blockIdx.x].x];
    if (og_rebuild_skip(s)) return;
    const int         nc = s.n_columns;
    const int         T  = blockDim.x;
    const int         per = (nc + T - 1) / T;
    const int         k0  = threadIdx.x * per;
    const int         k1  = min(k0 + per, nc);
    __shared__ int    sh[1024];
    __shared__ int    big;
    if (threadIdx.x == 0) big = 0;
    int local = 0, mx = 0;
    for (int k = k0; k < k1; k++) {
        const int n = s.n_atoms_col[k];
        local += (n + OG_NAC - 1) / OG_NAC;
        mx = max(mx, n);
    }
    sh[threadIdx.x] = local;
    __syncthreads();
    if (mx > OG_SORT_CAP) atomicOr(&big, 1);
    // Hillis-Steele inclusive scan over 1024 partial sums
    for (int off = 1; off < T; off <<= 1) {
        const int v = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    int cell = sh[threadIdx.x] - local;
    for (int k = k0; k < k1; k++) {
        const int n  = s.n_atoms_col[k];
        const int nz = (n + OG_NAC - 1) / OG_NAC;
        s.n_cells_z[k] = nz;
        s.head_cell[k] = n ? cell : -1;
        s.head_atom[k] = n ? cell * OG_NAC : -1;
        s.col_fill[k]  = 0;
        cell += nz;
    }
    if (threadIdx.x == T - 1) {
        int total = sh[T - 1];
        s.head_cell[nc] = total;
        s.head_atom[nc] = total * OG_NAC;
        int fl = 0;
        if (total > s.cap_cells) fl |= OG_F_CELL_OVERFLOW;
        if (big) fl |= OG_F_COL_OVERFLOW;
        if (fl) {
            atomicOr(&s.cnt->flags, fl);
            total = 0;  // nothing downstream runs on a grid that does not fit
        }
        s.cnt->n_cells = total;
    }
}

// -------------------------------------------------------------- scatter --
__global__ void __launch_bounds__(OG_THREADS) k_grid_scatter(const og_dev_sys *__restrict__ S,
                                                             const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         i  = bm.y * OG_THREADS + threadIdx.x;
    if (s.cnt->n_cells == 0) return;
    const bool in   = i < s.n_atoms;
    const int  col  = in ? s.col_of_atom[i] : 0;
    const int  slot = og_wave_count_add(s.col_fill, col, in);
    if (in) s.tmp_ids[s.head_atom[col] + slot] = i;
}

__device__ void og_column_cells(const og_dev_sys &s, int col, int nz);

// ----------------------------------------------------------------- sort --
// One block per column.  Ascending float z with the atom id as tie-break
// (the reference Hoare quicksort, MiniCell.cpp:636-691, orders distinct z
// identically; exact float ties inside a column are the only difference),
// then padding to a multiple of 8 with copies of the last real atom
// (MiniCell.cpp:574-580), grid-order coordinates rounded through float
// (MiniCell.cpp:414-415), cell z extents and uniform-z bins
// (MiniCell.cpp:884-936).
__global__ void __launch_bounds__(OG_THREADS) k_grid_sort(const og_dev_sys *__restrict__ S,
                                                          const int2 *__restrict__ map) {
    const int2        bm  = map[blockIdx.x];
    const og_dev_sys &s   = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         col = bm.y;
    __shared__ uint64_t keys[OG_SORT_CAP];
    if (s.cnt->n_cells == 0) return;
    const int n = s.n_atoms_col[col];
    if (n == 0) return;
    const int base = s.head_atom[col];
    int       P    = 1;
    while (P < n) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) {
        if (t < n) {
            const int id = s.tmp_ids[base + t];
            keys[t]      = ((uint64_t)s.zkey[id] << 32) | (uint32_t)id;
        } else {
            keys[t] = ~0ull;
        }
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < P; t += blockDim.x) {
                const int u = t ^ j;
                if (u > t) {
                    const uint64_t a = keys[t], b = keys[u];
                    const bool     up = (t & k) == 0;
                    if ((a > b) == up) {
                        keys[t] = b;
                        keys[u] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int nz    = s.n_cells_z[col];
    const int last  = (int)(keys[n - 1] & 0xffffffffu);
    const int cap   = s.cap_atom_array;
    for (int t = threadIdx.x; t < nz * OG_NAC; t += blockDim.x) {
        const int slot = base + t;
        const int id   = t < n ? (int)(keys[t] & 0xffffffffu) : -1;
        const int src  = t < n ? id : last;
        const float x = (float)s.crd[3 * src], y = (float)s.crd[3 * src + 1], z = (float)s.crd[3 * src + 2];
        s.atomids[slot]      = id;
        s.gcrd[3 * slot]     = (double)x;
        s.gcrd[3 * slot + 1] = (double)y;
        s.gcrd[3 * slot + 2] = (double)z;
        s.gxyzq[slot]        = make_float4(x, y, z, id >= 0 ? s.charge[src] : 0.0f);
        s.gtype[slot]        = id >= 0 ? s.atom_type[src] : 0;
        s.gcomp[slot]        = id >= 0 ? s.comp[src] : -1;
        if (id >= 0) s.rev[id] = slot;
        (void)cap;
    }
    __syncthreads();
    og_column_cells(s, col, nz);
}

// Cells of one grid column from grid-order coordinates: cell -> column and z
// index, z extents from the first and last slot of each cell (dummies copy
// the last real atom, so the last slot holds the top), and the uniform-z bins
// the cell overlaps (MiniCell::set_uniform_grid, MiniCell.cpp:884-936).
// Called by one block per column after the column's slots are written.
__device__ void og_column_cells(const og_dev_sys &s, int col, int nz) {
    const int hc = s.head_cell[col];
    for (int c = threadIdx.x; c < nz; c += blockDim.x) {
        const int   cell = hc + c;
        const float zmin = (float)s.gcrd[3 * (cell * OG_NAC) + 2];
        const float zmax = (float)s.gcrd[3 * (cell * OG_NAC + OG_NAC - 1) + 2];
        s.cell_col[cell]  = col;
        s.cell_z[cell]    = c;
        s.cell_zmin[cell] = zmin;
        s.cell_zmax[cell] = zmax;
        const float a = zmin - s.lb[2];
        const float b = zmax - s.lb[2];
        // (int) truncation as MiniCell.cpp:912-913; the clamp only guards
        // coordinates outside the box (the reference indexes out of bounds there)
        const int u0 = min(max((int)(a / s.L_z_uni), 0), s.n_uni_z);
        const int u1 = min(max((int)(b / s.L_z_uni), u0), s.n_uni_z);
        s.cell_u0[cell] = u0;
        s.cell_u1[cell] = u1;
        for (int u = u0; u <= u1; u++) {
            // uni id = x*n_uni_z + y*n_uni_z*nx + z  (MiniCell.cpp:1596-1598); a top that
            // rounds onto the upper face spills into the next column's first bin, as there.
            const int bin = col * s.n_uni_z + u;
            atomicMin(&s.u2c_min[bin], c);
            atomicMax(&s.u2c_max[bin], c);
        }
    }
}

// Cell tables of a grid built on the host (the reference call points, where
// MiniCell sorts the atoms and uploads atomids, column heads and float
// coordinates): one block per column.
__global__ void __launch_bounds__(OG_THREADS) k_cells_from_grid(const og_dev_sys *__restrict__ S,
                                                                const int2 *__restrict__ map) {
    const int2        bm  = map[blockIdx.x];
    const og_dev_sys &s   = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         col = bm.y;
    if (s.head_cell[col] < 0) return;
    og_column_cells(s, col, s.n_cells_z[col]);
}

// ---------------------------------------------------------- enumeration --
// MiniCell::enumerate_cell_pairs (MiniCell.cpp:938-1149) for one cell1 per
// wave: lane k takes neighbour column k in the reference loop order
// (dx outer, dy inner), walks the three z images and the candidate cells in
// ascending id, so the per-lane prefix sum reproduces the CPU list order.
__device__ __forceinline__ bool og_valid_pair(int c1, int c2) {
    // MiniCell::check_valid_pair (MiniCell.cpp:1150-1160)
    const bool o1 = (c1 & 1) != 0, o2 = (c2 & 1) != 0;
    if (o1) return !((c2 < c1 && !o2) || (c2 > c1 && o2));
    return !((c2 < c1 && o2) || (c2 > c1 && !o2));
}

// Squared z gap between cell2 [z2b, z2t] (image-shifted) and cell1 [z1b, z1t]
// (MiniCell.cpp:1095-1104)
__device__ __forceinline__ float og_zgap2(float z2t, float z2b, float z1b, float z1t) {
    float d2;
    if (z2t < z1b) {
        d2 = z1b - z2t;
        d2 = d2 * d2;
    } else if (z2b > z1t) {
        d2 = z2b - z1t;
        d2 = d2 * d2;
    } else {
        d2 = 0.0f;
    }
    return d2;
}

// MODE 1: count the cell's pairs and stage them (cell2 | image << 24, list
// order) in cp_stage[c1 * OG_ENUM_STAGE ...] as far as they fit;
// MODE 2: write them at cp_head[c1] (cells whose pairs did not fit).
template <int MODE>
__device__ void og_enum_cell(const og_dev_sys &s, int c1, int *out_count) {
    constexpr bool WRITE = MODE != 0;
    const int lane  = threadIdx.x & 63;
    const int col1  = s.cell_col[c1];
    const int x1    = col1 % s.nx, y1 = col1 / s.nx;
    const float z1b = s.cell_zmin[c1], z1t = s.cell_zmax[c1];
    const int   u0  = s.cell_u0[c1], u1 = s.cell_u1[c1];
    const int   wy  = 2 * s.nny + 1;
    const int   ncand = (2 * s.nnx + 1) * wy;
    const int   nuz = s.n_uni_z;
    // z-bin windows per image (MiniCell.cpp:1014-1038)
    int fz[3], lz[3];
    int tf = u0 - 2;
    if (tf < 0) {
        fz[0] = nuz + tf;
        lz[0] = nuz - 1;
        tf    = 0;
    } else {
        fz[0] = lz[0] = -1;
    }
    fz[1]  = tf;
    int tl = u1 + 2;
    if (tl >= nuz) {
        fz[2] = 0;
        lz[2] = tl - nuz;
        lz[1] = nuz - 1;
    } else {
        fz[2] = lz[2] = -1;
        lz[1]         = tl;
    }
    int carry = MODE == 2 ? s.cp_head[c1] : 0;
    int total = 0;
    for (int r0 = 0; r0 < ncand; r0 += 64) {
        const int k     = r0 + lane;
        int       mine  = 0;
        int       c2lo[3], c2hi[3], img2 = 0;
        uint64_t  kept[3] = {0ull, 0ull, 0ull};
        float     dd01  = 0.0f;
        int       imgxy = 0;
        for (int ii = 0; ii < 3; ii++) c2lo[ii] = 0, c2hi[ii] = -1;
        if (k < ncand) {
            const int dx = k / wy - s.nnx, dy = k % wy - s.nny;
            int       rx = x1 + dx, ry = y1 + dy, cx = rx, cy = ry;
            if (rx < 0) {
                cx = s.nx + rx;
                imgxy |= 1;
            } else if (rx >= s.nx) {
                cx = rx - s.nx;
                imgxy |= 2;
            }
            if (ry < 0) {
                cy = s.ny + ry;
                imgxy |= 4;
            } else if (ry >= s.ny) {
                cy = ry - s.ny;
                imgxy |= 8;
            }
            float d0 = dx > 0 ? (float)(dx - 1) * s.L_cell[0] : dx < 0 ? (float)(dx + 1) * s.L_cell[0] : 0.0f;
            float d1 = dy > 0 ? (float)(dy - 1) * s.L_cell[1] : dy < 0 ? (float)(dy + 1) * s.L_cell[1] : 0.0f;
            d0        = d0 * d0;
            d1        = d1 * d1;
            dd01      = d0 + d1;
            const int col2  = cx + cy * s.nx;
            const int head2 = s.head_cell[col2];
            for (int ii = 0; ii < 3; ii++) {
                if (fz[ii] < 0 || head2 < 0) continue;
                int fu = col2 * nuz + fz[ii], lu = col2 * nuz + lz[ii];
                int f = -1, l = -1;
                if (s.gap_bug) {
                    // stop at the first empty bin after a non-empty one (MiniCell.cpp:1063-1077)
                    int stg = 0;
                    for (int u = fu; u <= lu; u++) {
                        const bool full = s.u2c_min[u] != OG_EMPTY_BIN;
                        if (stg == 0 && full) {
                            stg = 1;
                            f   = u;
                            l   = lu;
                        } else if (stg == 1 && !full) {
                            l = u - 1;
                            break;
                        }
                    }
                } else {
                    for (int u = fu; u <= lu; u++)
                        if (s.u2c_min[u] != OG_EMPTY_BIN) {
                            if (f < 0) f = u;
                            l = u;
                        }
                }
                if (f < 0) continue;
                // a cell top rounding onto the upper face spills into the next
                // column's first bin (see k_grid_sort); keep the range in this column
                const int zc0 = s.u2c_min[f], zc1 = min(s.u2c_max[l], s.n_cells_z[col2] - 1);
                if (zc0 > zc1) continue;
                c2lo[ii] = head2 + zc0;
                c2hi[ii] = head2 + zc1;
            }
            for (int ii = 0; ii < 3; ii++) {
                const float sh = (float)(ii - 1) * s.L[2];
                for (int c2 = c2lo[ii]; c2 <= c2hi[ii]; c2++) {
                    if (dd01 + og_zgap2(s.cell_zmax[c2] + sh, s.cell_zmin[c2] + sh, z1b, z1t) > s.cutoff_pair2) continue;
                    if (!og_valid_pair(c1, c2)) continue;
                    mine++;
                    // the kept cells of this image as bits from c2lo (the write below
                    // walks them instead of testing every cell again)
                    if (c2 - c2lo[ii] < 64) kept[ii] |= 1ull << (c2 - c2lo[ii]);
                }
            }
        }
        // exclusive prefix of per-lane counts in lane order
        int incl = mine;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const int round_total = __shfl(incl, 63, 64);
        if (WRITE && mine) {
            int p = carry + incl - mine;
            for (int ii = 0; ii < 3; ii++) {
                const float sh = (float)(ii - 1) * s.L[2];
                img2           = ii == 0 ? 16 : ii == 2 ? 32 : 0;
                auto emit = [&](int c2) {
                    const int rec = c2 | ((imgxy | img2) << 24);
                    if (MODE == 1) {
                        if (p < OG_ENUM_STAGE) s.cp_stage[(size_t)c1 * OG_ENUM_STAGE + p] = rec;
                    } else if (p < s.cap_cell_pairs) {
                        s.cp_c2[p] = rec;
                    }
                    p++;
                };
                if (c2hi[ii] - c2lo[ii] < 64) {
                    // ascending cell order, as the count
                    for (uint64_t m = kept[ii]; m; m &= m - 1) emit(c2lo[ii] + __builtin_ctzll(m));
                } else {
                    for (int c2 = c2lo[ii]; c2 <= c2hi[ii]; c2++) {
                        if (dd01 + og_zgap2(s.cell_zmax[c2] + sh, s.cell_zmin[c2] + sh, z1b, z1t) > s.cutoff_pair2)
                            continue;
                        if (!og_valid_pair(c1, c2)) continue;
                        emit(c2);
                    }
                }
            }
        }
        carry += round_total;
        total += round_total;
    }
    if (MODE == 1 && lane == 0) *out_count = total;
}

__global__ void __launch_bounds__(OG_THREADS) k_enum_count(const og_dev_sys *__restrict__ S,
                                                           const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         c1 = __builtin_amdgcn_readfirstlane(bm.y * (OG_THREADS / 64) + (threadIdx.x >> 6));
    if (c1 >= s.cnt->n_cells) return;
    og_enum_cell<1>(s, c1, &s.cp_count[c1]);
}

__global__ void __launch_bounds__(OG_THREADS) k_enum_write(const og_dev_sys *__restrict__ S,
                                                           const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int         c1 = __builtin_amdgcn_readfirstlane(bm.y * (OG_THREADS / 64) + (threadIdx.x >> 6));
    if (c1 >= s.cnt->n_cells) return;
    const int n = s.cp_count[c1];
    if (n > OG_ENUM_STAGE) {  // did not fit the stage: enumerate again, writing in place
        og_enum_cell<2>(s, c1, nullptr);
        return;
    }
    const int  lane = threadIdx.x & 63;
    const int  p0   = s.cp_head[c1];
    const int *src  = s.cp_stage + (size_t)c1 * OG_ENUM_STAGE;
    for (int k = lane; k < n; k += 64)
        if (p0 + k < s.cap_cell_pairs) s.cp_c2[p0 + k] = src[k];
}

// Exclusive scan of per-cell pair counts -> cp_head (one 1024-thread block per system).
// The same scan over many blocks of 1024 cells (a large system has 150 k
// cells): k_enum_scan_part writes each cell's exclusive prefix within its
// block and the block total; k_enum_scan_add adds the sum of the earlier
// blocks' totals, clamps to the list capacity and records the total.
__global__ void __launch_bounds__(1024) k_enum_scan_part(const og_dev_sys *__restrict__ S,
                                                         const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int n = s.cnt->n_cells;
    if (bm.y * 1024 > n) return;
    const int k    = bm.y * 1024 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ int wsum[16];
    const int      v    = k < n ? s.cp_count[k] : 0;
    int            incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < w; q++) before += wsum[q];
    if (k <= n) s.cp_head[k] = before + incl - v;
    if (threadIdx.x == 1023) s.cp_bsum[bm.y] = before + incl;
}

__global__ void __launch_bounds__(1024) k_enum_scan_add(const og_dev_sys *__restrict__ S,
                                                        const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int n = s.cnt->n_cells;
    if (bm.y * 1024 > n) return;
    const int k    = bm.y * 1024 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ long long wsum[16];
    long long o = 0;
    for (int q = threadIdx.x; q < bm.y; q += 1024) o += s.cp_bsum[q];
    for (int off = 32; off > 0; off >>= 1) o += __shfl_xor(o, off, 64);
    if (lane == 0) wsum[w] = o;
    __syncthreads();
    long long offset = 0;
    for (int q = 0; q < 16; q++) offset += wsum[q];
    if (k > n) return;
    const long long p = offset + s.cp_head[k];
    s.cp_head[k]      = (int)min(p, (long long)s.cap_cell_pairs);
    if (k == n) {
        if (p > s.cap_cell_pairs) {
            atomicOr(&s.cnt->flags, OG_F_CP_OVERFLOW);
            s.cnt->n_cell_pairs = 0;
        } else {
            s.cnt->n_cell_pairs = (int)p;
        }
    }
}

__global__ void __launch_bounds__(1024) k_enum_scan(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const og_dev_sys &s   = S[map[blockIdx.x].x];
    if (og_rebuild_skip(s)) return;
    const int         n   = s.cnt->n_cells;
    const int         T   = blockDim.x;
    const int         per = (n + T - 1) / T;
    const int         k0  = threadIdx.x * per, k1 = min(k0 + per, n);
    const int         lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ long long wsum[16];
    long long local = 0;
    for (int k = k0; k < k1; k++) local += s.cp_count[k];
    // inclusive scan within the wave (shuffles), then over the 16 wave totals
    long long incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        const long long v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    long long before = 0;
    for (int q = 0; q < w; q++) before += wsum[q];
    long long p = before + incl - local;
    for (int k = k0; k < k1; k++) {
        s.cp_head[k] = (int)min(p, (long long)s.cap_cell_pairs);
        p += s.cp_count[k];
    }
    if (threadIdx.x == T - 1) {
        const long long total = before + incl;
        s.cp_head[n]          = (int)min(total, (long long)s.cap_cell_pairs);
        if (total > s.cap_cell_pairs) {
            atomicOr(&s.cnt->flags, OG_F_CP_OVERFLOW);
            s.cnt->n_cell_pairs = 0;
        } else {
            s.cnt->n_cell_pairs = (int)total;
        }
    }
}

// ---------------------------------------------------------------- masks --
// MiniCell::set_cell_pair_bitmask (MiniCell.cpp:1198-1237): bit a2*8+a1 set
// (= skip) for a dummy atom, the lower triangle of a self pair, or an nb15off
// partner (the partner list read up to its first -1).  One wave per cell1.
// The exclusion bits are found from the atoms' side: the cell's partner
// lists are read once (lane = 8 * atom + partner) and each partner's grid
// slot becomes a (cell2, bit) entry in LDS; each lane then builds the masks
// of 64 of the cell's pairs from the dummy flags, the self-pair triangle and
// the entries whose cell2 matches.  Same bits as testing every atom pair.
__global__ void __launch_bounds__(OG_THREADS) k_masks(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    __shared__ int exq[OG_THREADS / 64][OG_NAC * OG_MAX_NB15];  // (cell2 << 6) | bit, per wave
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_rebuild_skip(s)) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c1   = __builtin_amdgcn_readfirstlane(bm.y * (OG_THREADS / 64) + w);
    if (c1 >= s.cnt->n_cells || s.cnt->n_cell_pairs == 0) return;
    int      *ex = exq[w];
    const int p0 = __builtin_amdgcn_readfirstlane(s.cp_head[c1]);
    const int p1 = __builtin_amdgcn_readfirstlane(s.cp_head[c1 + 1]);

    // exclusion entries: lane = 8 * a1 + (partner index within this round of 8)
    const int a1  = lane >> 3;
    const int id1 = s.atomids[c1 * OG_NAC + a1];
    bool      open = id1 >= 0;  // a1's list has not ended yet
    int       nex  = 0;
    for (int kb = 0; kb < OG_MAX_NB15; kb += 8) {
        const int      k    = kb + (lane & 7);
        const int      t    = open ? s.nb15off[(size_t)id1 * OG_MAX_NB15 + k] : -1;
        const uint64_t neg  = __ballot(open && t < 0);
        const unsigned term = (unsigned)(neg >> (8 * a1)) & 0xffu;  // -1 entries of a1 in this round
        const bool     keep = open && (lane & 7) < (term ? __builtin_ctz(term) : 8);
        open                = open && term == 0;
        const uint64_t kbal = __ballot(keep);
        if (keep) {
            const int g2   = s.rev[t];
            const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(kbal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)kbal, 0u));
            ex[nex + rank] = ((g2 >> 3) << 6) | ((g2 & 7) << 3) | a1;
        }
        nex += __builtin_popcountll(kbal);
        if (__ballot(open) == 0) break;
    }

    // dummy atoms of cell1: bit a1 of every row byte
    const int      idl  = s.atomids[c1 * OG_NAC + (lane & 7)];
    const unsigned dmy1 = (unsigned)__ballot(idl < 0) & 0xffu;
    const uint64_t cols = (uint64_t)dmy1 * 0x0101010101010101ull;
    // self pair: a1 >= a2, i.e. bits a2..7 of row byte a2
    uint64_t self = 0;
    for (int a2 = 0; a2 < OG_NAC; a2++) self |= (uint64_t)((0xffu << a2) & 0xffu) << (8 * a2);
    const int4 *A4 = (const int4 *)s.atomids;
    for (int base = p0; base < p1; base += 64) {
        const int p = base + lane;
        if (p < p1) {
            const int  c2 = s.cp_c2[p] & 0xffffff;
            const int4 u = A4[2 * c2], v = A4[2 * c2 + 1];
            uint64_t   m = cols | (c2 == c1 ? self : 0ull);
            // dummy atoms of cell2: the whole row byte
            m |= (u.x < 0 ? 0xffull : 0ull) | (u.y < 0 ? 0xffull << 8 : 0ull) | (u.z < 0 ? 0xffull << 16 : 0ull) |
                 (u.w < 0 ? 0xffull << 24 : 0ull) | (v.x < 0 ? 0xffull << 32 : 0ull) |
                 (v.y < 0 ? 0xffull << 40 : 0ull) | (v.z < 0 ? 0xffull << 48 : 0ull) |
                 (v.w < 0 ? 0xffull << 56 : 0ull);
            for (int e = 0; e < nex; e++) {
                const int q = ex[e];
                if ((q >> 6) == c2) m |= 1ull << (q & 63);
            }
            s.cp_mask[p] = m;
        }
    }
}

// ------------------------------------------------------------- launchers --
void og_launch_grid_clear(const og_launch &cols, const og_launch &) {
    hipLaunchKernelGGL(k_grid_clear, dim3(cols.n_blocks), dim3(OG_THREADS), 0, cols.stream, cols.sys, cols.map);
}
void og_launch_grid_bin(const og_launch &a) {
    hipLaunchKernelGGL(k_grid_bin, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_grid_scan(const og_launch &a) {
    hipLaunchKernelGGL(k_grid_scan, dim3(a.n_blocks), dim3(1024), 0, a.stream, a.sys, a.map);
}
void og_launch_grid_scatter(const og_launch &a) {
    hipLaunchKernelGGL(k_grid_scatter, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_grid_sort(const og_launch &a) {
    hipLaunchKernelGGL(k_grid_sort, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_cells_from_grid(const og_launch &a) {
    hipLaunchKernelGGL(k_cells_from_grid, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_enum_count(const og_launch &a) {
    hipLaunchKernelGGL(k_enum_count, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_enum_scan_blocks(const og_launch &a) {
    hipLaunchKernelGGL(k_enum_scan_part, dim3(a.n_blocks), dim3(1024), 0, a.stream, a.sys, a.map);
    hipLaunchKernelGGL(k_enum_scan_add, dim3(a.n_blocks), dim3(1024), 0, a.stream, a.sys, a.map);
}
void og_launch_enum_scan(const og_launch &a) {
    hipLaunchKernelGGL(k_enum_scan, dim3(a.n_blocks), dim3(1024), 0, a.stream, a.sys, a.map);
}
void og_launch_enum_write(const og_launch &a) {
    hipLaunchKernelGGL(k_enum_write, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_masks(const og_launch &a) {
    hipLaunchKernelGGL(k_masks, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
