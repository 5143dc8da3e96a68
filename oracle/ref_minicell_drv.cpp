// ref_minicell_drv.cpp -- test driver that calls the reference MiniCell class
// (src/celeste/core/MiniCell.cpp, compiled from the reference tree by
// oracle/Makefile) in the sequence SubBox::set_nsgrid uses
// (SubBox.cpp:487-545) and dumps its grid and cell-pair list.
//
// TEST INFRASTRUCTURE ONLY.  Built into oracle/_ref/, never shipped.
//
// Two driver choices, neither of which changes a result:
//  * MiniCell is constructed in zeroed storage, so the capacity member the
//    reference leaves uninitialised (max_n_atoms_exbox, MiniCell.cpp:1386)
//    starts at 0;
//  * the CPU build never allocates MiniCell::crd (MiniCell.cpp:93-103); the
//    driver allocates it through the public get_crd() reference;
//  * expected_num_density is passed as a large value so the column/cell
//    capacities (MiniCell.cpp:244-270) cannot overflow on the test systems.
//
// Input (binary, native endian):  int n_atoms; double box[12]; double nsgrid_cutoff;
//   int gap_unused; double crd[n*3]; int nb15off[n*32]
// Output: int n_cells; int n_columns; int n_cell_pairs; int atomids[n_cells*8];
//   double grid_crd[n_cells*8*3]; int idx_xy_head_cell[n_columns+1];
//   int idx_head_cell_pairs[n_cells+1]; CellPair cell_pairs[n_cell_pairs] (5 int32 each)
#include "MiniCell.h"
#include "PBC.h"
#include <cstdio>
#include <cstdlib>
#include <new>
#include <vector>

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE *fi = std::fopen(argv[1], "rb");
    if (!fi) return 3;
    int    n_atoms, unused;
    double box[12], nscut;
    if (std::fread(&n_atoms, sizeof(int), 1, fi) != 1) return 4;
    if (std::fread(box, sizeof(double), 12, fi) != 12) return 4;
    if (std::fread(&nscut, sizeof(double), 1, fi) != 1) return 4;
    if (std::fread(&unused, sizeof(int), 1, fi) != 1) return 4;
    std::vector<double> crd((size_t)n_atoms * 3);
    std::vector<int>    nb15off((size_t)n_atoms * 32);
    if (std::fread(crd.data(), sizeof(double), crd.size(), fi) != crd.size()) return 4;
    if (std::fread(nb15off.data(), sizeof(int), nb15off.size(), fi) != nb15off.size()) return 4;
    std::fclose(fi);

    PBC pbc;
    pbc.set_pbc(box);
    void     *mem = std::calloc(1, sizeof(MiniCell));
    MiniCell *mc  = new (mem) MiniCell();
    real      nscut_pw = (real)(real_pw)nscut;
    mc->set_grid_parameters(n_atoms, nscut_pw, &pbc, 32, nb15off.data(), 10.0);
    int  nbox[3]  = {1, 1, 1};
    real box_l[3] = {pbc.L[0], pbc.L[1], pbc.L[2]};
    mc->set_box_info(nbox, box_l);
    mc->set_max_n_atoms_region();
    mc->set_grid_xy();
    mc->alloc_variables();
    int   cap            = mc->get_max_n_atom_array();
    real *grid_crd       = new real[(size_t)cap * 3]();
    mc->get_crd()        = grid_crd;
    std::vector<int> ids(n_atoms);
    for (int i = 0; i < n_atoms; i++) ids[i] = i;
    mc->set_crds_to_homebox(crd.data(), ids.data(), n_atoms);
    mc->set_atoms_into_grid_xy();
    mc->enumerate_cell_pairs();

    int n_cells = mc->get_n_cells(), n_cols = mc->get_n_columns(), n_cp = mc->get_n_cell_pairs();
    FILE *fo = std::fopen(argv[2], "wb");
    if (!fo) return 5;
    std::fwrite(&n_cells, sizeof(int), 1, fo);
    std::fwrite(&n_cols, sizeof(int), 1, fo);
    std::fwrite(&n_cp, sizeof(int), 1, fo);
    std::fwrite(mc->get_atomids(), sizeof(int), (size_t)n_cells * 8, fo);
    std::fwrite(grid_crd, sizeof(double), (size_t)n_cells * 8 * 3, fo);
    std::fwrite(mc->get_idx_xy_head_cell(), sizeof(int), (size_t)n_cols + 1, fo);
    std::fwrite(mc->get_idx_head_cell_pairs(), sizeof(int), (size_t)n_cells + 1, fo);
    CellPair *cp = mc->get_cell_pairs();
    for (int i = 0; i < n_cp; i++) {
        int rec[5] = {cp[i].cell_id1, cp[i].cell_id2, cp[i].image, cp[i].pair_mask[0], cp[i].pair_mask[1]};
        std::fwrite(rec, sizeof(int), 5, fo);
    }
    std::fclose(fo);
    std::printf("n_cells %d n_columns %d n_cell_pairs %d\n", n_cells, n_cols, n_cp);
    std::fflush(stdout);
    std::_Exit(0); // skip the destructor: it frees the driver-owned crd with the wrong allocator pairing
}
