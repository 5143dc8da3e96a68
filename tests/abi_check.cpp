// Compile-time check of the drop-in boundary: every call point of the
// reference's GPU path, as its callers declare it (src/celeste/core/SubBox.cpp:5-78,
// MiniCell.cpp:7-35), written with the og_ prefix and T*& as T**, must be the
// exact type of the declaration in include/og_celeste.h (or og_engine.h for
// cuda_set_device / cuda_thread_sync).  Built with g++ -fsyntax-only by
// tests/test_cls_abi.py; a mismatch is a compile error naming the function.
#include <type_traits>

#include "og_celeste.h"

typedef float  real_pw;  // CelesteObject.h:11-24 (single-precision pair build)
typedef double real_fc;

#define SAME(f, ...) static_assert(std::is_same<decltype(&f), __VA_ARGS__>::value, #f " differs from the reference prototype")

// SubBox.cpp:9-13 cuda_alloc_atom_info
SAME(og_alloc_atom_info, int (*)(int, int, int, int));
// SubBox.cpp:15 cuda_free_atom_info
SAME(og_free_atom_info, int (*)());
// SubBox.cpp:16 cuda_memcpy_htod_atom_info(real_pw *&, int *&)
SAME(og_memcpy_htod_atom_info, int (*)(real_pw **, int **));
// SubBox.cpp:18-24 cuda_set_cell_constant
SAME(og_set_cell_constant, int (*)(const int, const int, const int, const int *, const int, const real_pw *, const int *));
// SubBox.cpp:26 cuda_set_constant
SAME(og_set_constant, int (*)(real_pw, real_pw, int));
// SubBox.cpp:27-33 cuda_alloc_set_lj_params (seven arguments as its caller declares it)
SAME(og_alloc_set_lj_params, int (*)(real_pw *, real_pw *, int, int *, const int, const int, const int));
// SubBox.cpp:34 cuda_free_lj_params
SAME(og_free_lj_params, int (*)());
// SubBox.cpp:36 cuda_memcpy_htod_atomids(int *&, int *&)
SAME(og_memcpy_htod_atomids, int (*)(int **, int **));
// SubBox.cpp:37 cuda_set_pbc
SAME(og_set_pbc, int (*)(real_pw *, real_pw *));
// SubBox.cpp:38 cuda_memcpy_htod_crd(real_pw *&)
SAME(og_memcpy_htod_crd, int (*)(real_pw **));
// SubBox.cpp:40-42 cuda_set_atominfo, cuda_set_crd, cuda_pairwise_ljzd(const bool)
SAME(og_set_atominfo, int (*)());
SAME(og_set_crd, int (*)());
SAME(og_pairwise_ljzd, int (*)(const bool));
// SubBox.cpp:44 cuda_memcpy_dtoh_work(real_fc *&, real_fc *&, int, int)
SAME(og_memcpy_dtoh_work, int (*)(real_fc **, real_fc **, int, int));
// SubBox.cpp:46 cuda_pair_sync
SAME(og_pair_sync, int (*)());
// SubBox.cpp:50 cuda_zerodipole_constant
SAME(og_zerodipole_constant, int (*)(real_pw, real_pw, real_pw));
// SubBox.cpp:52-54 cuda_hostfree_atom_type_charge / cuda_hostalloc_atom_type_charge(int *&, real_pw *&, const int)
SAME(og_hostfree_atom_type_charge, int (*)(int *, real_pw *));
SAME(og_hostalloc_atom_type_charge, int (*)(int **, real_pw **, const int));
// SubBox.cpp:57-60 cuda_enumerate_cell_pairs(int *&, const int, const int, const int *)
SAME(og_enumerate_cell_pairs, int (*)(int **, const int, const int, const int *));
// SubBox.cpp:62-71 the HPS + Debye-Hueckel call points
SAME(og_hps_constant, int (*)(real_pw));
SAME(og_alloc_set_hps_params, int (*)(real_pw *, real_pw *, int));
SAME(og_free_hps_params, int (*)());
SAME(og_debye_huckel_constant, int (*)(real_pw, real_pw, real_pw));
SAME(og_pairwise_hps_dh, int (*)(const bool));
// MiniCell.cpp:8-12 cuda_hostalloc_atom_info(real_pw *&, int *&, real_fc *&, real_fc *&, int)
SAME(og_hostalloc_atom_info, int (*)(real_pw **, int **, real_fc **, real_fc **, int));
// MiniCell.cpp:13-18 cuda_hostalloc_cell_info(int *&, int)
SAME(og_hostalloc_cell_info, int (*)(int **, int));
// MiniCell.cpp:19 cuda_hostfree_atom_info(real_pw *, int *, real_fc *&, real_fc *&)
SAME(og_hostfree_atom_info, int (*)(real_pw *, int *, real_fc **, real_fc **));
// MiniCell.cpp:20-22 cuda_hostfree_cell_info(int *)
SAME(og_hostfree_cell_info, int (*)(int *));

// SubBox.cpp:7 cuda_set_device(int) and SubBox.cpp:47 cuda_thread_sync() are
// the engine's og_set_device / og_thread_sync: same parameters, returning
// og_status, an enum of int size returned like an int.
SAME(og_set_device, og_status (*)(int));
SAME(og_thread_sync, og_status (*)());
static_assert(sizeof(og_status) == sizeof(int), "og_status is returned like an int");

// Not provided (outside the default GPU build this engine replaces):
// cudaMemoryTest (SubBox.cpp:6, a debug probe), cuda_init_cellinfo
// (SubBox.cpp:56, declared but never defined by the reference), and the
// F_ECP host-list variant cuda_hostalloc_cellpair_info /
// cuda_hostfree_cellpair_info / cuda_memcpy_htod_cell_pairs
// (MiniCell.cpp:25-32, SubBox.cpp:74-76).
int main() { return 0; }
