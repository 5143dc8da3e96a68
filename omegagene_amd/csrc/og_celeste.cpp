// og_celeste.cpp -- the reference's GPU call points (include/og_celeste.h) on
// the engine's kernels.
//
// celeste keeps one system per process in file-scope device globals
// (src/celeste/cuda/cuda_setup.h:77-133) fed by the host MiniCell.  This file
// keeps the same shape: one file-scope context whose og_dev_sys record is
// filled call by call (sizes from og_hostalloc_atom_info/og_alloc_atom_info,
// constants from og_set_pbc/og_set_constant/og_zerodipole_constant, the grid
// from og_set_cell_constant/og_memcpy_htod_atomids/og_memcpy_htod_crd) and
// then drives the batched kernels of og_grid.hip and og_pair.hip with a
// batch of one.  Compiled as HIP with -ffp-contract=off.
#include "og_internal.h"
#include "../../include/og_celeste.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

extern "C" og_status og_last_error(char *buf, int len);
void og_set_error(const char *msg);  // og_engine.cpp

namespace {

#define CS_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            char b_[384];                                                                       \
            std::snprintf(b_, sizeof b_, "%s failed: %s", #call, hipGetErrorString(e_));        \
            og_set_error(b_);                                                                   \
            return OG_E_DEVICE;                                                                 \
        }                                                                                       \
    } while (0)

#define CS_REQUIRE(cond, msg)       \
    do {                            \
        if (!(cond)) {              \
            og_set_error(msg);      \
            return OG_E_STATE;      \
        }                           \
    } while (0)

// grid-order work and pair energies in the reference layout
// (cuda_memcpy_dtoh_work, cuda_setup.cu:803-823): work[slot*3 + d]
__global__ void __launch_bounds__(OG_THREADS) k_cs_work(const og_dev_sys *__restrict__ S, double *__restrict__ work,
                                                        int n_atom_array) {
    const og_dev_sys &s = S[0];
    const int         g = blockIdx.x * OG_THREADS + threadIdx.x;
    if (g >= n_atom_array) return;
    const int cap = s.cap_atom_array;
    og_fx_read(s.fx, cap, og_fx_slot(g), work + 3 * (size_t)g);
}

// kernel_set_atominfo (cuda_setup.cu:364-379): types, charges and exclusion
// components into grid order, atom -> slot map
__global__ void __launch_bounds__(OG_THREADS) k_cs_set_atominfo(const og_dev_sys *__restrict__ S, int n_atom_array) {
    const og_dev_sys &s = S[0];
    const int         g = blockIdx.x * OG_THREADS + threadIdx.x;
    if (g >= s.cap_atom_array) return;
    const int id = g < n_atom_array ? s.atomids[g] : -1;
    s.gtype[g]   = id >= 0 ? s.atom_type[id] : 0;
    s.gcomp[g]   = id >= 0 ? s.comp[id] : -1;
    s.gxyzq[g].w = id >= 0 ? s.charge[id] : 0.0f;
    if (id >= 0) s.rev[id] = g;
}

// kernel_set_crd (cuda_setup.cu:380-402): float xyz into the pair stream and
// the grid coordinates the cell tables read
__global__ void __launch_bounds__(OG_THREADS) k_cs_set_crd(const og_dev_sys *__restrict__ S,
                                                           const float *__restrict__ crd, int n_atom_array) {
    const og_dev_sys &s = S[0];
    const int         g = blockIdx.x * OG_THREADS + threadIdx.x;
    if (g >= n_atom_array) return;
    const float x = crd[3 * g], y = crd[3 * g + 1], z = crd[3 * g + 2];
    s.gxyzq[g].x      = x;
    s.gxyzq[g].y      = y;
    s.gxyzq[g].z      = z;
    s.gcrd[3 * g]     = x;
    s.gcrd[3 * g + 1] = y;
    s.gcrd[3 * g + 2] = z;
}

struct CsCtx {
    og_dev_sys   h{};
    og_dev_sys  *d_sys = nullptr;
    og_dev_counters *d_cnt = nullptr;
    float       *d_crd_in = nullptr;     // host float crd staged on the device
    double      *d_work_grid = nullptr;  // [max_n_atom_array*3]
    hipStream_t  stream = nullptr;       // the reference's stream_pair_home
    std::vector<void *> allocs;
    // sizes and constants as the calls deliver them
    int   max_n_atom_array = 0, max_n_atoms_exbox = 0, max_n_cells = 0, max_n_cell_pairs = 0, n_columns_alloc = 0;
    int   n_atoms_system = 0, n_atom_array = 0, n_cells = 0, n_columns = 0, n_atoms_exbox = 0;
    int   n_types = 0;
    bool  have_pbc = false, have_const = false, have_zd = false, have_lj = false, have_alloc = false;
    bool  have_cellconst = false, have_ids = false, have_info = false, have_pairs = false;
    float cutoff = 0, cutoff_pair = 0;
    std::vector<int> comp_host;  // exclusion components (atom order)
    std::vector<float> lj6_host, lj12_host;  // the LJ tables as delivered (for the HPS table)
    bool  have_hps = false, have_dh = false;
    // int2 block maps
    int2 *map_cellw = nullptr, *map_col = nullptr, *map_slot = nullptr, *map_sys = nullptr;
    int   nb_cellw = 0, nb_col = 0, nb_slot = 0;
    std::vector<double *> pinned_d;
    std::vector<float *>  pinned_f;
    std::vector<int *>    pinned_i;
};
CsCtx g_cs;

template <class T> og_status cs_alloc(T **p, size_t n) {
    *p = nullptr;
    CS_HIP(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
    CS_HIP(hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)));
    g_cs.allocs.push_back((void *)*p);
    return OG_OK;
}

#define CS_TRY(x)                      \
    do {                               \
        og_status s_ = (x);            \
        if (s_ != OG_OK) return s_;    \
    } while (0)

og_status cs_push_sys() {
    CS_HIP(hipMemcpy(g_cs.d_sys, &g_cs.h, sizeof(og_dev_sys), hipMemcpyHostToDevice));
    return OG_OK;
}

og_status cs_maps() {
    og_dev_sys &h = g_cs.h;
    auto mk = [](int2 **d, int n, int *nb) -> og_status {
        if (*d) hipFree(*d), *d = nullptr;
        *nb = n;
        if (n == 0) return OG_OK;
        std::vector<int2> m(n);
        for (int b = 0; b < n; b++) m[b] = make_int2(0, b);
        CS_HIP(hipMalloc((void **)d, sizeof(int2) * n));
        CS_HIP(hipMemcpy(*d, m.data(), sizeof(int2) * n, hipMemcpyHostToDevice));
        return OG_OK;
    };
    const int n_bins = (h.n_columns + 1) * h.n_uni_z;
    CS_TRY(mk(&g_cs.map_cellw, (std::max(g_cs.n_cells, 1) + 3) / 4, &g_cs.nb_cellw));
    CS_TRY(mk(&g_cs.map_col, h.n_columns, &g_cs.nb_col));
    CS_TRY(mk(&g_cs.map_slot, (std::max(h.n_columns, n_bins) + OG_THREADS - 1) / OG_THREADS, &g_cs.nb_slot));
    int dummy = 0;
    CS_TRY(mk(&g_cs.map_sys, 1, &dummy));
    return OG_OK;
}

og_launch cs_launch(int2 *map, int nb) { return og_launch{g_cs.d_sys, map, nb, g_cs.stream}; }

// exclusion-connectivity components (union-find), as og_system_create does
std::vector<int> cs_components(int n, const int *tab) {
    std::vector<int> p(n);
    std::iota(p.begin(), p.end(), 0);
    auto find = [&](int x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    };
    for (int a = 0; a < n; a++)
        for (int k = 0; k < OG_MAX_NB15; k++) {
            const int b = tab[a * OG_MAX_NB15 + k];
            if (b < 0) break;
            const int ra = find(a), rb = find(b);
            if (ra != rb) p[std::max(ra, rb)] = std::min(ra, rb);
        }
    for (int a = 0; a < n; a++) p[a] = find(a);
    return p;
}

}  // namespace

extern "C" {

int og_hostalloc_atom_info(float **h_crd, int **h_atomids, double **h_work, double **h_energy,
                                 int max_n_atom_array) {
    if (!h_crd || !h_atomids || !h_work || !h_energy || max_n_atom_array <= 0) return OG_E_ARG;
    g_cs.max_n_atom_array = max_n_atom_array;
    CS_HIP(hipHostMalloc((void **)h_crd, sizeof(float) * 3 * (size_t)max_n_atom_array, hipHostMallocDefault));
    CS_HIP(hipHostMalloc((void **)h_atomids, sizeof(int) * (size_t)max_n_atom_array, hipHostMallocDefault));
    CS_HIP(hipHostMalloc((void **)h_work, sizeof(double) * 3 * (size_t)max_n_atom_array, hipHostMallocDefault));
    CS_HIP(hipHostMalloc((void **)h_energy, sizeof(double) * 2, hipHostMallocDefault));
    return OG_OK;
}

int og_hostfree_atom_info(float *h_crd, int *h_atomids, double **h_work, double **h_energy) {
    // work and energy come by reference (real_fc *&, MiniCell.cpp:19): freed and set to null
    if (!h_work || !h_energy) return OG_E_ARG;
    CS_HIP(hipHostFree(h_crd));
    CS_HIP(hipHostFree(h_atomids));
    CS_HIP(hipHostFree(*h_work));
    CS_HIP(hipHostFree(*h_energy));
    *h_work   = nullptr;
    *h_energy = nullptr;
    return OG_OK;
}

int og_hostalloc_cell_info(int **h_idx_xy_head_cell, int n_columns) {
    if (!h_idx_xy_head_cell || n_columns <= 0) return OG_E_ARG;
    // one extra entry: MiniCell writes idx_xy_head_cell[n_columns] (MiniCell.cpp:852)
    CS_HIP(hipHostMalloc((void **)h_idx_xy_head_cell, sizeof(int) * ((size_t)n_columns + 1), hipHostMallocDefault));
    return OG_OK;
}

int og_hostfree_cell_info(int *h_idx_xy_head_cell) {
    CS_HIP(hipHostFree(h_idx_xy_head_cell));
    return OG_OK;
}

int og_hostalloc_atom_type_charge(int **h_atom_type, float **h_charge, int n_atoms_system) {
    if (!h_atom_type || !h_charge || n_atoms_system <= 0) return OG_E_ARG;
    g_cs.n_atoms_system = n_atoms_system;
    CS_HIP(hipHostMalloc((void **)h_atom_type, sizeof(int) * (size_t)n_atoms_system, hipHostMallocDefault));
    CS_HIP(hipHostMalloc((void **)h_charge, sizeof(float) * (size_t)n_atoms_system, hipHostMallocDefault));
    return OG_OK;
}

int og_hostfree_atom_type_charge(int *h_atom_type, float *h_charge) {
    CS_HIP(hipHostFree(h_atom_type));
    CS_HIP(hipHostFree(h_charge));
    return OG_OK;
}

int og_alloc_atom_info(int max_n_atoms_exbox, int max_n_cells, int max_n_cell_pairs, int n_columns) {
    CS_REQUIRE(g_cs.max_n_atom_array > 0, "og_alloc_atom_info: call og_hostalloc_atom_info first");
    if (max_n_atoms_exbox <= 0 || max_n_cells <= 0 || max_n_cell_pairs <= 0 || n_columns <= 0) return OG_E_ARG;
    CS_REQUIRE(!g_cs.have_alloc, "og_alloc_atom_info: already allocated (og_free_atom_info first)");
    if (!g_cs.stream) CS_HIP(hipStreamCreateWithFlags(&g_cs.stream, hipStreamNonBlocking));
    og_dev_sys &h        = g_cs.h;
    g_cs.max_n_atoms_exbox = max_n_atoms_exbox;
    g_cs.max_n_cells       = max_n_cells;
    g_cs.max_n_cell_pairs  = max_n_cell_pairs;
    g_cs.n_columns_alloc   = n_columns;
    // whole groups of four cells: the fixed-point gradient layout (og_fx_slot) permutes within them
    const int cap          = (std::max(g_cs.max_n_atom_array, max_n_cells * OG_NAC) + 31) / 32 * 32;
    h.cap_atom_array       = cap;
    h.cap_cells            = cap / OG_NAC;
    h.cap_cell_pairs       = max_n_cell_pairs;
    h.n_atoms              = max_n_atoms_exbox;
    h.gap_bug              = 1;
    h.log_stride           = 1;  // energies on every evaluation, as cuda_pairwise_ljzd
    CS_TRY(cs_alloc(&g_cs.d_sys, 1));
    CS_TRY(cs_alloc(&g_cs.d_cnt, 1));
    h.cnt = g_cs.d_cnt;
    CS_TRY(cs_alloc(&g_cs.d_crd_in, 3 * (size_t)cap));
    CS_TRY(cs_alloc(&g_cs.d_work_grid, 3 * (size_t)cap));
    CS_TRY(cs_alloc(&h.atomids, cap));
    CS_TRY(cs_alloc(&h.rev, max_n_atoms_exbox));
    CS_TRY(cs_alloc(&h.gcrd, 3 * (size_t)cap));
    CS_TRY(cs_alloc(&h.gxyzq, cap));
    CS_TRY(cs_alloc(&h.gtype, cap));
    CS_TRY(cs_alloc(&h.gcomp, cap));
    CS_TRY(cs_alloc(&h.charge, max_n_atoms_exbox));
    CS_TRY(cs_alloc(&h.atom_type, max_n_atoms_exbox));
    CS_TRY(cs_alloc(&h.comp, max_n_atoms_exbox));
    CS_TRY(cs_alloc(&h.head_cell, (size_t)n_columns + 1));
    CS_TRY(cs_alloc(&h.n_cells_z, (size_t)n_columns + 1));
    CS_TRY(cs_alloc(&h.n_atoms_col, (size_t)n_columns + 1));
    CS_TRY(cs_alloc(&h.col_fill, (size_t)n_columns + 1));
    CS_TRY(cs_alloc(&h.cell_col, h.cap_cells));
    CS_TRY(cs_alloc(&h.cell_z, h.cap_cells));
    CS_TRY(cs_alloc(&h.cell_zmin, h.cap_cells));
    CS_TRY(cs_alloc(&h.cell_zmax, h.cap_cells));
    CS_TRY(cs_alloc(&h.cell_u0, h.cap_cells));
    CS_TRY(cs_alloc(&h.cell_u1, h.cap_cells));
    CS_TRY(cs_alloc(&h.cp_count, (size_t)h.cap_cells + 1));
    CS_TRY(cs_alloc(&h.cp_stage, (size_t)h.cap_cells * OG_ENUM_STAGE));
    CS_TRY(cs_alloc(&h.cp_head, (size_t)h.cap_cells + 1));
    CS_TRY(cs_alloc(&h.cp_c2, max_n_cell_pairs));
    CS_TRY(cs_alloc(&h.cp_mask, max_n_cell_pairs));
    CS_TRY(cs_alloc(&h.fx, OG_FX_WORDS(cap)));
    CS_TRY(cs_alloc(&h.e_acc, (size_t)OG_EACC_SLOTS * OG_EACC_STRIDE));
    CS_TRY(cs_alloc(&h.energy, OG_E_NTERMS));
    g_cs.have_alloc = true;
    return OG_OK;
}

int og_free_atom_info(void) {
    if (g_cs.stream) hipStreamSynchronize(g_cs.stream);
    for (void *p : g_cs.allocs) hipFree(p);
    g_cs.allocs.clear();
    for (int2 **m : {&g_cs.map_cellw, &g_cs.map_col, &g_cs.map_slot, &g_cs.map_sys})
        if (*m) hipFree(*m), *m = nullptr;
    const int keep_arr = g_cs.max_n_atom_array, keep_sys = g_cs.n_atoms_system;
    hipStream_t st = g_cs.stream;
    g_cs = CsCtx();
    g_cs.max_n_atom_array = keep_arr;  // the pinned host buffers outlive the device state
    g_cs.n_atoms_system   = keep_sys;
    g_cs.stream           = st;
    return OG_OK;
}

int og_alloc_set_lj_params(float *h_lj_6term, float *h_lj_12term, int n_lj_types, int *h_nb15off,
                           const int max_n_nb15off, const int max_n_atoms, const int max_n_atom_array) {
    (void)max_n_atoms;       // passed by SubBox.cpp:1693-1694; the sizes come from og_alloc_atom_info
    (void)max_n_atom_array;
    CS_REQUIRE(g_cs.have_alloc, "og_alloc_set_lj_params: call og_alloc_atom_info first");
    if (!h_lj_6term || !h_lj_12term || n_lj_types <= 0 || !h_nb15off || max_n_nb15off < 0) return OG_E_ARG;
    if (max_n_nb15off > OG_MAX_NB15) {
        og_set_error("og_alloc_set_lj_params: more than 32 exclusion partners per atom");
        return OG_E_ARG;
    }
    og_dev_sys &h = g_cs.h;
    const int   n = g_cs.max_n_atoms_exbox;
    if (h.lj6) return OG_E_STATE;
    h.n_types     = n_lj_types;
    g_cs.n_types  = n_lj_types;
    CS_TRY(cs_alloc(&h.lj6, (size_t)n_lj_types * n_lj_types));
    CS_TRY(cs_alloc(&h.lj12, (size_t)n_lj_types * n_lj_types));
    CS_HIP(hipMemcpy(h.lj6, h_lj_6term, sizeof(float) * n_lj_types * n_lj_types, hipMemcpyHostToDevice));
    CS_HIP(hipMemcpy(h.lj12, h_lj_12term, sizeof(float) * n_lj_types * n_lj_types, hipMemcpyHostToDevice));
    {
        const size_t       nt2 = (size_t)n_lj_types * n_lj_types;
        std::vector<float2> ljp(nt2);
        // the pair kernel's scaled entries {6 A6, 12 A12}
        for (size_t k = 0; k < nt2; k++) ljp[k] = make_float2(6.0f * h_lj_6term[k], 12.0f * h_lj_12term[k]);
        g_cs.lj6_host.assign(h_lj_6term, h_lj_6term + nt2);
        g_cs.lj12_host.assign(h_lj_12term, h_lj_12term + nt2);
        CS_TRY(cs_alloc(&h.lj_pair, nt2));
        CS_HIP(hipMemcpy(h.lj_pair, ljp.data(), sizeof(float2) * nt2, hipMemcpyHostToDevice));
    }
    // the reference stride is max_n_nb15off (MmSystem.cpp:478-525); the kernels read 32
    std::vector<int> tab((size_t)n * OG_MAX_NB15, -1);
    for (int a = 0; a < n; a++)
        for (int k = 0; k < max_n_nb15off; k++) {
            const int b = h_nb15off[(size_t)a * max_n_nb15off + k];
            if (b < -1 || b >= n) {
                og_set_error("og_alloc_set_lj_params: nb15off entry out of range");
                return OG_E_ARG;
            }
            tab[(size_t)a * OG_MAX_NB15 + k] = b;
        }
    g_cs.comp_host = cs_components(n, tab.data());
    CS_TRY(cs_alloc(&h.nb15off, tab.size()));
    CS_HIP(hipMemcpy(h.nb15off, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    CS_HIP(hipMemcpy(h.comp, g_cs.comp_host.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    g_cs.have_lj = true;
    return OG_OK;
}

int og_free_lj_params(void) {
    og_dev_sys &h = g_cs.h;
    for (void **p : {(void **)&h.lj6, (void **)&h.lj12, (void **)&h.lj_pair, (void **)&h.nb15off, (void **)&h.lj_hps}) {
        if (*p) {
            auto it = std::find(g_cs.allocs.begin(), g_cs.allocs.end(), *p);
            if (it != g_cs.allocs.end()) g_cs.allocs.erase(it);
            hipFree(*p);
            *p = nullptr;
        }
    }
    g_cs.have_lj  = false;
    g_cs.have_hps = false;
    return OG_OK;
}

// cuda_alloc_set_hps_params (cuda_setup.cu:1230-1244): the HPS switch distance and lambda per type pair,
// joined here with the LJ table into the {6 A6, 12 A12, switch, lambda} entries the pair kernel reads
int og_alloc_set_hps_params(float *h_hps_cutoff, float *h_hps_lambda, int n_lj_types) {
    CS_REQUIRE(g_cs.have_lj, "og_alloc_set_hps_params: call og_alloc_set_lj_params first");
    if (!h_hps_cutoff || !h_hps_lambda || n_lj_types != g_cs.n_types) return OG_E_ARG;
    og_dev_sys  &h   = g_cs.h;
    const size_t nt2 = (size_t)n_lj_types * n_lj_types;
    std::vector<float4> t4(nt2);
    for (size_t k = 0; k < nt2; k++)
        t4[k] = make_float4(6.0f * g_cs.lj6_host[k], 12.0f * g_cs.lj12_host[k], h_hps_cutoff[k], h_hps_lambda[k]);
    if (!h.lj_hps) CS_TRY(cs_alloc(&h.lj_hps, nt2));
    CS_HIP(hipMemcpy(h.lj_hps, t4.data(), sizeof(float4) * nt2, hipMemcpyHostToDevice));
    g_cs.have_hps = true;
    return OG_OK;
}

// cuda_free_hps_params (cuda_setup.cu:1245-1252)
int og_free_hps_params(void) {
    og_dev_sys &h = g_cs.h;
    if (h.lj_hps) {
        auto it = std::find(g_cs.allocs.begin(), g_cs.allocs.end(), (void *)h.lj_hps);
        if (it != g_cs.allocs.end()) g_cs.allocs.erase(it);
        hipFree(h.lj_hps);
        h.lj_hps = nullptr;
    }
    g_cs.have_hps = false;
    return OG_OK;
}

// cuda_hps_constant (cuda_setup.cu:1254-1258)
int og_hps_constant(float hps_eps) {
    g_cs.h.hps_eps = hps_eps;
    return OG_OK;
}

// cuda_debye_huckel_constant (cuda_setup.cu:1259-1270): the inverse Debye length in double from the
// float arguments, kept as float; constants CelesteObject.cpp:32-40
int og_debye_huckel_constant(float dielect, float temperature, float ionic_strength) {
    if (!(dielect > 0.0f) || !(temperature > 0.0f) || !(ionic_strength > 0.0f)) return OG_E_ARG;
    og_dev_sys &h = g_cs.h;
    h.dh_kappa    = (float)(1.0 / (sqrt(8.85418782e-12 * (double)dielect * 1.380658e-23 * (double)temperature /
                                        (2 * 6.0221413e23 * 1.60217657e-19 * 1.60217657e-19 * (double)ionic_strength)) *
                                   1e10));
    h.dh_diel_inv = (float)(1.0 / (double)dielect);
    g_cs.have_dh  = true;
    return OG_OK;
}

int og_set_pbc(float *l, float *lb) {
    if (!l || !lb) return OG_E_ARG;
    og_dev_sys &h = g_cs.h;
    for (int k = 0; k < 3; k++) {
        h.L[k]  = l[k];
        h.lb[k] = lb[k];
        h.ub[k] = lb[k] + l[k];
    }
    g_cs.have_pbc = true;
    return OG_OK;
}

int og_set_constant(float cutoff, float cutoff_pairlist, int n_atomtypes) {
    if (!(cutoff > 0.0f) || !(cutoff_pairlist > 0.0f) || n_atomtypes <= 0) return OG_E_ARG;
    og_dev_sys &h     = g_cs.h;
    g_cs.cutoff       = cutoff;
    g_cs.cutoff_pair  = cutoff_pairlist;
    h.cutoff          = cutoff;
    h.cutoff2         = cutoff * cutoff;
    h.cutoff_pair2    = cutoff_pairlist * cutoff_pairlist;
    g_cs.have_const   = true;
    return OG_OK;
}

int og_zerodipole_constant(float zcore, float bcoeff, float fcoeff) {
    og_dev_sys &h = g_cs.h;
    h.zcore       = zcore;
    h.bcoeff      = bcoeff;
    h.fcoeff      = fcoeff;
    h.alpha       = 0.0f;  // the reference GPU form is alpha = 0 (cuda_setup.cu:513-520)
    h.piewald     = 0.0f;
    g_cs.have_zd  = true;
    return OG_OK;
}

int og_memcpy_htod_atom_info(float **h_charge_orig, int **h_atomtype_orig) {
    CS_REQUIRE(g_cs.have_alloc && g_cs.have_lj,
               "og_memcpy_htod_atom_info: call og_alloc_atom_info and og_alloc_set_lj_params first");
    if (!h_charge_orig || !h_atomtype_orig || !*h_charge_orig || !*h_atomtype_orig) return OG_E_ARG;
    const int n = std::min(g_cs.n_atoms_system > 0 ? g_cs.n_atoms_system : g_cs.max_n_atoms_exbox,
                           g_cs.max_n_atoms_exbox);
    for (int i = 0; i < n; i++)
        if ((*h_atomtype_orig)[i] < 0 || (*h_atomtype_orig)[i] >= g_cs.n_types) {
            og_set_error("og_memcpy_htod_atom_info: atom type out of range");
            return OG_E_ARG;
        }
    CS_HIP(hipMemcpy(g_cs.h.charge, *h_charge_orig, sizeof(float) * n, hipMemcpyHostToDevice));
    CS_HIP(hipMemcpy(g_cs.h.atom_type, *h_atomtype_orig, sizeof(int) * n, hipMemcpyHostToDevice));
    g_cs.have_info = true;
    return OG_OK;
}

int og_set_cell_constant(int n_cells, int n_atoms_exbox, int n_atom_array, const int *n_cells_xyz,
                               int n_columns, const float *l_cell_xyz, const int *n_neighbor_xyz) {
    CS_REQUIRE(g_cs.have_alloc && g_cs.have_pbc && g_cs.have_const,
               "og_set_cell_constant: call og_alloc_atom_info, og_set_pbc and og_set_constant first");
    if (!n_cells_xyz || !l_cell_xyz || !n_neighbor_xyz || n_cells <= 0 || n_columns <= 0) return OG_E_ARG;
    og_dev_sys &h = g_cs.h;
    if (n_atom_array != n_cells * OG_NAC || n_atom_array > h.cap_atom_array || n_cells > h.cap_cells ||
        n_columns + 1 > g_cs.n_columns_alloc + 1 || n_columns != n_cells_xyz[0] * n_cells_xyz[1] ||
        n_atoms_exbox > g_cs.max_n_atoms_exbox) {
        og_set_error("og_set_cell_constant: grid sizes do not fit the allocation");
        return OG_E_CAPACITY;
    }
    g_cs.n_cells       = n_cells;
    g_cs.n_atom_array  = n_atom_array;
    g_cs.n_columns     = n_columns;
    g_cs.n_atoms_exbox = n_atoms_exbox;
    h.n_columns        = n_columns;
    h.nx               = n_cells_xyz[0];
    h.ny               = n_cells_xyz[1];
    h.nnx              = n_neighbor_xyz[0];
    h.nny              = n_neighbor_xyz[1];
    h.L_cell[0]        = l_cell_xyz[0];
    h.L_cell[1]        = l_cell_xyz[1];
    // MiniCell::set_grid_parameters (MiniCell.cpp:1299-1300), float arithmetic
    {
#pragma clang fp contract(off)
        const float half = (float)(g_cs.cutoff_pair * 0.5);
        h.n_uni_z        = (int)(h.L[2] / half);
        h.L_z_uni        = h.L[2] / (float)h.n_uni_z;
    }
    if (h.n_uni_z < 2 || h.nnx > h.nx || h.nny > h.ny) {
        og_set_error("og_set_cell_constant: box too small for the pair-list radius");
        return OG_E_ARG;
    }
    // uniform-z bin tables are sized by the geometry: (re)allocate on change
    const size_t n_bins = (size_t)(n_columns + 2) * h.n_uni_z;
    for (int **p : {&h.u2c_min, &h.u2c_max}) {
        if (*p) {
            auto it = std::find(g_cs.allocs.begin(), g_cs.allocs.end(), (void *)*p);
            if (it != g_cs.allocs.end()) g_cs.allocs.erase(it);
            hipFree(*p);
        }
        CS_TRY(cs_alloc(p, n_bins));
    }
    og_dev_counters c{};
    c.n_cells  = n_cells;
    c.step_end = OG_NO_STEP_END;
    CS_HIP(hipMemcpy(g_cs.d_cnt, &c, sizeof c, hipMemcpyHostToDevice));
    CS_TRY(cs_maps());
    CS_TRY(cs_push_sys());
    g_cs.have_cellconst = true;
    g_cs.have_pairs     = false;
    return OG_OK;
}

int og_memcpy_htod_atomids(int **h_atomids, int **h_idx_xy_head_cell) {
    CS_REQUIRE(g_cs.have_cellconst, "og_memcpy_htod_atomids: call og_set_cell_constant first");
    if (!h_atomids || !h_idx_xy_head_cell || !*h_atomids || !*h_idx_xy_head_cell) return OG_E_ARG;
    og_dev_sys &h  = g_cs.h;
    const int   nc = g_cs.n_columns, na = g_cs.n_atom_array;
    const int  *ids  = *h_atomids;
    const int  *head = *h_idx_xy_head_cell;
    // cells per column from the heads (empty column = -1, MiniCell.cpp:807-810)
    std::vector<int> nz(nc + 1, 0);
    int              next = g_cs.n_cells;
    for (int col = nc - 1; col >= 0; col--) {
        if (head[col] < 0) continue;
        if (head[col] > next) {
            og_set_error("og_memcpy_htod_atomids: column heads are not ascending");
            return OG_E_ARG;
        }
        nz[col] = next - head[col];
        next    = head[col];
    }
    for (int g = 0; g < na; g++)
        if (ids[g] < -1 || ids[g] >= g_cs.max_n_atoms_exbox) {
            og_set_error("og_memcpy_htod_atomids: atom id out of range");
            return OG_E_ARG;
        }
    std::vector<int> hd(head, head + nc);
    hd.push_back(g_cs.n_cells);
    CS_HIP(hipMemset(h.atomids, 0xff, sizeof(int) * h.cap_atom_array));
    CS_HIP(hipMemcpy(h.atomids, ids, sizeof(int) * na, hipMemcpyHostToDevice));
    CS_HIP(hipMemcpy(h.head_cell, hd.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
    CS_HIP(hipMemcpy(h.n_cells_z, nz.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
    g_cs.have_ids   = true;
    g_cs.have_pairs = false;
    return OG_OK;
}

int og_set_atominfo(void) {
    CS_REQUIRE(g_cs.have_ids && g_cs.have_info, "og_set_atominfo: call og_memcpy_htod_atom_info and og_memcpy_htod_atomids first");
    CS_TRY(cs_push_sys());
    const int cap = g_cs.h.cap_atom_array;
    hipLaunchKernelGGL(k_cs_set_atominfo, dim3((cap + OG_THREADS - 1) / OG_THREADS), dim3(OG_THREADS), 0, g_cs.stream,
                       g_cs.d_sys, g_cs.n_atom_array);
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(g_cs.stream));
    return OG_OK;
}

int og_memcpy_htod_crd(float **h_crd) {
    CS_REQUIRE(g_cs.have_cellconst, "og_memcpy_htod_crd: call og_set_cell_constant first");
    if (!h_crd || !*h_crd) return OG_E_ARG;
    CS_HIP(hipMemcpy(g_cs.d_crd_in, *h_crd, sizeof(float) * 3 * g_cs.n_atom_array, hipMemcpyHostToDevice));
    return OG_OK;
}

int og_set_crd(void) {
    CS_REQUIRE(g_cs.have_cellconst, "og_set_crd: call og_set_cell_constant first");
    CS_TRY(cs_push_sys());
    const int na = g_cs.n_atom_array;
    hipLaunchKernelGGL(k_cs_set_crd, dim3((na + OG_THREADS - 1) / OG_THREADS), dim3(OG_THREADS), 0, g_cs.stream,
                       g_cs.d_sys, g_cs.d_crd_in, na);
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(g_cs.stream));
    return OG_OK;
}

int og_enumerate_cell_pairs(int **h_atomids, int n_cells, int n_neighbor_col, const int *idx_atom_cell_xy) {
    (void)h_atomids;
    (void)idx_atom_cell_xy;  // the column of each cell comes from the uploaded column heads
    CS_REQUIRE(g_cs.have_ids && g_cs.have_lj, "og_enumerate_cell_pairs: grid or exclusions not uploaded");
    og_dev_sys &h = g_cs.h;
    if (n_cells != g_cs.n_cells || n_neighbor_col != (2 * h.nnx + 1) * (2 * h.nny + 1)) {
        og_set_error("og_enumerate_cell_pairs: n_cells / n_neighbor_col differ from og_set_cell_constant");
        return OG_E_ARG;
    }
    CS_TRY(cs_push_sys());
    og_dev_counters c{};
    c.n_cells  = n_cells;
    c.step_end = OG_NO_STEP_END;
    CS_HIP(hipMemcpy(g_cs.d_cnt, &c, sizeof c, hipMemcpyHostToDevice));
    og_launch_grid_clear(cs_launch(g_cs.map_slot, g_cs.nb_slot), cs_launch(g_cs.map_slot, g_cs.nb_slot));
    og_launch_cells_from_grid(cs_launch(g_cs.map_col, g_cs.nb_col));
    og_launch_enum_count(cs_launch(g_cs.map_cellw, g_cs.nb_cellw));
    og_launch_enum_scan(cs_launch(g_cs.map_sys, 1));
    og_launch_enum_write(cs_launch(g_cs.map_cellw, g_cs.nb_cellw));
    og_launch_masks(cs_launch(g_cs.map_cellw, g_cs.nb_cellw));
    CS_HIP(hipGetLastError());
    CS_HIP(hipStreamSynchronize(g_cs.stream));
    CS_HIP(hipMemcpy(&c, g_cs.d_cnt, sizeof c, hipMemcpyDeviceToHost));
    if (c.flags & OG_F_CP_OVERFLOW) {
        og_set_error("og_enumerate_cell_pairs: more cell pairs than max_n_cell_pairs");
        return OG_E_CAPACITY;
    }
    g_cs.have_pairs = true;
    return OG_OK;
}

int og_pairwise_ljzd(bool flg_mod_15mask) {
    (void)flg_mod_15mask;
    CS_REQUIRE(g_cs.have_pairs && g_cs.have_zd, "og_pairwise_ljzd: no cell-pair list (og_enumerate_cell_pairs)");
    og_dev_sys &h = g_cs.h;
    CS_TRY(cs_push_sys());
    CS_HIP(hipMemsetAsync(h.fx, 0, sizeof(long long) * OG_FX_WORDS(h.cap_atom_array), g_cs.stream));
    og_launch_pair(cs_launch(g_cs.map_cellw, g_cs.nb_cellw), OG_FORM_ZD, 4);
    og_launch_energy_reduce(cs_launch(g_cs.map_sys, 1));
    CS_HIP(hipGetLastError());
    return OG_OK;
}

// cuda_pairwise_hps_dh (cuda_setup.cu:1478-1493): hydrophobicity-scale LJ + Debye-Hueckel on the same
// cell-pair list, results through og_memcpy_dtoh_work as og_pairwise_ljzd
int og_pairwise_hps_dh(bool flg_mod_15mask) {
    (void)flg_mod_15mask;
    CS_REQUIRE(g_cs.have_pairs, "og_pairwise_hps_dh: no cell-pair list (og_enumerate_cell_pairs)");
    CS_REQUIRE(g_cs.have_hps && g_cs.have_dh,
               "og_pairwise_hps_dh: call og_alloc_set_hps_params and og_debye_huckel_constant first");
    og_dev_sys &h = g_cs.h;
    CS_TRY(cs_push_sys());
    CS_HIP(hipMemsetAsync(h.fx, 0, sizeof(long long) * OG_FX_WORDS(h.cap_atom_array), g_cs.stream));
    og_launch_pair(cs_launch(g_cs.map_cellw, g_cs.nb_cellw), OG_FORM_HPS_DH, 4);
    og_launch_energy_reduce(cs_launch(g_cs.map_sys, 1));
    CS_HIP(hipGetLastError());
    return OG_OK;
}

int og_pair_sync(void) {
    if (g_cs.stream) CS_HIP(hipStreamSynchronize(g_cs.stream));
    return OG_OK;
}

int og_memcpy_dtoh_work(double **h_work, double **h_energy, int n_atoms, int n_atom_array) {
    (void)n_atoms;
    CS_REQUIRE(g_cs.have_pairs, "og_memcpy_dtoh_work: no pair evaluation");
    if (!h_work || !h_energy || !*h_work || !*h_energy || n_atom_array != g_cs.n_atom_array) return OG_E_ARG;
    hipLaunchKernelGGL(k_cs_work, dim3((n_atom_array + OG_THREADS - 1) / OG_THREADS), dim3(OG_THREADS), 0,
                       g_cs.stream, g_cs.d_sys, g_cs.d_work_grid, n_atom_array);
    CS_HIP(hipGetLastError());
    CS_HIP(hipMemcpyAsync(*h_work, g_cs.d_work_grid, sizeof(double) * 3 * n_atom_array, hipMemcpyDeviceToHost,
                          g_cs.stream));
    CS_HIP(hipMemcpyAsync(*h_energy, g_cs.h.energy, sizeof(double) * 2, hipMemcpyDeviceToHost, g_cs.stream));
    CS_HIP(hipStreamSynchronize(g_cs.stream));
    og_dev_counters c;
    CS_HIP(hipMemcpy(&c, g_cs.d_cnt, sizeof c, hipMemcpyDeviceToHost));
    if (c.flags & OG_F_NAN) {
        // reported once: the flag would otherwise stop every later evaluation
        const int zero = 0;
        CS_HIP(hipMemcpy(&g_cs.d_cnt->flags, &zero, sizeof(int), hipMemcpyHostToDevice));
        og_set_error("og_memcpy_dtoh_work: non-finite or out-of-range gradient");
        return OG_E_NAN;
    }
    return OG_OK;
}

int og_get_celeste_cell_pairs(int32_t *cp5, int *idx_head, int *n_cell_pairs) {
    CS_REQUIRE(g_cs.have_pairs, "og_get_celeste_cell_pairs: no cell-pair list");
    og_dev_counters c;
    CS_HIP(hipMemcpy(&c, g_cs.d_cnt, sizeof c, hipMemcpyDeviceToHost));
    if (n_cell_pairs) *n_cell_pairs = c.n_cell_pairs;
    std::vector<int> head(c.n_cells + 1);
    CS_HIP(hipMemcpy(head.data(), g_cs.h.cp_head, sizeof(int) * (c.n_cells + 1), hipMemcpyDeviceToHost));
    if (idx_head) std::memcpy(idx_head, head.data(), sizeof(int) * (c.n_cells + 1));
    if (cp5 && c.n_cell_pairs) {
        std::vector<uint32_t> c2(c.n_cell_pairs);
        std::vector<uint64_t> mk(c.n_cell_pairs);
        CS_HIP(hipMemcpy(c2.data(), g_cs.h.cp_c2, sizeof(int) * c.n_cell_pairs, hipMemcpyDeviceToHost));
        CS_HIP(hipMemcpy(mk.data(), g_cs.h.cp_mask, sizeof(uint64_t) * c.n_cell_pairs, hipMemcpyDeviceToHost));
        for (int c1 = 0; c1 < c.n_cells; c1++)
            for (int p = head[c1]; p < head[c1 + 1]; p++) {
                cp5[5 * p]     = c1;
                cp5[5 * p + 1] = (int32_t)(c2[p] & 0xffffffu);
                cp5[5 * p + 2] = (int32_t)(c2[p] >> 24);
                cp5[5 * p + 3] = (int32_t)(uint32_t)(mk[p] & 0xffffffffu);
                cp5[5 * p + 4] = (int32_t)(uint32_t)(mk[p] >> 32);
            }
    }
    return OG_OK;
}

}  // extern "C"
