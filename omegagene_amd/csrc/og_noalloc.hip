// og_noalloc.hip -- the call-point pair evaluation (og_pairwise_ljzd followed
// by og_memcpy_dtoh_work) for callers that cannot allocate device memory.
//
// The caller owns every input and output buffer on the device.  What the
// kernels need beyond them -- the system record the pair kernel reads, its
// counters, the block map, the fixed-point gradient accumulators and the
// energy accumulators -- lives in one file-scope device array, `scratch`,
// sized for OG_NOALLOC_MAX_ATOM_ARRAY grid slots.  Each GPU holds its own
// copy for the life of the process; its device address is taken once per
// device with hipGetSymbolAddress and only that address is used from the
// host.  The call is four kernel launches on the caller's stream: set-up
// (record, counters, block map, zeroed accumulators), the pair kernel of
// og_pair.hip, the energy reduction, and the conversion to FP64 work.  It
// calls no allocator, no memset and no copy, so it can be captured into a
// HIP graph.  Calls that overlap on one GPU share the array and race.
#include "og_internal.h"
#include "../../include/og_celeste.h"

#include <cstdio>
#include <mutex>

void og_set_error(const char *msg);  // og_engine.cpp

namespace {

constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }

constexpr int    kMaxSlots  = OG_NOALLOC_MAX_ATOM_ARRAY;
constexpr int    kMaxCells  = kMaxSlots / OG_NAC;
constexpr size_t kOffSys    = 0;
constexpr size_t kOffCnt    = kOffSys + na_round(sizeof(og_dev_sys));
constexpr size_t kOffEacc   = kOffCnt + na_round(sizeof(og_dev_counters));
constexpr size_t kOffEnergy = kOffEacc + na_round(sizeof(long long) * OG_EACC_SLOTS * OG_EACC_STRIDE);
constexpr size_t kOffMap    = kOffEnergy + na_round(sizeof(double) * OG_E_NTERMS);
constexpr size_t kOffFx     = kOffMap + na_round(sizeof(int2) * kMaxCells);
constexpr size_t kBytes     = kOffFx + sizeof(long long) * OG_FX_WORDS(kMaxSlots);
constexpr int    kWaves     = 4;  // waves (cells) per pair-kernel block
static_assert(kMaxSlots % 32 == 0, "the fixed-point layout permutes slots within groups of 32");

}  // namespace

// about 11.3 MiB per GPU at 262,144 slots
static __device__ __align__(256) unsigned char scratch[kBytes];

namespace {

// Record, counters and block map into the scratch array; accumulators zeroed.
__global__ void __launch_bounds__(OG_THREADS) k_na_setup(unsigned char *base, og_dev_sys sys, int n_cells) {
    const int g      = blockIdx.x * OG_THREADS + threadIdx.x;
    const int stride = gridDim.x * OG_THREADS;
    long long *fx    = (long long *)(base + kOffFx);
    for (size_t k = g; k < OG_FX_WORDS(sys.cap_atom_array); k += stride) fx[k] = 0;
    long long *eacc = (long long *)(base + kOffEacc);
    for (int k = g; k < OG_EACC_SLOTS * OG_EACC_STRIDE; k += stride) eacc[k] = 0;
    int2 *map = (int2 *)(base + kOffMap);
    for (int k = g; k < (n_cells + kWaves - 1) / kWaves; k += stride) map[k] = make_int2(0, k);
    if (g == 0) {
        *(og_dev_sys *)(base + kOffSys) = sys;
        og_dev_counters c{};
        c.n_cells    = n_cells;
        c.step_end   = OG_NO_STEP_END;
        c.noise_step = -1;
        *(og_dev_counters *)(base + kOffCnt) = c;
        double *e = (double *)(base + kOffEnergy);
        for (int t = 0; t < OG_E_NTERMS; t++) e[t] = 0.0;
    }
}

// Grid-order FP64 work, the energies {vdw, ele} and the status flags out to
// the caller (the layout of og_memcpy_dtoh_work).
__global__ void __launch_bounds__(OG_THREADS) k_na_out(const og_dev_sys *__restrict__ S, double *__restrict__ work,
                                                       double *__restrict__ energy, int *__restrict__ flags,
                                                       int n_atom_array) {
    const og_dev_sys &s = S[0];
    const int         g = blockIdx.x * OG_THREADS + threadIdx.x;
    if (g < n_atom_array) og_fx_read(s.fx, s.cap_atom_array, og_fx_slot(g), work + 3 * (size_t)g);
    if (g == 0) {
        energy[0] = s.energy[OG_E_VDW];
        energy[1] = s.energy[OG_E_ELE];
        if (flags) *flags = s.cnt->flags;
    }
}

std::mutex     g_na_mu;
unsigned char *g_na_ptr[OG_NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

int na_fail(int code, const char *msg) {
    og_set_error(msg);
    return code;
}

}  // namespace

extern "C" int og_pairwise_ljzd_noalloc(const float *d_xyzq, const int *d_atom_type, const float *d_lj_pair,
                                        int n_lj_types, const int *d_idx_head_cell_pairs,
                                        const uint32_t *d_cell_pair_c2, const uint64_t *d_cell_pair_mask,
                                        int n_cells, const float *l, float cutoff, float zcore, float bcoeff,
                                        float fcoeff, double *d_work, double *d_energy, int *d_flags,
                                        void *stream) {
    if (!d_xyzq || !d_atom_type || !d_lj_pair || !d_idx_head_cell_pairs || !d_cell_pair_c2 || !d_cell_pair_mask ||
        !l || !d_work || !d_energy)
        return na_fail(OG_E_ARG, "og_pairwise_ljzd_noalloc: null buffer");
    if (n_cells <= 0 || n_lj_types <= 0 || !(cutoff > 0.0f) || !(l[0] > 0.0f) || !(l[1] > 0.0f) || !(l[2] > 0.0f))
        return na_fail(OG_E_ARG, "og_pairwise_ljzd_noalloc: n_cells, n_lj_types, cutoff and box must be positive");
    if (n_cells > kMaxCells) {
        char b[160];
        std::snprintf(b, sizeof b,
                      "og_pairwise_ljzd_noalloc: %d cells (%d grid slots) exceed the fixed scratch maximum of %d slots",
                      n_cells, n_cells * OG_NAC, kMaxSlots);
        return na_fail(OG_E_CAPACITY, b);
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= OG_NOALLOC_MAX_DEVICES)
        return na_fail(OG_E_DEVICE, "og_pairwise_ljzd_noalloc: no current device");
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_na_mu);
        if (!g_na_ptr[dev]) {
            void            *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                char b[200];
                std::snprintf(b, sizeof b, "og_pairwise_ljzd_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                              hipGetErrorString(e));
                return na_fail(OG_E_DEVICE, b);
            }
            g_na_ptr[dev] = (unsigned char *)p;
        }
        base = g_na_ptr[dev];
    }
    const int n_atom_array = n_cells * OG_NAC;
    og_dev_sys h{};
    h.n_atoms        = n_atom_array;
    h.cap_atom_array = (n_atom_array + 31) / 32 * 32;
    h.cap_cells      = h.cap_atom_array / OG_NAC;
    h.n_types        = n_lj_types;
    h.gap_bug        = 1;
    h.log_stride     = 1;  // energies on every evaluation, as og_pairwise_ljzd
    for (int k = 0; k < 3; k++) h.L[k] = l[k];
    h.cutoff   = cutoff;
    h.cutoff2  = cutoff * cutoff;
    h.zcore    = zcore;
    h.bcoeff   = bcoeff;
    h.fcoeff   = fcoeff;
    h.alpha    = 0.0f;  // the reference GPU form is alpha = 0 (cuda_setup.cu:513-520)
    h.piewald  = 0.0f;
    h.gxyzq    = (float4 *)d_xyzq;
    h.gtype    = (int *)d_atom_type;
    h.lj_pair  = (float2 *)d_lj_pair;
    h.lj_unscaled = 1;  // the caller's table is the reference's {A6, A12}
    h.cp_head  = (int *)d_idx_head_cell_pairs;
    h.cp_c2    = (int *)d_cell_pair_c2;
    h.cp_mask  = (uint64_t *)d_cell_pair_mask;
    h.fx       = (long long *)(base + kOffFx);
    h.e_acc    = (long long *)(base + kOffEacc);
    h.energy   = (double *)(base + kOffEnergy);
    h.cnt      = (og_dev_counters *)(base + kOffCnt);
    const og_dev_sys *d_sys = (const og_dev_sys *)(base + kOffSys);
    const int2       *d_map = (const int2 *)(base + kOffMap);
    hipStream_t       st    = (hipStream_t)stream;

    const int nb_setup = (4 * h.cap_atom_array + OG_THREADS - 1) / OG_THREADS;
    hipLaunchKernelGGL(k_na_setup, dim3(nb_setup), dim3(OG_THREADS), 0, st, base, h, n_cells);
    og_launch_pair(og_launch{d_sys, d_map, (n_cells + kWaves - 1) / kWaves, st}, 0, kWaves);
    og_launch_energy_reduce(og_launch{d_sys, d_map, 1, st});
    hipLaunchKernelGGL(k_na_out, dim3((n_atom_array + OG_THREADS - 1) / OG_THREADS), dim3(OG_THREADS), 0, st, d_sys,
                       d_work, d_energy, d_flags, n_atom_array);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        char b[200];
        std::snprintf(b, sizeof b, "og_pairwise_ljzd_noalloc: launch failed: %s", hipGetErrorString(e));
        return na_fail(OG_E_DEVICE, b);
    }
    return OG_OK;
}
