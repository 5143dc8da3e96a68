// og_engine.cpp -- host side of the MI355X engine: system/batch lifetime,
// device memory layout, block maps, the per-step launch plan, hipGraph capture
// and the C ABI declared in include/og_engine.h.
#include "og_internal.h"
#include "../../include/og_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>


namespace {

thread_local std::string g_last_error;

void set_err(const char *fmt, ...) {
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

#define OG_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);      \
            return OG_E_DEVICE;                                                                      \
        }                                                                                            \
    } while (0)

const double CHARGE_COEFF = 332.06378;  // CelesteObject.cpp:36
const double EPS_REF      = 1e-10;      // CelesteObject.cpp:31
const double PI_REF       = 3.14159265; // CelesteObject.cpp:26

// Grid geometry of MiniCell::set_grid_parameters/set_box_info/set_grid_xy
// (MiniCell.cpp:1279-1303, :1327-1362, :221-290), float semantics.
struct GridGeom {
    float L[3], lb[3], ub[3];
    float cutoff_pair, cutoff_pair_half, cutoff_pair2;
    int   nx, ny, nnx, nny, n_columns, n_uni_z;
    float L_cell[2], L_z_uni;
};

#pragma clang fp contract(off)
// edge_scale: the column width relative to MiniCell's (8 atoms per cube of
// the mean density); 1 reproduces the reference geometry (and list order)
int grid_geom(GridGeom &g, int n_atoms, const double *box, double nscut, double edge_scale = 1.0) {
    g.L[0] = (float)box[0];
    g.L[1] = (float)box[4];
    g.L[2] = (float)box[8];
    for (int i = 0; i < 3; i++) {
        const float Lh = (float)(g.L[i] * 0.5);
        g.lb[i]        = (float)(box[9 + i] - (double)Lh);
        g.ub[i]        = g.lb[i] + g.L[i];
    }
    g.cutoff_pair      = (float)nscut;
    g.cutoff_pair_half = (float)(g.cutoff_pair * 0.5);
    g.cutoff_pair2     = g.cutoff_pair * g.cutoff_pair;
    g.n_uni_z          = (int)(g.L[2] / g.cutoff_pair_half);
    g.L_z_uni          = g.L[2] / (float)g.n_uni_z;
    const float volume = g.L[0] * g.L[1] * g.L[2];
    float       w      = powf((float)((float)OG_NAC / ((float)n_atoms / volume)), (float)(1.0 / 3.0));
    if (edge_scale != 1.0) w = (float)(w * edge_scale);
    if (w > g.cutoff_pair_half) w = g.cutoff_pair_half;
    g.nx        = (int)floorf(g.L[0] / w);
    g.ny        = (int)floorf(g.L[1] / w);
    g.L_cell[0] = g.L[0] / (float)g.nx;
    g.L_cell[1] = g.L[1] / (float)g.ny;
    g.nnx       = (int)ceilf(g.cutoff_pair / g.L_cell[0]);
    g.nny       = (int)ceilf(g.cutoff_pair / g.L_cell[1]);
    g.n_columns = g.nx * g.ny;
    if (g.nnx > g.nx || g.nny > g.ny || g.n_uni_z < 2) return -1;
    return 0;
}

// ZeroMultipoleSum::set_zms_params (ZeroMultipoleSum.cpp:108-190), float members.
struct ZdConst {
    float cutoff, alpha, fcoeff, bcoeff, zcore, piewald;
    // higher orders: polynomial coefficients c_k of r^(2k+2) and the cores, index 0 quadrupole ... 2 hexadecapole
    float zc[3][4], zcc[3][4], zmcore[3];
};
ZdConst zd_const(double cutoff, double alpha) {
    ZdConst z;
    z.cutoff       = (float)cutoff;
    z.alpha        = (float)alpha;
    float c2       = z.cutoff * z.cutoff;
    float c3       = c2 * z.cutoff;
    float tmp      = z.cutoff * z.alpha;
    float errorfc  = erfcf(tmp);
    float expterm  = expf(-tmp * tmp);
    z.piewald      = 0.0f;
    if (z.alpha < EPS_REF) {
        z.fcoeff = (float)(1.0 / c3);
        z.bcoeff = (float)(0.5 * z.fcoeff);
    } else {
        float root_pi = (float)sqrt(PI_REF);
        z.piewald     = (float)(2.0 / root_pi * z.alpha);
        z.fcoeff      = errorfc / c3 + z.piewald * expterm / c2;
        z.bcoeff      = (float)(0.5 * z.fcoeff);
    }
    z.zcore = errorfc / z.cutoff + z.bcoeff * c2;
    // derivative coefficients of erfc(a r)/r at the cutoff; the expressions keep the reference's double
    // promotions (each rounds to a float member on assignment), including its 4th-derivative series
    const float c = z.cutoff, c4 = c2 * c2, c5 = c3 * c2, c6 = c3 * c3, c7 = c3 * c4;
    float       hz, d1, d2, d3, d4;
    if (z.alpha < EPS_REF) {
        hz = (float)(2.0 / c3);
        d1 = (float)(1.0 / c2);
        d2 = (float)(2.0 / c3);
        d3 = (float)(6.0 / c4);
        d4 = (float)(24.0 / c5);
    } else {
        const float rp = (float)sqrt(PI_REF);
        const float t3 = tmp * tmp * tmp, t5 = t3 * tmp * tmp, t7 = t5 * tmp * tmp;
        hz = (float)((errorfc + expterm * (tmp + t3) * 2.0 / rp) * 2.0 / c3);
        d1 = (float)((errorfc + expterm * tmp * 2.0 / rp) / c2);
        d2 = (float)((errorfc + expterm * (tmp + t3) * 2.0 / rp) * 2.0 / c3);
        d3 = (float)((errorfc + expterm * (tmp + 2.0 / 3.0 * t3 + 2.0 / 3.0 * t5) * 2.0 / rp) * 6.0 / c4);
        d4 = (float)((errorfc + expterm * (tmp + 2.0 / 3.0 * t3 + 1.0 / 6.0 * t5 + 1.0 / 3.0 + t7) * 2.0 / rp) * 24.0 / c5);
    }
    float(&q)[4] = z.zc[0], (&o)[4] = z.zc[1], (&h)[4] = z.zc[2];
    q[0] = (float)(-3.0 / 4.0 * z.fcoeff - hz / 4.0);
    q[1] = (float)(z.fcoeff / 8.0 / c2 + hz / 8.0 / c2);
    q[2] = q[3] = 0.0f;
    o[0] = (float)(-15.0 / 16.0 * d1 / c - 7.0 / 16.0 * d2 - 1.0 / 16.0 * d3 * c);
    o[1] = (float)(5.0 / 16.0 * d1 / c3 + 5.0 / 16.0 * d2 / c2 + 1.0 / 16.0 * d3 / c);
    o[2] = (float)(-1.0 / 16.0 * d1 / c5 - 1.0 / 16.0 * d2 / c4 - 1.0 / 48.0 * d3 / c3);
    o[3] = 0.0f;
    h[0] = (float)(-35.0 / 32.0 * d1 / c - 19.0 / 32.0 * d2 - 1.0 / 96.0 * d4 * c2 - 1.0 / 8.0 * d3 * c);
    h[1] = (float)(1.0 / 64.0 * d4 + 35.0 / 64.0 * d1 / c3 + 35.0 / 64.0 * d2 / c2 + 5.0 / 32.0 * d3 / c);
    h[2] = (float)(-7.0 / 32.0 * d1 / c5 - 7.0 / 32.0 * d2 / c4 - 1.0 / 12.0 * d3 / c3 - 1.0 / 96.0 * d4 / c2);
    h[3] = (float)(5.0 / 128.0 * d2 / c6 + 5.0 / 128.0 * d1 / c7 + 1.0 / 64.0 * d3 / c5 + 1.0 / 384.0 * d4 / c4);
    for (int m = 0; m < 3; m++)
        for (int k = 0; k < 4; k++) z.zcc[m][k] = (float)((2.0 * (k + 1)) * z.zc[m][k]);
    z.zmcore[0] = errorfc / c - q[0] * c2 - q[1] * c4;
    z.zmcore[1] = errorfc / c - o[0] * c2 - o[1] * c4 - o[2] * c6;
    z.zmcore[2] = errorfc / c - h[0] * c2 - h[1] * c4 - h[2] * c6 - h[3] * c6 * c2;
    return z;
}

// ZeroMultipoleSum::cal_self_energy (ZeroMultipoleSum.cpp:43-106): per-atom FP64
// terms summed into a float member (d_self_mon).
double self_energy(const ZdConst &z, float zmcore_f, int n, const float *q, int n_ex, const int *ex) {
    std::vector<double> dwork(n, 0.0);
    for (int i = 0; i < n_ex; i++) {
        dwork[ex[2 * i]] += q[ex[2 * i + 1]];
        dwork[ex[2 * i + 1]] += q[ex[2 * i]];
    }
    const double zmcore = zmcore_f;  // the order's core (ZeroMultipoleSum.cpp:72-80)
    float        mon    = 0.0f;
    for (int i = 0; i < n; i++) {
        const float  pq = -z.piewald * q[i] * q[i];
        const double e  = ((double)pq - zmcore * q[i] * (q[i] + dwork[i])) * 0.5 * CHARGE_COEFF;
        mon             = (float)((double)mon + e);
    }
    return (double)mon;
}
#pragma clang fp contract(on)

// exclusion-connectivity components (union-find over nb15off)
std::vector<int> components(int n, const int *nb15off) {
    std::vector<int> p(n);
    std::iota(p.begin(), p.end(), 0);
    auto find = [&](int x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    };
    for (int a = 0; a < n; a++)
        for (int k = 0; k < OG_MAX_NB15; k++) {
            const int b = nb15off[a * OG_MAX_NB15 + k];
            if (b < 0) break;
            int ra = find(a), rb = find(b);
            if (ra != rb) p[std::max(ra, rb)] = std::min(ra, rb);
        }
    for (int a = 0; a < n; a++) p[a] = find(a);
    return p;
}

template <class T> hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e == hipSuccess) e = hipMemset(*p, 0, sizeof(T) * n);
    return e;
}

int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }
// pair-kernel work units: cells plus the split cells' second halves, at most (the split depends on the
// system's cell count, which the capacity bounds)
int pair_units(const og_dev_sys &h) { return h.cap_cells + (h.cap_cells >= 2 * h.pair_split ? h.pair_split : 0); }

// MmSystem::revise_crd_inbox (MmSystem.cpp:808-816): coordinates into [lb, ub)
og_status wrap_inbox(const GridGeom &g, int n, double *crd) {
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            double &x = crd[3 * (size_t)i + k];
            if (!std::isfinite(x)) {
                set_err("non-finite input coordinate (atom %d)", i);
                return OG_E_NAN;
            }
            if (x < (double)g.lb[k] || x >= (double)g.ub[k]) x -= std::floor((x - g.lb[k]) / g.L[k]) * g.L[k];
            while (x < (double)g.lb[k]) x += g.L[k];
            while (x >= (double)g.ub[k]) x -= g.L[k];
        }
    return OG_OK;
}

}  // namespace

// error text for the call-point layer (og_celeste.cpp), read by og_last_error
void og_set_error(const char *msg) { g_last_error = msg; }

// ----------------------------------------------------------------- types --
enum { KIND_ATOM = 0, KIND_SLOT, KIND_CELLW, KIND_COL, KIND_SYS, KIND_MISC, KIND_TRIO, KIND_COMA, KIND_LSTEP,
       KIND_BTERM,  // OG_THREADS threads per block over the bonded terms (angles, torsions, impropers, 1-4)
       KIND_QUAD,   // OG_QUAD_THREADS threads per block over the four-atom SHAKE groups
       KIND_SCANB,  // blocks of 1024 cells over cap_cells + 1 (multi-block offset scan)
       KIND_PAIR,   // 16 pair-kernel maps: index log2(waves per block) + 4 * order (0 system order,
                    // 1 round-robin across systems, 2 each system's blocks on one XCD, 3 the system-order
                    // list cut into eight contiguous parts, one per XCD)
       KIND_N = KIND_PAIR + 16 };
constexpr int OG_N_XCD = 8;  // MI355X: workgroup i of a launch runs on XCD i mod 8
constexpr int OG_PAIR_DEFAULT = 2;  // 4 waves per block, blocks in system order

static_assert(OG_MC_VMCMD == OG_EXT_VMCMD && OG_MC_VAUS == OG_EXT_VAUS && OG_MC_NPOLY == OG_MC_MAX_ORDER + 1,
              "device extended-ensemble constants follow og_engine.h");

struct og_system {
    og_dev_sys          h{};           // host copy of the descriptor (device pointers)
    og_dev_sys         *d_self = nullptr;
    og_dev_counters    *d_cnt  = nullptr;
    og_params           prm{};
    GridGeom            geom{};
    std::vector<void *> allocs;
    int                 pair_form = 0;
    int                 d_free     = 0;
    long long           host_step  = 0;  // steps completed (host mirror of cnt->step)
    bool                grid_valid = false;
    double              pair_ms    = 0.0;
    int                 pair_launches = 0;
    bool                time_pair     = false;
    long long           step_end      = OG_NO_STEP_END;  // host mirror of og_dev_counters::step_end
    struct og_exec     *ctx = nullptr;   // private batch-of-one
    uint64_t            dev_bytes = 0;   // device memory allocated for the system
    og_status           status = OG_OK;  // outcome of the system's last run or evaluation
};

struct og_exec {
    std::vector<og_system *> members;
    og_dev_sys              *d_sys = nullptr;
    int2                    *d_map[KIND_N]{};
    int                      n_blocks[KIND_N]{};
    hipStream_t              stream = nullptr;
    bool                     dirty  = true;
    std::map<std::string, hipGraphExec_t> graphs;  // by rebuild pattern: '1' = rebuild before that step
    std::vector<hipEvent_t>  ev;          // pair-kernel timing events
    int                      pair_form = 0;
    bool                     any_langevin = false, any_leapfrog = false, any_com = false, any_constraint = false;
    bool                     any_vcmd = false, any_thermo = false;
    bool                     any_vmcmd = false;  // a member runs V-McMD
    bool                     any_bonded = false, any_quad = false;
    bool                     fused = false;       // every member takes the fused Langevin step
    int                      pair_shape = OG_PAIR_DEFAULT;  // pair-kernel launch shape (KIND_PAIR + pair_shape)
    int                      intvl = 1;
    int                      graph_steps = 0;     // steps per captured graph (0: automatic)
};

struct og_batch {
    og_exec ctx;
};

namespace {

void free_graphs(og_exec &x) {
    for (auto &kv : x.graphs) hipGraphExecDestroy(kv.second);
    x.graphs.clear();
}

og_status exec_release(og_exec &x) {
    free_graphs(x);
    for (auto &e : x.ev) hipEventDestroy(e);
    x.ev.clear();
    for (int k = 0; k < KIND_N; k++)
        if (x.d_map[k]) hipFree(x.d_map[k]), x.d_map[k] = nullptr;
    if (x.d_sys) hipFree(x.d_sys), x.d_sys = nullptr;
    if (x.stream) hipStreamDestroy(x.stream), x.stream = nullptr;
    return OG_OK;
}

// (re)build the descriptor array and block maps of an execution context
og_status exec_prepare(og_exec &x) {
    if (!x.dirty) return OG_OK;
    free_graphs(x);
    if (!x.stream) OG_HIP(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
    if (x.d_sys) hipFree(x.d_sys), x.d_sys = nullptr;
    for (int k = 0; k < KIND_N; k++)
        if (x.d_map[k]) hipFree(x.d_map[k]), x.d_map[k] = nullptr;
    const int ns = (int)x.members.size();
    if (ns == 0) {
        x.dirty = false;
        return OG_OK;
    }
    std::vector<og_dev_sys> hs(ns);
    std::vector<int2>       maps[KIND_N];
    x.pair_form   = x.members[0]->pair_form;
    x.any_langevin = x.any_leapfrog = x.any_com = x.any_constraint = x.any_vcmd = x.any_thermo = false;
    x.any_bonded = x.any_quad = x.any_vmcmd = false;
    x.fused        = true;
    x.intvl        = x.members[0]->prm.nsgrid_update_intvl;
    for (int si = 0; si < ns; si++) {
        og_system *s = x.members[si];
        if (s->pair_form != x.pair_form) {
            set_err("systems in one batch must share the electrostatics form (alpha = 0 or > 0)");
            return OG_E_ARG;
        }
        if (s->prm.nsgrid_update_intvl != x.intvl) {
            set_err("systems in one batch must share --nsgrid-update-intvl");
            return OG_E_ARG;
        }
        hs[si] = s->h;
        const og_dev_sys &d = s->h;
        x.any_langevin |= d.integrator == 1;
        x.any_leapfrog |= d.integrator == 0;
        x.any_com |= d.com_cancel && d.n_com_groups > 0;
        x.any_constraint |= d.constraint != 0;
        x.any_vcmd |= d.vc_n_states > 0 || d.mc_mode == OG_MC_VAUS;
        x.any_vmcmd |= d.mc_mode == OG_MC_VMCMD;
        x.fused &= d.mc_mode != OG_MC_VMCMD;  // V-McMD scales the whole gradient by the step's potential energy
        x.any_thermo |= d.thermostat == OG_THERMOSTAT_SCALING;
        x.any_bonded |= d.n_bterms > 0;
        x.any_quad |= d.constraint != 0 && d.n_quads > 0;
        x.fused &= d.integrator == 1 && d.n_water_blocks > 0;
        const int n_bins = (d.n_columns + 1) * d.n_uni_z;
        int       cnt[KIND_N];
        cnt[KIND_ATOM]  = cdiv(d.n_atoms, OG_THREADS);
        cnt[KIND_SLOT]  = cdiv(std::max(d.n_columns, n_bins), OG_THREADS);
        cnt[KIND_CELLW] = cdiv(d.cap_cells, OG_THREADS / 64);
        cnt[KIND_COL]   = d.n_columns;
        cnt[KIND_SYS]   = 1;
        cnt[KIND_MISC]  = cdiv(d.n_pairs_c, OG_THREADS);  // SHAKE pair constraints
        cnt[KIND_TRIO]  = cdiv(std::max(d.n_trios, d.n_settle), OG_TRIO_THREADS);
        cnt[KIND_COMA] = 0;  // filled below
        cnt[KIND_LSTEP] = d.n_water_blocks ? d.n_water_blocks + 1 : 0;  // free-atom block + water blocks
        cnt[KIND_BTERM] = cdiv(d.n_bterms, OG_THREADS);
        cnt[KIND_QUAD]  = d.constraint ? cdiv(d.n_quads, OG_QUAD_THREADS) : 0;
        cnt[KIND_SCANB] = d.cap_cells / 1024 + 1;
        for (int k = 0; k < KIND_PAIR; k++)
            for (int b = 0; b < cnt[k]; b++) maps[k].push_back(make_int2(si, b));
        for (int lw = 0; lw < 4; lw++) {  // pair kernel, blocks in system order
            const int nb = cdiv(pair_units(d), 1 << lw);
            for (int b = 0; b < nb; b++) maps[KIND_PAIR + lw].push_back(make_int2(si, b));
        }
    }
    // pair kernel, blocks round-robin across the systems of a batch
    for (int lw = 0; lw < 4; lw++) {
        int most = 0;
        for (int si = 0; si < ns; si++) most = std::max(most, cdiv(pair_units(x.members[si]->h), 1 << lw));
        for (int b = 0; b < most; b++)
            for (int si = 0; si < ns; si++)
                if (b < cdiv(pair_units(x.members[si]->h), 1 << lw)) maps[KIND_PAIR + 4 + lw].push_back(make_int2(si, b));
    }
    // pair kernel, each system's blocks on one XCD (launch positions p with p mod 8 fixed), so a system's
    // coordinates and lists stay in that XCD's L2; systems go to the XCD with the fewest blocks so far, largest
    // first, and once an XCD's list runs out its positions take blocks of the longest remaining list
    for (int lw = 0; lw < 4; lw++) {
        std::vector<int> nb(ns), order(ns);
        for (int si = 0; si < ns; si++) nb[si] = cdiv(pair_units(x.members[si]->h), 1 << lw), order[si] = si;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nb[a] > nb[b]; });
        std::vector<std::vector<int2>> lists(OG_N_XCD);
        std::vector<long long>         load(OG_N_XCD, 0);
        for (int si : order) {
            const int xc = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[xc] += nb[si];
            for (int b = 0; b < nb[si]; b++) lists[xc].push_back(make_int2(si, b));
        }
        std::vector<size_t> at(OG_N_XCD, 0);
        size_t              total = 0;
        for (auto &l : lists) total += l.size();
        for (size_t p = 0; p < total; p++) {
            int xc = (int)(p % OG_N_XCD);
            if (at[xc] >= lists[xc].size()) {
                size_t most = 0;
                for (int k = 0; k < OG_N_XCD; k++)
                    if (lists[k].size() - at[k] > most) most = lists[k].size() - at[k], xc = k;
            }
            maps[KIND_PAIR + 8 + lw].push_back(lists[xc][at[xc]++]);
        }
    }
    // pair kernel, the system-order list in eight contiguous parts, part k at the launch positions of XCD k:
    // neighbouring cells (and so the atoms their rows share) stay on one XCD's L2
    for (int lw = 0; lw < 4; lw++) {
        const std::vector<int2> &all = maps[KIND_PAIR + lw];
        const size_t             T   = all.size();
        size_t                   start[OG_N_XCD + 1];
        start[0] = 0;
        for (int k = 0; k < OG_N_XCD; k++) start[k + 1] = start[k] + T / OG_N_XCD + ((size_t)k < T % OG_N_XCD);
        for (size_t i = 0; i < T / OG_N_XCD + 1; i++)
            for (int k = 0; k < OG_N_XCD; k++)
                if (start[k] + i < start[k + 1]) maps[KIND_PAIR + 12 + lw].push_back(all[start[k] + i]);
    }
    // COM-apply map from per-system totals (kept on the system)
    maps[KIND_COMA].clear();
    for (int si = 0; si < ns; si++) {
        og_system *s = x.members[si];
        const int  n = (s->h.com_cancel && s->h.n_com_groups) ? s->h.n_atom_blocks : 0;  // com atoms <= n_atoms
        for (int b = 0; b < n; b++) maps[KIND_COMA].push_back(make_int2(si, b));
    }
    OG_HIP(hipMalloc((void **)&x.d_sys, sizeof(og_dev_sys) * ns));
    OG_HIP(hipMemcpy(x.d_sys, hs.data(), sizeof(og_dev_sys) * ns, hipMemcpyHostToDevice));
    for (int k = 0; k < KIND_N; k++) {
        x.n_blocks[k] = (int)maps[k].size();
        if (x.n_blocks[k]) {
            OG_HIP(hipMalloc((void **)&x.d_map[k], sizeof(int2) * maps[k].size()));
            OG_HIP(hipMemcpy(x.d_map[k], maps[k].data(), sizeof(int2) * maps[k].size(), hipMemcpyHostToDevice));
        }
    }
    x.dirty = false;
    return OG_OK;
}

og_launch L(const og_exec &x, int kind) { return og_launch{x.d_sys, x.d_map[kind], x.n_blocks[kind], x.stream}; }

void enqueue_rebuild(og_exec &x) {
    og_launch_grid_clear(L(x, KIND_SLOT), L(x, KIND_SLOT));
    og_launch_grid_bin(L(x, KIND_ATOM));
    og_launch_grid_scan(L(x, KIND_SYS));
    og_launch_grid_scatter(L(x, KIND_ATOM));
    og_launch_grid_sort(L(x, KIND_COL));
    og_launch_enum_count(L(x, KIND_CELLW));
    og_launch_enum_scan_blocks(L(x, KIND_SCANB));
    og_launch_enum_write(L(x, KIND_CELLW));
    og_launch_masks(L(x, KIND_CELLW));
}

// reduce: reduce the energies now (a single-point call); an MD step reduces
// them in its closing kernel instead
void enqueue_forces(og_exec &x, hipEvent_t e0, hipEvent_t e1, bool reduce = false) {
    if (e0) hipEventRecord(e0, x.stream);
    og_launch_pair(L(x, KIND_PAIR + x.pair_shape), x.pair_form, 1 << (x.pair_shape & 3));
    if (e1) hipEventRecord(e1, x.stream);
    // angles, torsions, impropers, 1-4 pairs: each atom's shares, summed with its pair gradient below (or in
    // the fused step kernel)
    if (x.any_bonded) og_launch_bonded(L(x, KIND_BTERM));
    // an MD step of a fused context converts the gradient inside its step kernel
    if (reduce || !x.fused) og_launch_forces_to_work(L(x, KIND_ATOM));
    if (reduce) og_launch_energy_reduce(L(x, KIND_SYS));
}

void enqueue_integrate(og_exec &x) {
    if (x.fused) {
        og_launch_langevin_step(L(x, KIND_LSTEP));
        og_launch_step_close(L(x, KIND_SYS));
        return;
    }
    // the extended-ensemble bias acts on the gradient of the step
    // (DynamicsMode.cpp:504-512, :824-834)
    if (x.any_vcmd) og_launch_vcmd(L(x, KIND_SYS));
    if (x.any_vmcmd) og_launch_vmcmd(L(x, KIND_SYS), L(x, KIND_ATOM));
    if (x.any_leapfrog) {
        og_launch_leapfrog_vel(L(x, KIND_ATOM));
        if (x.any_com) {
            og_launch_com_sum(L(x, KIND_SYS));
            og_launch_com_apply(L(x, KIND_COMA));
        }
        // SubBox::update_thermostat + apply_thermostat (DynamicsMode.cpp:525-532)
        if (x.any_thermo) og_launch_thermostat(L(x, KIND_ATOM));
        og_launch_positions(L(x, KIND_ATOM));
    }
    if (x.any_langevin) og_launch_langevin_positions(L(x, KIND_ATOM));
    if (x.any_constraint) og_launch_constraints(L(x, KIND_TRIO), L(x, KIND_MISC));
    // four-atom groups after pairs and trios (ConstraintShake::apply_constraint, ConstraintShake.cpp:10-16)
    if (x.any_quad) og_launch_shake_quad(L(x, KIND_QUAD));
    og_launch_finish(L(x, KIND_ATOM));
    og_launch_step_close(L(x, KIND_SYS));
}

// Ask for a grid rebuild of this system before its next step whatever its
// phase (a new system, new coordinates, an explicit rebuild); the step's
// closing kernel (or the single-point reduction) clears the request.
og_status request_rebuild(og_system *s, int on = 1) {
    OG_HIP(hipMemcpy(&s->d_cnt->rebuild_now, &on, sizeof(int), hipMemcpyHostToDevice));
    return OG_OK;
}

void enqueue_step(og_exec &x, bool rebuild, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (rebuild) enqueue_rebuild(x);
    enqueue_forces(x, e0, e1);
    enqueue_integrate(x);
}

// Rebuild pattern of the next K steps: step k gets the rebuild kernels when
// some running member is due then (its own step a multiple of the interval)
// or, at k = 0, when a rebuild was requested; the kernels skip the members
// that are not due, so members at different phases share the launches.
std::string rebuild_pattern(const og_exec &x, int K, long long offset, bool forced) {
    std::string p((size_t)K, '0');
    for (const og_system *s : x.members) {
        const long long st = std::min(s->host_step + offset, s->step_end);
        if (st >= s->step_end) continue;  // frozen for this chunk
        for (int k = 0; k < K && st + k < s->step_end; k++)
            if ((st + k) % x.intvl == 0) p[(size_t)k] = '1';
    }
    if (forced && K > 0) p[0] = '1';
    return p;
}

// One graph per rebuild pattern (its length is the chunk's step count).
og_status get_graph(og_exec &x, const std::string &pattern, hipGraphExec_t *out) {
    auto it = x.graphs.find(pattern);
    if (it != x.graphs.end()) {
        *out = it->second;
        return OG_OK;
    }
    if (x.graphs.size() >= 32) free_graphs(x);  // bound the cache (serving mixes phases)
    hipGraph_t g;
    OG_HIP(hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
    for (char c : pattern) enqueue_step(x, c == '1');
    OG_HIP(hipStreamEndCapture(x.stream, &g));
    hipGraphExec_t ge;
    OG_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    x.graphs[pattern] = ge;
    *out              = ge;
    return OG_OK;
}

// Reads every member's counters: its step, and a status from its flags.  A
// faulted member stands still on the device (og_frozen); its sticky state is
// cleared here so it can be run again once the caller has acted on the
// error, and its grid is rebuilt first.  Returns the first member's error.
og_status check_flags(og_exec &x) {
    og_status first = OG_OK;
    for (og_system *s : x.members) {
        og_dev_counters c;
        OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
        s->host_step = c.step;
        og_status st = OG_OK;
        char      msg[256];
        if (c.flags & (OG_F_CP_OVERFLOW | OG_F_CELL_OVERFLOW | OG_F_COL_OVERFLOW)) {
            st = OG_E_CAPACITY;
            snprintf(msg, sizeof msg, "device list overflow (flags 0x%x): cells %d, cell pairs cap %d", c.flags,
                     c.n_cells, s->h.cap_cell_pairs);
        } else if (c.flags & OG_F_NAN) {
            st = OG_E_NAN;
            snprintf(msg, sizeof msg, "non-finite coordinate or force (flags 0x%x) at step %lld", c.flags, c.step);
        } else if (c.shake_fail) {
            st = OG_E_NOCONV;
            snprintf(msg, sizeof msg, "SHAKE did not converge for %d constraint groups", c.shake_fail);
        }
        s->status = st;
        // the device flags are sticky: clear them once reported
        if (c.flags & ~OG_F_OUTSIDE_BOX) OG_HIP(hipMemset(&s->d_cnt->flags, 0, sizeof(int)));
        if (c.shake_fail) OG_HIP(hipMemset(&s->d_cnt->shake_fail, 0, sizeof(int)));
        if (st == OG_E_CAPACITY || st == OG_E_NAN) {
            // a step stopped part-way: drop its partial sums, rebuild before the next one
            OG_HIP(hipMemset(s->h.fx, 0, sizeof(long long) * OG_FX_WORDS(s->h.cap_atom_array)));
            OG_HIP(hipMemset(s->h.e_acc, 0, sizeof(long long) * OG_EACC_SLOTS * OG_EACC_STRIDE));
            s->grid_valid = false;
        }
        if (st != OG_OK && first == OG_OK) {
            first = st;
            set_err("%s", msg);
        }
    }
    return first;
}

// Run n steps for every member as graph replays of up to graph_steps (100 by
// default) steps; a member at any rebuild phase can join.  With timing on,
// steps are launched directly with HIP events around each pair launch
// (events inside a captured graph cannot be timed).
og_status exec_run(og_exec &x, int n_steps, bool timing) {
    og_status st = exec_prepare(x);
    if (st != OG_OK) return st;
    if (x.members.empty() || n_steps <= 0) return OG_OK;
    bool                     forced = false;
    std::vector<og_system *> running;
    for (og_system *s : x.members) {
        if (s->host_step >= s->step_end) continue;
        running.push_back(s);
        if (!s->grid_valid) {  // new system or new coordinates: rebuild before its first step
            st = request_rebuild(s);
            if (st != OG_OK) return st;
            forced = true;
        }
    }
    if (running.empty()) return OG_OK;
    for (og_system *s : x.members) {
        s->pair_ms       = 0.0;
        s->pair_launches = 0;
    }
    int done = 0;
    if (timing) {
        for (int k = (int)x.ev.size(); k < 2 * n_steps; k++) {
            hipEvent_t e;
            OG_HIP(hipEventCreate(&e));
            x.ev.push_back(e);
        }
        const std::string p = rebuild_pattern(x, n_steps, 0, forced);
        for (int k = 0; k < n_steps; k++) enqueue_step(x, p[(size_t)k] == '1', x.ev[2 * k], x.ev[2 * k + 1]);
        OG_HIP(hipStreamSynchronize(x.stream));
        double ms = 0.0;
        for (int k = 0; k < n_steps; k++) {
            float t = 0.0f;
            OG_HIP(hipEventElapsedTime(&t, x.ev[2 * k], x.ev[2 * k + 1]));
            ms += t;
        }
        for (og_system *s : x.members) {
            s->pair_ms += ms;
            s->pair_launches += n_steps;
        }
        done = n_steps;
    }
    const int kmax = x.graph_steps > 0 ? x.graph_steps : 100;
    while (done < n_steps) {
        const int      K = std::min(kmax, n_steps - done);
        hipGraphExec_t ge;
        st = get_graph(x, rebuild_pattern(x, K, done, forced && done == 0), &ge);
        if (st != OG_OK) return st;
        OG_HIP(hipGraphLaunch(ge, x.stream));
        done += K;
    }
    OG_HIP(hipStreamSynchronize(x.stream));
    for (og_system *s : running) s->grid_valid = true;  // the members that stood still keep their state
    return check_flags(x);
}

// ------------------------------------------------------------ allocation --
// Cell-pair list capacity from the candidate geometry: cells per column
// height, candidate cells per cell1 within the list radius, half of them
// owned, 25% headroom.
int pair_capacity(const GridGeom &g, int n, int cap_cells, double capacity_factor) {
    const double per_col = (double)n / g.n_columns;
    const double hcell   = (double)g.L[2] * OG_NAC / std::max(per_col, 1.0);
    double       cand    = 0.0;
    for (int dx = -g.nnx; dx <= g.nnx; dx++)
        for (int dy = -g.nny; dy <= g.nny; dy++) {
            const double ex = std::max(0, std::abs(dx) - 1) * (double)g.L_cell[0];
            const double ey = std::max(0, std::abs(dy) - 1) * (double)g.L_cell[1];
            const double r2 = (double)g.cutoff_pair2 - ex * ex - ey * ey;
            if (r2 < 0) continue;
            cand += (2.0 * std::sqrt(r2) + 2.0 * hcell) / hcell + 1.0;
        }
    const double est = 0.5 * cand * (double)cap_cells;
    const double fac = capacity_factor > 0 ? capacity_factor : 1.0;
    return (int)std::min(2.0e9, est * 1.25 * fac + 4096.0);
}

og_status sys_alloc(og_system *s, const og_system_desc *d, const og_params *p) {
    og_dev_sys &h = s->h;
    auto        A = [&](auto **ptr, size_t n) -> bool {
        hipError_t e = dalloc(ptr, n);
        if (e != hipSuccess) {
            set_err("hipMalloc of %zu elements failed: %s", n, hipGetErrorString(e));
            return false;
        }
        s->allocs.push_back((void *)*ptr);
        s->dev_bytes += n * sizeof(**ptr);
        return true;
    };
    const int n = d->n_atoms;
    h.n_atoms   = n;
    const GridGeom &g = s->geom;
    h.n_columns = g.n_columns;
    h.nx = g.nx, h.ny = g.ny, h.nnx = g.nnx, h.nny = g.nny, h.n_uni_z = g.n_uni_z;
    h.cap_atom_array = (n + g.n_columns * OG_NAC + 31) / 32 * 32;  // whole groups of four cells (og_fx_slot)
    h.cap_cells      = h.cap_atom_array / OG_NAC;
    h.pair_split     = OG_PAIR_SPLIT;
    h.cap_cell_pairs   = pair_capacity(g, n, h.cap_cells, p->capacity_factor);
    h.n_types          = d->n_lj_types;
    h.n_excess         = d->n_excess;
    h.n_bonds          = d->n_bonds;
    h.n_pairs_c        = p->constraint ? d->n_shake_pairs : 0;
    h.n_trios          = p->constraint ? d->n_shake_trios : 0;
    h.n_settle         = p->constraint == OG_CONSTRAINT_SHAKE_SETTLE ? d->n_settle : 0;
    h.n_quads          = p->constraint ? d->n_shake_quads : 0;
    h.n_angles         = d->n_angles;
    h.n_torsions       = d->n_torsions;
    h.n_impros         = d->n_impros;
    h.n_nb14           = d->n_nb14;
    h.n_bterms         = h.n_angles + h.n_torsions + h.n_impros + h.n_nb14;
    h.n_com_groups     = p->com_cancel ? d->n_com_groups : 0;
    h.integrator       = p->integrator;
    h.constraint       = p->constraint;
    h.com_cancel       = p->com_cancel && d->n_com_groups > 0;
    h.gap_bug          = p->gap_bug;
    h.excess_bug       = p->excess_bug;
    h.rebuild_intvl    = p->nsgrid_update_intvl;
    h.max_loops        = p->const_max_loops;
    h.log_stride       = std::max(1, p->log_stride);
    h.log_cap          = std::max(1, p->log_cap);
    for (int k = 0; k < 3; k++) h.L[k] = g.L[k], h.lb[k] = g.lb[k], h.ub[k] = g.ub[k];
    h.L_cell[0] = g.L_cell[0], h.L_cell[1] = g.L_cell[1];
    h.L_z_uni      = g.L_z_uni;
    h.cutoff_pair2 = g.cutoff_pair2;
    ZdConst z = zd_const(p->cutoff, p->ele_alpha);
    h.cutoff  = z.cutoff;
    h.cutoff2 = z.cutoff * z.cutoff;
    h.zcore = z.zcore, h.bcoeff = z.bcoeff, h.fcoeff = z.fcoeff, h.piewald = z.piewald, h.alpha = z.alpha;
    const bool dh    = p->electrostatic == OG_ELE_DEBYE_HUCKEL;
    const int  order = p->electrostatic == OG_ELE_ZERO_QUADRUPOLE  ? 1
                       : p->electrostatic == OG_ELE_ZERO_OCTUPOLE  ? 2
                       : p->electrostatic == OG_ELE_ZERO_HEXADECAPOLE ? 3
                                                                    : 0;
    const bool alpha_on = z.alpha > 0.0f;
    s->pair_form = dh ? OG_FORM_HPS_DH
                      : order ? (alpha_on ? OG_FORM_ZM_ALPHA : OG_FORM_ZM) : (alpha_on ? OG_FORM_ZD_ALPHA : OG_FORM_ZD);
    float zmcore = z.zcore;
    h.zm_order   = order;
    if (order) {
        zmcore    = z.zmcore[order - 1];
        h.zm_core = zmcore;
        for (int k = 0; k < 4; k++) {
            h.zm_c[k]  = z.zc[order - 1][k];
            h.zm_cc[k] = z.zcc[order - 1][k];
            // ZeroMultipoleSum::calc_zero08pole (:392-408) takes the hexadecapole coefficients for its gradient
            h.zm_pg[k] = (order == 2 && alpha_on && p->zo_grad_bug) ? (k < 3 ? z.zcc[2][k] : 0.0f) : h.zm_cc[k];
        }
    }
    if (dh) {
        // ZeroMultipoleSum constructor (ZeroMultipoleSum.cpp:21-24): float members from the double cfg values;
        // constants CelesteObject.cpp:32-40
        h.dh_diel_inv = (float)(1.0 / p->dh_dielectric);
        h.dh_kappa    = (float)(1.0 / (sqrt(8.85418782e-12 * p->dh_dielectric * 1.380658e-23 * p->dh_temperature /
                                            (2 * 6.0221413e23 * 1.60217657e-19 * 1.60217657e-19 *
                                             p->dh_ionic_strength)) * 1e10));
        h.dh_grad_fix = p->dh_grad_fix != 0;
        h.hps_eps     = p->nonbond == OG_NONBOND_HPS ? (float)p->hps_epsilon : 0.0f;
        // excluded pairs add nothing under Debye-Hueckel (calc_null, ZeroMultipoleSum.cpp:456-466) and the
        // self energy has no core term (zmcore = 0 outside the zero-multipole modes, :75-82)
        h.n_excess = 0;
        z.zcore    = 0.0f;
        zmcore     = 0.0f;
    }
    h.dt          = p->time_step;
    h.tol         = p->const_tolerance;
    h.temperature = p->temperature;
    h.gamma       = p->langevin_gamma;
    h.thermostat    = p->thermostat;
    if (h.thermostat == OG_THERMOSTAT_SCALING) {
        // SubBox::init_thermostat (SubBox.cpp:1660-1679): tau replaced by the step when the step is smaller;
        // ThermostatObject::set_temperature_coeff (Thermostat.cpp:13-17)
        double tau = p->berendsen_tau;
        if (p->time_step < tau) tau = p->time_step;
        h.th_tau_inv    = 1.0 / tau;
        h.th_temp_init  = p->temperature_init < 0 ? p->temperature : p->temperature_init;
        h.heating_steps = p->heating_steps;
        h.th_temp_coeff = (2.0 * 4.184 * 1e+3) / (8.31451 * (double)d->d_free) * ((1e7 / (4.184 * 1e3)) * 0.5);
    }
    h.settle_dOH  = d->settle_dOH;
    h.settle_dHH  = d->settle_dHH;
    h.seed_lo     = (uint32_t)p->random_seed;
    h.seed_hi     = (uint32_t)(p->random_seed >> 32) ^ 0x9E3779B9u;
    h.vc_n_states = d->vc_n_states;
    h.vc_interval = std::max(1, d->vc_interval);
    const bool groups_on = d->vc_n_states > 0 || d->mc_mode == OG_EXT_VAUS;  // VcMD or V-AUS groups
    h.vc_na       = groups_on ? d->vc_n_a : 0;
    h.vc_nb       = groups_on ? d->vc_n_b : 0;
    h.mc_mode     = d->mc_mode;
    h.mc_n_states = d->mc_mode ? d->mc_n_states : 0;
    h.mc_interval = std::max(1, d->mc_interval);
    // ExtendedVMcMD::set_temperature (Extend.cpp:119-122) with the input section's temperature
    h.mc_constk = (8.31451 / 4.184) * 1e-3 * (double)d->mc_temperature;
    h.vc_sigma    = p->enhance_sigma;
    h.vc_coef     = p->enhance_recov_coef;
    // ExtendedVcMD::set_temperature (Extend.cpp:573-576): k_B T in kcal/mol
    h.vc_constk = (8.31451 / 4.184) * 1e-3 * p->temperature;

    std::vector<float> q(n), m(n), minv(n);
    for (int i = 0; i < n; i++) {
        q[i]    = (float)d->charge[i];
        m[i]    = (float)d->mass[i];
        minv[i] = (float)(1.0 / (double)m[i]);  // SubBox.cpp:691 (real_pw mass_inv = 1.0 / mass)
    }
    h.e_self = self_energy(z, zmcore, n, q.data(), h.n_excess, d->excess);

    const int cap = h.cap_atom_array;
    bool      ok  = A(&h.crd, 3 * n) && A(&h.crd_prev, 3 * n) && A(&h.crd_prev2, 3 * n) && A(&h.vel, 3 * n) &&
              A(&h.vel_next, 3 * n) && A(&h.work, 3 * n) && A(&h.mass, n) && A(&h.mass_inv, n) && A(&h.charge, n) &&
              A(&h.atom_type, n) && A(&h.nb15off, (size_t)n * OG_MAX_NB15) && A(&h.comp, n) &&
              A(&h.lj6, (size_t)h.n_types * h.n_types) && A(&h.lj12, (size_t)h.n_types * h.n_types) &&
              A(&h.lj_pair, (size_t)h.n_types * h.n_types) &&
              A(&h.lj_hps, s->pair_form == OG_FORM_HPS_DH ? (size_t)h.n_types * h.n_types : 1) &&
              A(&h.excess, 2 * (size_t)h.n_excess) && A(&h.bond_ids, 2 * (size_t)h.n_bonds) &&
              A(&h.bond_k, h.n_bonds) && A(&h.bond_r0, h.n_bonds) && A(&h.trio_ids, 3 * (size_t)h.n_trios) &&
              A(&h.trio_d2, 3 * (size_t)h.n_trios) && A(&h.pairc_ids, 2 * (size_t)h.n_pairs_c) &&
              A(&h.pairc_d2, h.n_pairs_c) && A(&h.settle_ids, 3 * (size_t)h.n_settle) &&
              A(&h.com_off, h.n_com_groups + 1) && A(&h.com_atoms, h.n_com_groups ? d->com_group_off[h.n_com_groups] : 1) &&
              A(&h.com_minv, h.n_com_groups + 1) && A(&h.col_of_atom, n) && A(&h.zkey, n) &&
              A(&h.n_atoms_col, h.n_columns) && A(&h.col_fill, h.n_columns) && A(&h.head_cell, h.n_columns + 1) &&
              A(&h.head_atom, h.n_columns + 1) && A(&h.n_cells_z, h.n_columns) && A(&h.tmp_ids, cap) &&
              A(&h.atomids, cap) && A(&h.rev, n) && A(&h.gcrd, 3 * (size_t)cap) && A(&h.gxyzq, cap) &&
              A(&h.gtype, cap) && A(&h.gcomp, cap) && A(&h.cell_col, h.cap_cells) && A(&h.cell_z, h.cap_cells) &&
              A(&h.cell_zmin, h.cap_cells) && A(&h.cell_zmax, h.cap_cells) && A(&h.cell_u0, h.cap_cells) &&
              A(&h.cell_u1, h.cap_cells) && A(&h.u2c_min, (size_t)(h.n_columns + 2) * h.n_uni_z) &&
              A(&h.u2c_max, (size_t)(h.n_columns + 2) * h.n_uni_z) && A(&h.cp_count, h.cap_cells + 1) && A(&h.cp_stage, (size_t)h.cap_cells * OG_ENUM_STAGE) &&
              A(&h.cp_head, h.cap_cells + 1) && A(&h.cp_bsum, h.cap_cells / 1024 + 1) && A(&h.cp_c2, h.cap_cell_pairs) && A(&h.cp_mask, h.cap_cell_pairs) &&
              A(&h.fx, OG_FX_WORDS(cap)) && A(&h.e_acc, (size_t)OG_EACC_SLOTS * OG_EACC_STRIDE);
    if (!ok) return OG_E_DEVICE;
    const int n_tor     = h.n_torsions + h.n_impros;
    const int n_contrib = 3 * h.n_angles + 4 * n_tor + 2 * h.n_nb14;
    if (h.n_bterms) {
        ok = A(&h.ang_ids, std::max(3 * (size_t)h.n_angles, (size_t)1)) &&
             A(&h.ang_par, std::max((size_t)h.n_angles, (size_t)1)) &&
             A(&h.tor_ids, std::max(4 * (size_t)n_tor, (size_t)1)) && A(&h.tor_par, std::max((size_t)n_tor, (size_t)1)) &&
             A(&h.nb14_ids, std::max(2 * (size_t)h.n_nb14, (size_t)1)) &&
             A(&h.nb14_tp, std::max((size_t)h.n_nb14, (size_t)1)) &&
             A(&h.nb14_coef, std::max((size_t)h.n_nb14, (size_t)1)) && A(&h.bcontrib, 3 * (size_t)n_contrib) &&
             A(&h.bc_off, (size_t)n + 1) && A(&h.bc_at, (size_t)n_contrib);
        if (!ok) return OG_E_DEVICE;
    }
    if (h.n_quads && !(A(&h.quad_ids, 4 * (size_t)h.n_quads) && A(&h.quad_d2, 6 * (size_t)h.n_quads)))
        return OG_E_DEVICE;
    h.n_atom_blocks = cdiv(n, OG_THREADS);
    h.n_ke_parts    = std::max(h.n_atom_blocks, cdiv(h.n_settle, OG_WATERS_PER_BLOCK) + 1);
    if (h.thermostat == OG_THERMOSTAT_SCALING && !A(&h.th_part, h.n_atom_blocks)) return OG_E_DEVICE;
    ok = A(&h.e_atom, std::max(h.n_ke_parts, 1)) &&
         A(&h.com_mom, 3 * (size_t)(h.n_com_groups + 1)) && A(&h.energy, OG_E_NTERMS) &&
         A(&h.elog, (size_t)h.log_cap * OG_E_NTERMS) && A(&s->d_cnt, 1) && A(&s->d_self, 1);
    if (!ok) return OG_E_DEVICE;
    h.cnt = s->d_cnt;
    {
        og_dev_counters c0{};
        c0.step_end   = OG_NO_STEP_END;
        c0.noise_step = -1;
        OG_HIP(hipMemcpy(s->d_cnt, &c0, sizeof c0, hipMemcpyHostToDevice));
    }

    // uploads
    // LJ types are renumbered by frequency (most atoms first, ties by the input number) so the pair kernel's
    // LDS copy of the leading block of the table covers most pairs; the table entries are the same floats.
    const int nt = h.n_types;
    std::vector<int> tcount(nt, 0), torder(nt), tnew(nt), type_new(n);
    for (int i = 0; i < n; i++) tcount[d->atom_type[i]]++;
    std::iota(torder.begin(), torder.end(), 0);
    std::stable_sort(torder.begin(), torder.end(), [&](int a, int b) { return tcount[a] > tcount[b]; });
    for (int k = 0; k < nt; k++) tnew[torder[k]] = k;
    for (int i = 0; i < n; i++) type_new[i] = tnew[d->atom_type[i]];
    std::vector<float> lj6f((size_t)nt * nt), lj12f(lj6f.size());
    for (int a = 0; a < nt; a++)
        for (int b = 0; b < nt; b++) {
            const size_t src = (size_t)torder[a] * nt + torder[b], dst = (size_t)a * nt + b;
            lj6f[dst]  = (float)d->lj6[src];
            lj12f[dst] = (float)d->lj12[src];
        }
    std::vector<int> comp = components(n, d->nb15off);
    std::vector<double> crd0(d->crd, d->crd + 3 * (size_t)n);
    og_status           wst = wrap_inbox(g, n, crd0.data());
    if (wst != OG_OK) return wst;
    OG_HIP(hipMemcpy(h.crd, crd0.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.crd_prev, crd0.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.crd_prev2, crd0.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.vel_next, d->vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.vel, d->vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.mass, m.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.mass_inv, minv.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.charge, q.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.atom_type, type_new.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.nb15off, d->nb15off, sizeof(int) * (size_t)n * OG_MAX_NB15, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.comp, comp.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.lj6, lj6f.data(), sizeof(float) * lj6f.size(), hipMemcpyHostToDevice));
    OG_HIP(hipMemcpy(h.lj12, lj12f.data(), sizeof(float) * lj12f.size(), hipMemcpyHostToDevice));
    {
        std::vector<float2> ljp(lj6f.size());
        // the pair kernel's scaled entries {6 A6, 12 A12}
        for (size_t k = 0; k < ljp.size(); k++) ljp[k] = make_float2(6.0f * lj6f[k], 12.0f * lj12f[k]);
        OG_HIP(hipMemcpy(h.lj_pair, ljp.data(), sizeof(float2) * ljp.size(), hipMemcpyHostToDevice));
    }
    if (s->pair_form == OG_FORM_HPS_DH) {
        // {6 A6, 12 A12, switch distance, lambda} in the renumbered type order; plain LJ (nonbond lennard-jones)
        // is the HPS form with lambda 1 and switch distance 0
        const bool          hps = p->nonbond == OG_NONBOND_HPS;
        std::vector<float4> t4((size_t)nt * nt);
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                const size_t src = (size_t)torder[a] * nt + torder[b], dst = (size_t)a * nt + b;
                t4[dst] = make_float4(6.0f * lj6f[dst], 12.0f * lj12f[dst], hps ? (float)d->hps_cutoff[src] : 0.0f,
                                      hps ? (float)d->hps_lambda[src] : 1.0f);
            }
        OG_HIP(hipMemcpy(h.lj_hps, t4.data(), sizeof(float4) * t4.size(), hipMemcpyHostToDevice));
    }
    if (h.n_excess) OG_HIP(hipMemcpy(h.excess, d->excess, sizeof(int) * 2 * h.n_excess, hipMemcpyHostToDevice));
    if (h.n_bonds) {
        OG_HIP(hipMemcpy(h.bond_ids, d->bond_ids, sizeof(int) * 2 * h.n_bonds, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.bond_k, d->bond_k, sizeof(double) * h.n_bonds, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.bond_r0, d->bond_r0, sizeof(double) * h.n_bonds, hipMemcpyHostToDevice));
    }
    {
        // per-atom lists of the excess pairs and bonds each atom takes part in,
        // in ascending list order (the per-atom gradient sums follow it)
        // (entries carry both atoms and the pair index, one 16-byte load each)
        auto csr = [&](const int *pairs, int np, int **off_d, int4 **at_d) -> og_status {
            std::vector<int>  off(n + 1, 0);
            std::vector<int4> at(2 * (size_t)np);
            for (int t = 0; t < np; t++) off[pairs[2 * t] + 1]++, off[pairs[2 * t + 1] + 1]++;
            for (int i = 0; i < n; i++) off[i + 1] += off[i];
            std::vector<int> fill(off.begin(), off.end() - 1);
            for (int t = 0; t < np; t++) {
                const int4 e            = make_int4(pairs[2 * t], pairs[2 * t + 1], t, 0);
                at[fill[pairs[2 * t]]++] = e;
                at[fill[pairs[2 * t + 1]]++] = e;
            }
            if (!A(off_d, n + 1) || !A(at_d, std::max<size_t>(at.size(), 1))) return OG_E_DEVICE;
            OG_HIP(hipMemcpy(*off_d, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
            if (np) OG_HIP(hipMemcpy(*at_d, at.data(), sizeof(int4) * at.size(), hipMemcpyHostToDevice));
            return OG_OK;
        };
        og_status st = csr(d->excess, h.n_excess, &h.exc_off, &h.exc_at);
        if (st != OG_OK) return st;
        st = csr(d->bond_ids, h.n_bonds, &h.bond_off, &h.bond_at);
        if (st != OG_OK) return st;
    }
    if (h.n_bterms) {
        // term tables; 1-4 type pairs in the renumbered LJ table; per-atom contribution lists in slot order
        std::vector<int>     ids4(4 * (size_t)n_tor), tp(h.n_nb14), off(n + 1, 0), at(n_contrib);
        std::vector<double2> ap(h.n_angles), cf(h.n_nb14);
        std::vector<double4> tpar(n_tor);
        for (int t = 0; t < h.n_angles; t++) ap[t] = make_double2(d->angle_k[t], d->angle_theta0[t]);
        for (int t = 0; t < n_tor; t++) {
            const bool    im  = t >= h.n_torsions;
            const int     u   = im ? t - h.n_torsions : t;
            const int    *src = im ? d->impro_ids : d->torsion_ids;
            for (int k = 0; k < 4; k++) ids4[4 * (size_t)t + k] = src[4 * (size_t)u + k];
            tpar[t] = im ? make_double4(d->impro_ene[u], (double)d->impro_overlaps[u], (double)d->impro_symmetry[u],
                                        d->impro_phase[u])
                         : make_double4(d->torsion_ene[u], (double)d->torsion_overlaps[u],
                                        (double)d->torsion_symmetry[u], d->torsion_phase[u]);
        }
        for (int t = 0; t < h.n_nb14; t++) {
            tp[t] = tnew[d->nb14_types[2 * t]] * nt + tnew[d->nb14_types[2 * t + 1]];
            cf[t] = make_double2(d->nb14_coeff_vdw[t], d->nb14_coeff_ele[t]);
        }
        // slot -> atom, in slot order
        std::vector<int> slot_atom;
        slot_atom.reserve(n_contrib);
        for (int t = 0; t < h.n_angles; t++)
            for (int k = 0; k < 3; k++) slot_atom.push_back(d->angle_ids[3 * (size_t)t + k]);
        for (int t = 0; t < n_tor; t++)
            for (int k = 0; k < 4; k++) slot_atom.push_back(ids4[4 * (size_t)t + k]);
        for (int t = 0; t < h.n_nb14; t++)
            for (int k = 0; k < 2; k++) slot_atom.push_back(d->nb14_ids[2 * (size_t)t + k]);
        for (int c = 0; c < n_contrib; c++) off[slot_atom[c] + 1]++;
        for (int i = 0; i < n; i++) off[i + 1] += off[i];
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int c = 0; c < n_contrib; c++) at[fill[slot_atom[c]]++] = c;
        if (h.n_angles) {
            OG_HIP(hipMemcpy(h.ang_ids, d->angle_ids, sizeof(int) * 3 * h.n_angles, hipMemcpyHostToDevice));
            OG_HIP(hipMemcpy(h.ang_par, ap.data(), sizeof(double2) * ap.size(), hipMemcpyHostToDevice));
        }
        if (n_tor) {
            OG_HIP(hipMemcpy(h.tor_ids, ids4.data(), sizeof(int) * ids4.size(), hipMemcpyHostToDevice));
            OG_HIP(hipMemcpy(h.tor_par, tpar.data(), sizeof(double4) * tpar.size(), hipMemcpyHostToDevice));
        }
        if (h.n_nb14) {
            OG_HIP(hipMemcpy(h.nb14_ids, d->nb14_ids, sizeof(int) * 2 * h.n_nb14, hipMemcpyHostToDevice));
            OG_HIP(hipMemcpy(h.nb14_tp, tp.data(), sizeof(int) * tp.size(), hipMemcpyHostToDevice));
            OG_HIP(hipMemcpy(h.nb14_coef, cf.data(), sizeof(double2) * cf.size(), hipMemcpyHostToDevice));
        }
        OG_HIP(hipMemcpy(h.bc_off, off.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.bc_at, at.data(), sizeof(int) * at.size(), hipMemcpyHostToDevice));
    }
    if (h.n_quads) {
        OG_HIP(hipMemcpy(h.quad_ids, d->shake_quad_ids, sizeof(int) * 4 * h.n_quads, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.quad_d2, d->shake_quad_d2, sizeof(double) * 6 * h.n_quads, hipMemcpyHostToDevice));
    }
    if (h.n_trios) {
        OG_HIP(hipMemcpy(h.trio_ids, d->shake_trio_ids, sizeof(int) * 3 * h.n_trios, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.trio_d2, d->shake_trio_d2, sizeof(double) * 3 * h.n_trios, hipMemcpyHostToDevice));
    }
    if (h.n_pairs_c) {
        OG_HIP(hipMemcpy(h.pairc_ids, d->shake_pair_ids, sizeof(int) * 2 * h.n_pairs_c, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.pairc_d2, d->shake_pair_d2, sizeof(double) * h.n_pairs_c, hipMemcpyHostToDevice));
    }
    if (h.n_settle) OG_HIP(hipMemcpy(h.settle_ids, d->settle_ids, sizeof(int) * 3 * h.n_settle, hipMemcpyHostToDevice));
    // The fused Langevin step (one kernel per MD step after the pair kernel)
    // takes SETTLE waters one per thread, so each water must interact with no
    // other atom through exclusions and carry no bond, SHAKE or VcMD group;
    // the other atoms (at most OG_FREE_MAX) go through one block.
    h.n_free         = 0;
    h.n_water_blocks = 0;
    if (h.integrator == 1 && h.n_settle > 0 && !p->split_step) {
        std::vector<int> water(n, -1);
        bool             fuse = true;
        for (int t = 0; t < h.n_settle && fuse; t++)
            for (int k = 0; k < 3 && fuse; k++) {
                int &w = water[d->settle_ids[3 * t + k]];
                fuse   = w < 0;
                w      = t;
            }
        auto outside = [&](const int *ids, long long cnt) {
            for (long long k = 0; k < cnt; k++)
                if (water[ids[k]] >= 0) return false;
            return true;
        };
        for (int t = 0; t < h.n_excess && fuse; t++) fuse = water[d->excess[2 * t]] == water[d->excess[2 * t + 1]];
        fuse = fuse && outside(d->bond_ids, 2LL * h.n_bonds) && outside(d->shake_pair_ids, 2LL * h.n_pairs_c) &&
               outside(d->shake_trio_ids, 3LL * h.n_trios) && outside(d->vc_atoms_a, h.vc_na) &&
               outside(d->vc_atoms_b, h.vc_nb) && outside(d->angle_ids, 3LL * h.n_angles) &&
               outside(d->torsion_ids, 4LL * h.n_torsions) && outside(d->impro_ids, 4LL * h.n_impros) &&
               outside(d->nb14_ids, 2LL * h.n_nb14) && h.n_quads == 0;  // the step kernel runs no four-atom SHAKE
        std::vector<int> free_ids;
        for (int i = 0; i < n && fuse; i++)
            if (water[i] < 0) free_ids.push_back(i);
        const int nwb = cdiv(h.n_settle, OG_WATERS_PER_BLOCK);
        if (fuse && (int)free_ids.size() <= OG_FREE_MAX) {
            if (!A(&h.free_ids, std::max<size_t>(free_ids.size(), 1))) return OG_E_DEVICE;
            if (!free_ids.empty())
                OG_HIP(hipMemcpy(h.free_ids, free_ids.data(), sizeof(int) * free_ids.size(), hipMemcpyHostToDevice));
            h.n_free         = (int)free_ids.size();
            h.n_water_blocks = nwb;
        }
    }
    if (h.n_com_groups) {
        const int ng = h.n_com_groups;
        OG_HIP(hipMemcpy(h.com_off, d->com_group_off, sizeof(int) * (ng + 1), hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.com_atoms, d->com_group_atoms, sizeof(int) * d->com_group_off[ng], hipMemcpyHostToDevice));
        std::vector<float> gminv(ng);
        for (int gi = 0; gi < ng; gi++) {
            float sum = 0.0f;  // MmSystem.cpp:572-582 float accumulation
            for (int k = d->com_group_off[gi]; k < d->com_group_off[gi + 1]; k++) sum += m[d->com_group_atoms[k]];
            gminv[gi] = (float)(1.0 / (double)sum);
        }
        OG_HIP(hipMemcpy(h.com_minv, gminv.data(), sizeof(float) * ng, hipMemcpyHostToDevice));
    }
    if (groups_on) {
        // VcMD / V-AUS groups: atoms, unwrapped coordinates (the launch set's group coordinates when given,
        // else the atoms' own), float group masses (MmSystem.cpp:572-582)
        const int ng = h.vc_na + h.vc_nb;
        ok = A(&h.vc_atoms, ng) && A(&h.vc_crdg, 3 * (size_t)ng);
        if (!ok) return OG_E_DEVICE;
        std::vector<int>    ids(d->vc_atoms_a, d->vc_atoms_a + h.vc_na);
        ids.insert(ids.end(), d->vc_atoms_b, d->vc_atoms_b + h.vc_nb);
        std::vector<double> cg(3 * (size_t)ng);
        float               ma = 0.0f, mb = 0.0f;
        for (int k = 0; k < ng; k++) {
            for (int j = 0; j < 3; j++)
                cg[3 * k + j] = d->vc_crd_init ? d->vc_crd_init[3 * (size_t)k + j] : crd0[3 * (size_t)ids[k] + j];
            (k < h.vc_na ? ma : mb) += m[ids[k]];
        }
        h.vc_minv_a = (double)(float)(1.0 / (double)ma);
        h.vc_minv_b = (double)(float)(1.0 / (double)mb);
        OG_HIP(hipMemcpy(h.vc_atoms, ids.data(), sizeof(int) * ng, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.vc_crdg, cg.data(), sizeof(double) * cg.size(), hipMemcpyHostToDevice));
    }
    if (h.vc_n_states) {
        const int ns = h.vc_n_states;
        ok = A(&h.vc_lo, ns) && A(&h.vc_hi, ns) && A(&h.vc_qinv, ns) && A(&h.vc_hist, ns);
        if (!ok) return OG_E_DEVICE;
        // ExtendedVcMD::set_default_q_cano (Extend.cpp:548-566): unset weights take the smallest positive one
        double qmin = 1e10;
        for (int k = 0; k < ns; k++)
            if (d->vc_q_cano[k] > 0 && d->vc_q_cano[k] < qmin) qmin = d->vc_q_cano[k];
        if (qmin == 1e10) qmin = 1.0;
        std::vector<double> qinv(ns);
        for (int k = 0; k < ns; k++) qinv[k] = 1.0 / (d->vc_q_cano[k] > 0 ? d->vc_q_cano[k] : qmin);
        OG_HIP(hipMemcpy(h.vc_lo, d->vc_min, sizeof(double) * ns, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.vc_hi, d->vc_max, sizeof(double) * ns, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.vc_qinv, qinv.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
    }
    if (h.mc_n_states) {
        const int ns = h.mc_n_states;
        ok = A(&h.mc_order, ns) && A(&h.mc_poly, (size_t)ns * OG_MC_NPOLY) && A(&h.mc_lo, ns) && A(&h.mc_hi, ns) &&
             A(&h.mc_pdown, ns) && A(&h.mc_pup, ns);
        if (!ok) return OG_E_DEVICE;
        OG_HIP(hipMemcpy(h.mc_order, d->mc_order, sizeof(int) * ns, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.mc_poly, d->mc_poly, sizeof(double) * ns * OG_MC_NPOLY, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.mc_lo, d->mc_lo, sizeof(double) * ns, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.mc_hi, d->mc_hi, sizeof(double) * ns, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.mc_pdown, d->mc_prob_down, sizeof(double) * ns, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(h.mc_pup, d->mc_prob_up, sizeof(double) * ns, hipMemcpyHostToDevice));
    }
    if (h.vc_n_states || h.mc_n_states) {
        og_dev_counters c0{};
        c0.vs         = h.vc_n_states ? d->vc_init_state : d->mc_init_state;
        c0.step_end   = OG_NO_STEP_END;
        c0.noise_step = -1;
        c0.mc_scale   = 1.0;
        OG_HIP(hipMemcpy(s->d_cnt, &c0, sizeof c0, hipMemcpyHostToDevice));
    }
    OG_HIP(hipMemcpy(s->d_self, &h, sizeof h, hipMemcpyHostToDevice));
    return OG_OK;
}

// Autotuner: times the pair kernel of an execution context over its launch
// shapes (1, 2, 4 or 8 waves per block; for a batch also blocks in system
// order, round-robin across systems, or each system on one XCD), `trials`
// launches each after one
// warm-up, and keeps the fastest.  The forces are fixed-point integer sums,
// so every shape gives the same bits and the choice never changes a result.
// The trial launches add into the fixed-point gradient and energy slots,
// which are zero between steps; they are cleared again afterwards.
og_status exec_autotune(og_exec &x, int trials, int *shape_out, double *ms_out) {
    og_status st = exec_prepare(x);
    if (st != OG_OK) return st;
    if (x.members.empty()) return OG_E_STATE;
    trials = std::max(trials, 1);
    bool need = false;
    for (og_system *s : x.members)
        if (!s->grid_valid) {  // a fresh system rebuilds on its first step anyway
            st = request_rebuild(s);
            if (st != OG_OK) return st;
            need = true;
        }
    if (need) enqueue_rebuild(x);
    struct Events {  // destroyed on every return path
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) hipEventDestroy(a);
            if (b) hipEventDestroy(b);
        }
    } ev;
    OG_HIP(hipEventCreate(&ev.a));
    OG_HIP(hipEventCreate(&ev.b));
    hipEvent_t e0 = ev.a, e1 = ev.b;
    double    best     = 1e300;
    int       best_shp = x.pair_shape;
    // shapes: waves per block in each block order (for one system only orders 0 and 3 differ)
    for (int shp = 0; shp < 16; shp++) {
        if (x.members.size() == 1 && shp >= 4 && shp < 12) continue;
        const og_launch a = L(x, KIND_PAIR + shp);
        const int       w = 1 << (shp & 3);
        og_launch_pair(a, x.pair_form, w);
        OG_HIP(hipEventRecord(e0, x.stream));
        for (int t = 0; t < trials; t++) og_launch_pair(a, x.pair_form, w);
        OG_HIP(hipEventRecord(e1, x.stream));
        OG_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        OG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms / trials < best) best = ms / trials, best_shp = shp;
    }
    for (og_system *s : x.members) {
        OG_HIP(hipMemsetAsync(s->h.fx, 0, sizeof(long long) * OG_FX_WORDS(s->h.cap_atom_array), x.stream));
        OG_HIP(hipMemsetAsync(s->h.e_acc, 0, sizeof(long long) * OG_EACC_SLOTS * OG_EACC_STRIDE, x.stream));
    }
    OG_HIP(hipStreamSynchronize(x.stream));
    if (need)
        for (og_system *s : x.members) {
            st = request_rebuild(s, 0);
            if (st != OG_OK) return st;
        }
    for (og_system *s : x.members) s->grid_valid = true;
    if (best_shp != x.pair_shape) {
        free_graphs(x);
        x.pair_shape = best_shp;
    }
    if (shape_out) *shape_out = best_shp;
    if (ms_out) *ms_out = best;
    return check_flags(x);
}

og_status exec_set_pair_shape(og_exec &x, int waves, int order) {
    int lw = -1;
    for (int k = 0; k < 4; k++)
        if (waves == 1 << k) lw = k;
    if (lw < 0) {
        set_err("pair kernel waves per block must be 1, 2, 4 or 8");
        return OG_E_ARG;
    }
    if (order < 0 || order > 3) {
        set_err("pair kernel block order must be 0 (system order), 1 (round-robin), 2 (one XCD per system) or 3 "
                "(contiguous part per XCD)");
        return OG_E_ARG;
    }
    const int shp = lw + 4 * order;
    if (shp != x.pair_shape) {
        if (x.stream) OG_HIP(hipStreamSynchronize(x.stream));
        free_graphs(x);
        x.pair_shape = shp;
    }
    return OG_OK;
}

og_status ensure_ctx(og_system *s) {
    if (!s->ctx) {
        s->ctx = new og_exec();
        s->ctx->members.push_back(s);
    }
    return exec_prepare(*s->ctx);
}

}  // namespace

// ================================================================ C ABI ==
extern "C" {

void og_params_default(og_params *p) {
    std::memset(p, 0, sizeof *p);
    p->cutoff              = 12.0;
    p->nsgrid_cutoff       = 13.0;   // Config.h:76 (cutoff + 1)
    p->nsgrid_update_intvl = 1;      // Config.h:77
    p->ele_alpha           = 0.0;
    p->time_step           = 0.0005; // Config.h:31
    p->integrator          = OG_INTEGRATOR_LEAPFROG_PRESTO;
    p->constraint          = OG_CONSTRAINT_NONE;
    p->const_tolerance     = 0.000001;
    p->const_max_loops     = 1000;
    p->temperature         = 300.0;
    p->langevin_gamma      = 0.0;
    p->random_seed         = 0;
    p->com_cancel          = 0;
    p->gap_bug             = 1;
    p->excess_bug          = 1;
    p->zo_grad_bug         = 1;
    p->log_stride          = 1;
    p->log_cap             = 1024;
    p->capacity_factor     = 1.0;
    p->thermostat          = OG_THERMOSTAT_NONE;
    p->berendsen_tau       = 0.0;   // Config.h:23
    p->temperature_init    = -1.0;  // Config.h:22
    p->heating_steps       = 0;     // Config.h:24
    p->nonbond             = OG_NONBOND_LJ;       // Config.h:96
    p->hps_epsilon         = 0.0;                 // Config.h:97
    p->electrostatic       = OG_ELE_ZERO_DIPOLE;  // Config.h:32
    p->dh_dielectric       = 0.0;                 // Config.h:92-94
    p->dh_ionic_strength   = 0.0;
    p->dh_temperature      = 0.0;
    p->dh_grad_fix         = 0;
    p->cell_edge_scale     = 1.0;
}

const char *og_status_string(og_status s) {
    switch (s) {
        case OG_OK: return "ok";
        case OG_E_ARG: return "invalid argument";
        case OG_E_CAPACITY: return "device list capacity exceeded";
        case OG_E_NAN: return "non-finite coordinate or force";
        case OG_E_NOCONV: return "constraint iteration did not converge";
        case OG_E_DEVICE: return "HIP runtime error";
        case OG_E_STATE: return "call out of order";
    }
    return "unknown";
}

og_status og_last_error(char *buf, int len) {
    if (buf && len > 0) {
        std::strncpy(buf, g_last_error.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return OG_OK;
}

og_status og_set_device(int device_id) {
    OG_HIP(hipSetDevice(device_id));
    return OG_OK;
}

og_status og_thread_sync(void) {
    OG_HIP(hipDeviceSynchronize());
    return OG_OK;
}

og_status og_system_footprint(const og_system_desc *d, const og_params *p, uint64_t *bytes) {
    if (!d || !p || !bytes) return OG_E_ARG;
    GridGeom g;
    if (grid_geom(g, d->n_atoms, d->box, p->nsgrid_cutoff, p->cell_edge_scale) != 0) return OG_E_ARG;
    // the allocations of sys_alloc, term by term (the fused step's free-atom list counted at its bound)
    const double n = d->n_atoms, nt = (double)d->n_lj_types * d->n_lj_types;
    const double cols = g.n_columns, cap = std::ceil((n + cols * OG_NAC) / 32.0) * 32.0, cells = cap / OG_NAC;
    const double pairs = pair_capacity(g, d->n_atoms, (int)cells, p->capacity_factor);
    const double ex = d->n_excess, nb = d->n_bonds;
    const double trios = p->constraint ? d->n_shake_trios : 0, pc = p->constraint ? d->n_shake_pairs : 0;
    const double settle = p->constraint == OG_CONSTRAINT_SHAKE_SETTLE ? d->n_settle : 0;
    const double ng = p->com_cancel ? d->n_com_groups : 0;
    const double ng_atoms = ng > 0 ? d->com_group_off[d->n_com_groups] : 1;
    const double ke_parts = std::max(std::ceil(n / OG_THREADS), std::floor(settle / OG_WATERS_PER_BLOCK) + 2);
    const double vc = (d->vc_n_states || d->mc_mode == OG_EXT_VAUS) ? d->vc_n_a + d->vc_n_b : 0, vs = d->vc_n_states;
    const double mcs = d->mc_mode ? d->mc_n_states : 0;
    const double log_rows = std::max(1, p->log_cap);
    double b = 0.0;
    b += n * (6 * 24 + 3 * 4 + 4 + 4.0 * OG_MAX_NB15 + 4 + 4 + 4 + 4);   // state, masses, types, table, comp, col, zkey, rev
    b += nt * (4 + 4 + 8 + (p->electrostatic == OG_ELE_DEBYE_HUCKEL ? 16 : 0));  // lj6, lj12, lj_pair, lj_hps
    b += ex * 8 + nb * (8 + 16) + trios * (12 + 24) + pc * (8 + 8) + settle * 12;
    b += 2 * (n + 1) * 4 + std::max(2 * ex, 1.0) * 16 + std::max(2 * nb, 1.0) * 16;      // per-atom pair lists
    b += (ng + 1) * 4 + ng_atoms * 4 + (ng + 1) * 4 + 3 * (ng + 1) * 8;
    b += cols * (4 * 6 + 8) + 2 * (cols + 2) * g.n_uni_z * 4;
    b += cap * (4 + 4 + 24 + 16 + 4 + 4 + 32);                            // tmp_ids, atomids, gcrd, gxyzq, gtype, gcomp, fx
    b += cells * (6 * 4 + 4.0 * OG_ENUM_STAGE + 8) + (cells / 1024 + 1) * 8;
    b += pairs * 12;
    b += 8.0 * OG_EACC_SLOTS * OG_EACC_STRIDE + ke_parts * 8 + OG_E_NTERMS * 8 + log_rows * OG_E_NTERMS * 8;
    b += std::min(n, (double)OG_FREE_MAX) * 4 + vc * 28 + vs * 32 + mcs * (4 + 8.0 * OG_MC_NPOLY + 32);
    {
        const double na = d->n_angles, ntr = (double)d->n_torsions + d->n_impros, n14 = d->n_nb14;
        const double nq = p->constraint ? d->n_shake_quads : 0;
        const double contrib = 3 * na + 4 * ntr + 2 * n14;
        if (na + ntr + n14 > 0) b += na * (12 + 16) + ntr * (16 + 32) + n14 * (8 + 4 + 16) + contrib * (24 + 4) + (n + 1) * 4;
        b += nq * (16 + 48);
    }
    if (p->thermostat == OG_THERMOSTAT_SCALING) b += std::ceil(n / OG_THREADS) * 8;   // thermostat partials
    b += sizeof(og_dev_counters) + sizeof(og_dev_sys);
    *bytes = (uint64_t)(b * 1.02) + 65536;                                // allocation granularity
    return OG_OK;
}

og_status og_system_create(const og_system_desc *d, const og_params *p, og_system **out) {
    if (!d || !p || !out || d->n_atoms <= 0 || !d->crd || !d->box || !d->nb15off) {
        set_err("og_system_create: missing arguments");
        return OG_E_ARG;
    }
    auto s = std::make_unique<og_system>();
    s->prm = *p;
    if (s->prm.nsgrid_update_intvl < 1) s->prm.nsgrid_update_intvl = 1;
    if (p->cell_edge_scale != 1.0 && (p->gap_bug || !(p->cell_edge_scale >= 0.5 && p->cell_edge_scale <= 2.0))) {
        // the reference's list order follows its own geometry: only the full z-bin scan may take another
        set_err("cell_edge_scale other than 1 needs gap_bug 0 and lies in [0.5, 2]");
        return OG_E_ARG;
    }
    if (grid_geom(s->geom, d->n_atoms, d->box, p->nsgrid_cutoff, p->cell_edge_scale) != 0) {
        set_err("box too small for the pair-list radius (needs at least 2*n_neighbors+1 columns and 2 z bins)");
        return OG_E_ARG;
    }
    {
        // every index the kernels dereference must be in range
        const int n = d->n_atoms;
        auto bad_ids = [n](const int *v, long long cnt) {
            for (long long k = 0; k < cnt; k++)
                if (v[k] < 0 || v[k] >= n) return true;
            return false;
        };
        if (d->n_lj_types <= 0 || !d->atom_type || !d->lj6 || !d->lj12 || !d->charge || !d->mass || !d->vel) {
            set_err("og_system_create: incomplete topology");
            return OG_E_ARG;
        }
        for (int i = 0; i < n; i++)
            if (d->atom_type[i] < 0 || d->atom_type[i] >= d->n_lj_types || !(d->mass[i] > 0.0)) {
                set_err("atom %d: type out of range or non-positive mass", i);
                return OG_E_ARG;
            }
        for (long long k = 0; k < (long long)n * OG_MAX_NB15; k++)
            if (d->nb15off[k] < -1 || d->nb15off[k] >= n) {
                set_err("nb15off entry %lld out of range", k);
                return OG_E_ARG;
            }
        if ((d->n_excess && bad_ids(d->excess, 2LL * d->n_excess)) ||
            (d->n_bonds && bad_ids(d->bond_ids, 2LL * d->n_bonds)) ||
            (p->constraint && d->n_shake_pairs && bad_ids(d->shake_pair_ids, 2LL * d->n_shake_pairs)) ||
            (p->constraint && d->n_shake_trios && bad_ids(d->shake_trio_ids, 3LL * d->n_shake_trios)) ||
            (p->constraint == OG_CONSTRAINT_SHAKE_SETTLE && d->n_settle && bad_ids(d->settle_ids, 3LL * d->n_settle)) ||
            (p->constraint && d->n_shake_quads && bad_ids(d->shake_quad_ids, 4LL * d->n_shake_quads))) {
            set_err("an atom index in the excess/bond/constraint lists is out of range");
            return OG_E_ARG;
        }
        if (d->n_angles < 0 || d->n_torsions < 0 || d->n_impros < 0 || d->n_nb14 < 0 || d->n_shake_quads < 0 ||
            (d->n_angles && (!d->angle_ids || !d->angle_k || !d->angle_theta0)) ||
            (d->n_torsions && (!d->torsion_ids || !d->torsion_ene || !d->torsion_overlaps || !d->torsion_symmetry ||
                               !d->torsion_phase)) ||
            (d->n_impros && (!d->impro_ids || !d->impro_ene || !d->impro_overlaps || !d->impro_symmetry ||
                             !d->impro_phase)) ||
            (d->n_nb14 && (!d->nb14_ids || !d->nb14_types || !d->nb14_coeff_vdw || !d->nb14_coeff_ele)) ||
            (p->constraint && d->n_shake_quads && !d->shake_quad_d2)) {
            set_err("og_system_create: a bonded-term count without its tables");
            return OG_E_ARG;
        }
        if ((d->n_angles && bad_ids(d->angle_ids, 3LL * d->n_angles)) ||
            (d->n_torsions && bad_ids(d->torsion_ids, 4LL * d->n_torsions)) ||
            (d->n_impros && bad_ids(d->impro_ids, 4LL * d->n_impros)) ||
            (d->n_nb14 && bad_ids(d->nb14_ids, 2LL * d->n_nb14))) {
            set_err("an atom index in the angle/torsion/improper/1-4 lists is out of range");
            return OG_E_ARG;
        }
        // torsion multiplicity indexes the reference's 0..10 table (ForceField.cpp:7-17); a term with
        // V > 1e-10 divides by its overlap count
        for (int pass = 0; pass < 2; pass++) {
            const int     nt  = pass ? d->n_impros : d->n_torsions;
            const int    *sym = pass ? d->impro_symmetry : d->torsion_symmetry;
            const int    *ov  = pass ? d->impro_overlaps : d->torsion_overlaps;
            const double *ene = pass ? d->impro_ene : d->torsion_ene;
            for (int t = 0; t < nt; t++)
                if (sym[t] < 0 || sym[t] > 10 || (ene[t] > 1e-10 && ov[t] == 0)) {
                    set_err("%s %d: multiplicity outside 0..10 or zero overlap count", pass ? "improper" : "torsion", t);
                    return OG_E_ARG;
                }
        }
        for (int t = 0; t < d->n_nb14; t++)
            if (d->nb14_types[2 * t] < 0 || d->nb14_types[2 * t] >= d->n_lj_types || d->nb14_types[2 * t + 1] < 0 ||
                d->nb14_types[2 * t + 1] >= d->n_lj_types) {
                set_err("1-4 pair %d: LJ type out of range", t);
                return OG_E_ARG;
            }
        if (p->constraint)
            for (long long k = 0; k < 6LL * d->n_shake_quads; k++)
                if (!(d->shake_quad_d2[k] > 0.0)) {
                    set_err("four-atom SHAKE group %lld: non-positive target distance", k / 6);
                    return OG_E_ARG;
                }
        if (p->constraint == OG_CONSTRAINT_SHAKE_SETTLE && d->n_settle &&
            !(d->settle_dOH > 0.0 && d->settle_dHH > 0.0 && d->settle_dHH < 2.0 * d->settle_dOH)) {
            set_err("SETTLE needs 0 < dHH < 2 dOH");
            return OG_E_ARG;
        }
    }
    if (p->thermostat == OG_THERMOSTAT_SCALING) {
        // the reference switches to ThermostatScaling::apply_thermostat_with_shake under constraints
        // (DynamicsModePresto::apply_constraint, DynamicsMode.cpp:564-576), which this engine does not run
        if (p->integrator != OG_INTEGRATOR_LEAPFROG_PRESTO || p->constraint != OG_CONSTRAINT_NONE) {
            set_err("the scaling thermostat runs with leapfrog-presto and no constraints");
            return OG_E_ARG;
        }
        if (!(p->berendsen_tau > 0.0) || d->d_free <= 0) {
            set_err("the scaling thermostat needs --berendsen-tau > 0 and positive degrees of freedom");
            return OG_E_ARG;
        }
    } else if (p->thermostat != OG_THERMOSTAT_NONE) {
        set_err("unknown thermostat %d", p->thermostat);
        return OG_E_ARG;
    }
    if (p->electrostatic == OG_ELE_DEBYE_HUCKEL) {
        if (!(p->dh_dielectric > 0.0 && p->dh_ionic_strength > 0.0 && p->dh_temperature > 0.0)) {
            set_err("Debye-Hueckel needs positive --debye-huckel-dielectric, -ionic-strength and -temperature");
            return OG_E_ARG;
        }
    } else if (p->electrostatic < OG_ELE_ZERO_DIPOLE || p->electrostatic > OG_ELE_ZERO_HEXADECAPOLE) {
        set_err("unknown electrostatic %d", p->electrostatic);
        return OG_E_ARG;
    }
    if (p->nonbond == OG_NONBOND_HPS) {
        // the device pairs the HPS switch with Debye-Hueckel only, as the reference's GPU kernel
        // (kernel_pairwise_hps_dh, cuda_setup.cu:1334-1477)
        if (p->electrostatic != OG_ELE_DEBYE_HUCKEL) {
            set_err("hydrophobicity-scale-lj runs with --electrostatic debye-huckel");
            return OG_E_ARG;
        }
        if (!d->hps_cutoff || !d->hps_lambda) {
            set_err("hydrophobicity-scale-lj needs the HPS cutoff and lambda tables");
            return OG_E_ARG;
        }
    } else if (p->nonbond != OG_NONBOND_LJ) {
        set_err("unknown nonbond %d", p->nonbond);
        return OG_E_ARG;
    }
    if (p->com_cancel && d->n_com_groups > 0) {
        // groups must be disjoint for the parallel removal
        std::vector<char> seen(d->n_atoms, 0);
        for (int k = 0; k < d->com_group_off[d->n_com_groups]; k++) {
            const int a = d->com_group_atoms[k];
            if (a < 0 || a >= d->n_atoms || seen[a]) {
                set_err("COM-cancel groups must be disjoint atom sets");
                return OG_E_ARG;
            }
            seen[a] = 1;
        }
    }
    if (d->vc_n_states) {
        // two disjoint, non-empty groups; ranges and weights for every state
        const int ns = d->vc_n_states;
        if (ns < 0 || !d->vc_min || !d->vc_max || !d->vc_q_cano || d->vc_init_state < 0 || d->vc_init_state >= ns ||
            d->vc_n_a <= 0 || d->vc_n_b <= 0 || !d->vc_atoms_a || !d->vc_atoms_b || d->vc_interval <= 0) {
            set_err("VcMD: states, initial state, interval or groups missing or out of range");
            return OG_E_ARG;
        }
        std::vector<char> seen(d->n_atoms, 0);
        for (int k = 0; k < d->vc_n_a + d->vc_n_b; k++) {
            const int a = k < d->vc_n_a ? d->vc_atoms_a[k] : d->vc_atoms_b[k - d->vc_n_a];
            if (a < 0 || a >= d->n_atoms || seen[a]) {
                set_err("VcMD groups must be disjoint sets of valid atoms");
                return OG_E_ARG;
            }
            seen[a] = 1;
        }
        for (int k = 0; k < ns; k++)
            if (!(d->vc_min[k] < d->vc_max[k])) {
                set_err("VcMD state %d: empty lambda range", k);
                return OG_E_ARG;
            }
    }
    if (d->mc_mode) {
        // V-McMD / V-AUS: one of the two, not beside VcMD; states with orders, ranges and probabilities;
        // V-AUS needs two disjoint non-empty groups
        const int ns = d->mc_n_states;
        if ((d->mc_mode != OG_EXT_VMCMD && d->mc_mode != OG_EXT_VAUS) || d->vc_n_states || ns <= 0 ||
            !d->mc_order || !d->mc_poly || !d->mc_lo || !d->mc_hi || !d->mc_prob_down || !d->mc_prob_up ||
            d->mc_init_state < 0 || d->mc_init_state >= ns || d->mc_interval <= 0 || !(d->mc_temperature > 0.0)) {
            set_err("V-McMD/V-AUS: mode, states, initial state, interval or temperature missing or out of range");
            return OG_E_ARG;
        }
        for (int k = 0; k < ns; k++)
            if (d->mc_order[k] < 0 || d->mc_order[k] > OG_MC_MAX_ORDER || !(d->mc_lo[k] <= d->mc_hi[k])) {
                set_err("V-McMD/V-AUS state %d: order outside 0..%d or empty lambda range", k, OG_MC_MAX_ORDER);
                return OG_E_ARG;
            }
        if (d->mc_mode == OG_EXT_VAUS) {
            if (d->vc_n_a <= 0 || d->vc_n_b <= 0 || !d->vc_atoms_a || !d->vc_atoms_b) {
                set_err("V-AUS needs its two groups (vc_atoms_a, vc_atoms_b)");
                return OG_E_ARG;
            }
            std::vector<char> seen(d->n_atoms, 0);
            for (int k = 0; k < d->vc_n_a + d->vc_n_b; k++) {
                const int a = k < d->vc_n_a ? d->vc_atoms_a[k] : d->vc_atoms_b[k - d->vc_n_a];
                if (a < 0 || a >= d->n_atoms || seen[a]) {
                    set_err("V-AUS groups must be disjoint sets of valid atoms");
                    return OG_E_ARG;
                }
                seen[a] = 1;
            }
        }
    }
    s->d_free = d->d_free;
    og_status st = sys_alloc(s.get(), d, p);
    if (st != OG_OK) {
        for (void *ptr : s->allocs) hipFree(ptr);
        return st;
    }
    *out = s.release();
    return OG_OK;
}

og_status og_system_destroy(og_system *s) {
    if (!s) return OG_E_ARG;
    if (s->ctx) {
        hipStreamSynchronize(s->ctx->stream);
        exec_release(*s->ctx);
        delete s->ctx;
    }
    for (void *p : s->allocs) hipFree(p);
    delete s;
    return OG_OK;
}

og_status og_nsgrid_update(og_system *s) {
    if (!s) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    if (st != OG_OK) return st;
    st = request_rebuild(s);
    if (st != OG_OK) return st;
    enqueue_rebuild(*s->ctx);
    OG_HIP(hipStreamSynchronize(s->ctx->stream));
    st = request_rebuild(s, 0);
    if (st != OG_OK) return st;
    s->grid_valid = true;
    return check_flags(*s->ctx);
}

og_status og_calc_energy_force(og_system *s, int rebuild, double *work_atom, double *energies) {
    if (!s) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    if (st != OG_OK) return st;
    og_exec &x = *s->ctx;
    // a single-point evaluation always carries its energies
    const int one = 1, zero = 0;
    OG_HIP(hipMemcpy(&s->d_cnt->force_energy, &one, sizeof(int), hipMemcpyHostToDevice));
    if (rebuild || !s->grid_valid) {
        st = request_rebuild(s);  // cleared by the energy reduction
        if (st != OG_OK) return st;
        enqueue_rebuild(x);
    }
    enqueue_forces(x, nullptr, nullptr, true);
    OG_HIP(hipStreamSynchronize(x.stream));
    OG_HIP(hipMemcpy(&s->d_cnt->force_energy, &zero, sizeof(int), hipMemcpyHostToDevice));
    s->grid_valid = true;
    st            = check_flags(x);
    if (st != OG_OK) return st;
    if (work_atom) OG_HIP(hipMemcpy(work_atom, s->h.work, sizeof(double) * 3 * s->h.n_atoms, hipMemcpyDeviceToHost));
    if (energies) {
        double e[OG_E_NTERMS];
        OG_HIP(hipMemcpy(e, s->h.energy, sizeof e, hipMemcpyDeviceToHost));
        for (int k = 0; k < OG_NTERMS; k++) energies[k] = e[k];
    }
    return OG_OK;
}

og_status og_md_run(og_system *s, int n_steps) {
    if (!s || n_steps < 0) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    if (st != OG_OK) return st;
    return exec_run(*s->ctx, n_steps, s->time_pair);
}

og_status og_get_state(og_system *s, double *crd, double *vel_next, double *vel_just) {
    if (!s) return OG_E_ARG;
    const size_t n = 3 * (size_t)s->h.n_atoms;
    if (crd) OG_HIP(hipMemcpy(crd, s->h.crd, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (vel_next) OG_HIP(hipMemcpy(vel_next, s->h.vel_next, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (vel_just) {
        std::vector<double> a(n), b(n);
        OG_HIP(hipMemcpy(a.data(), s->h.vel, sizeof(double) * n, hipMemcpyDeviceToHost));
        OG_HIP(hipMemcpy(b.data(), s->h.vel_next, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (s->h.integrator == 1)
            for (size_t k = 0; k < n; k++) vel_just[k] = a[k];
        else
            for (size_t k = 0; k < n; k++) vel_just[k] = (a[k] + b[k]) * 0.5;
    }
    return OG_OK;
}

og_status og_get_crd_prev(og_system *s, double *crd_prev) {
    if (!s || !crd_prev) return OG_E_ARG;
    if (s->ctx && s->ctx->stream) OG_HIP(hipStreamSynchronize(s->ctx->stream));
    OG_HIP(hipMemcpy(crd_prev, s->h.crd_prev, sizeof(double) * 3 * (size_t)s->h.n_atoms, hipMemcpyDeviceToHost));
    return OG_OK;
}

og_status og_set_langevin_noise(og_system *s, const double *zeta) {
    if (!s) return OG_E_ARG;
    if (s->ctx && s->ctx->stream) OG_HIP(hipStreamSynchronize(s->ctx->stream));
    long long step = -1;
    if (zeta) {
        const size_t n = 3 * (size_t)s->h.n_atoms;
        if (!s->h.noise) {
            double *p = nullptr;
            OG_HIP(hipMalloc((void **)&p, sizeof(double) * n));
            s->allocs.push_back(p);
            s->dev_bytes += sizeof(double) * n;
            s->h.noise = p;
            OG_HIP(hipMemcpy(s->d_self, &s->h, sizeof s->h, hipMemcpyHostToDevice));
            if (s->ctx) s->ctx->dirty = true;  // its descriptor copy carries the pointer
        }
        OG_HIP(hipMemcpy(s->h.noise, zeta, sizeof(double) * n, hipMemcpyHostToDevice));
        OG_HIP(hipMemcpy(&step, &s->d_cnt->step, sizeof step, hipMemcpyDeviceToHost));
    }
    OG_HIP(hipMemcpy(&s->d_cnt->noise_step, &step, sizeof step, hipMemcpyHostToDevice));
    return OG_OK;
}

og_status og_get_work(og_system *s, double *work_atom) {
    if (!s || !work_atom) return OG_E_ARG;
    if (s->ctx && s->ctx->stream) OG_HIP(hipStreamSynchronize(s->ctx->stream));
    OG_HIP(hipMemcpy(work_atom, s->h.work, sizeof(double) * 3 * (size_t)s->h.n_atoms, hipMemcpyDeviceToHost));
    return OG_OK;
}

og_status og_set_state(og_system *s, const double *crd, const double *vel_next) {
    if (!s) return OG_E_ARG;
    const size_t n = 3 * (size_t)s->h.n_atoms;
    if (crd) {
        // into the box as at creation; the grid is rebuilt before the next step
        std::vector<double> c(crd, crd + n);
        og_status           st = wrap_inbox(s->geom, s->h.n_atoms, c.data());
        if (st != OG_OK) return st;
        if (s->ctx && s->ctx->stream) OG_HIP(hipStreamSynchronize(s->ctx->stream));
        OG_HIP(hipMemcpy(s->h.crd, c.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        s->grid_valid = false;
    }
    if (vel_next) OG_HIP(hipMemcpy(s->h.vel_next, vel_next, sizeof(double) * n, hipMemcpyHostToDevice));
    return OG_OK;
}

og_status og_get_energies(og_system *s, double *energies) {
    if (!s || !energies) return OG_E_ARG;
    double e[OG_E_NTERMS];
    OG_HIP(hipMemcpy(e, s->h.energy, sizeof e, hipMemcpyDeviceToHost));
    for (int k = 0; k < OG_NTERMS; k++) energies[k] = e[k];
    return OG_OK;
}

og_status og_get_energy_log(og_system *s, double *out, int max_rows, int *n_rows) {
    if (!s || !out) return OG_E_ARG;
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    const long long nlog = (c.step + s->h.log_stride - 1) / s->h.log_stride;
    const int       n    = (int)std::min<long long>(std::min<long long>(nlog, s->h.log_cap), max_rows);
    std::vector<double> all((size_t)s->h.log_cap * OG_E_NTERMS);
    OG_HIP(hipMemcpy(all.data(), s->h.elog, sizeof(double) * all.size(), hipMemcpyDeviceToHost));
    // oldest kept row first
    const long long first = nlog - n;
    for (int r = 0; r < n; r++) {
        const long long idx = (first + r) % s->h.log_cap;
        std::memcpy(out + (size_t)r * OG_NTERMS, all.data() + idx * OG_E_NTERMS, sizeof(double) * OG_NTERMS);
    }
    if (n_rows) *n_rows = n;
    return OG_OK;
}

og_status og_get_step(og_system *s, long long *step) {
    if (!s || !step) return OG_E_ARG;
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    *step = c.step;
    return OG_OK;
}

og_status og_system_status(og_system *s, int *status) {
    if (!s || !status) return OG_E_ARG;
    *status = (int)s->status;
    return OG_OK;
}

og_status og_get_vcmd(og_system *s, int *vstate, double *counts, int n_states) {
    if (!s) return OG_E_ARG;
    if (s->h.vc_n_states == 0) {
        set_err("og_get_vcmd: the system has no VcMD block");
        return OG_E_STATE;
    }
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    if (vstate) *vstate = c.vs;
    if (counts) {
        if (n_states < s->h.vc_n_states) return OG_E_ARG;
        OG_HIP(hipMemcpy(counts, s->h.vc_hist, sizeof(double) * s->h.vc_n_states, hipMemcpyDeviceToHost));
    }
    return OG_OK;
}

og_status og_get_grid_info(og_system *s, int *n_cells, int *n_columns, int *n_cell_pairs, int *n_atom_array) {
    if (!s) return OG_E_ARG;
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    if (n_cells) *n_cells = c.n_cells;
    if (n_columns) *n_columns = s->h.n_columns;
    if (n_cell_pairs) *n_cell_pairs = c.n_cell_pairs;
    if (n_atom_array) *n_atom_array = c.n_cells * OG_NAC;
    return OG_OK;
}

og_status og_get_grid_geometry(og_system *s, int *n_cells_xyz, float *L_cell_xyz, int *n_neighbors_xyz,
                               int *max_n_cells, int *max_n_cell_pairs) {
    if (!s || !n_cells_xyz || !L_cell_xyz || !n_neighbors_xyz) return OG_E_ARG;
    n_cells_xyz[0]     = s->h.nx;
    n_cells_xyz[1]     = s->h.ny;
    n_cells_xyz[2]     = 0;
    L_cell_xyz[0]      = s->h.L_cell[0];
    L_cell_xyz[1]      = s->h.L_cell[1];
    L_cell_xyz[2]      = 0.0f;
    n_neighbors_xyz[0] = s->h.nnx;
    n_neighbors_xyz[1] = s->h.nny;
    n_neighbors_xyz[2] = 0;
    if (max_n_cells) *max_n_cells = s->h.cap_cells;
    if (max_n_cell_pairs) *max_n_cell_pairs = s->h.cap_cell_pairs;
    return OG_OK;
}

og_status og_get_grid(og_system *s, int *atomids, double *grid_crd, int *idx_xy_head_cell) {
    if (!s) return OG_E_ARG;
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    const size_t na = (size_t)c.n_cells * OG_NAC;
    if (atomids) OG_HIP(hipMemcpy(atomids, s->h.atomids, sizeof(int) * na, hipMemcpyDeviceToHost));
    if (grid_crd) OG_HIP(hipMemcpy(grid_crd, s->h.gcrd, sizeof(double) * 3 * na, hipMemcpyDeviceToHost));
    if (idx_xy_head_cell)
        OG_HIP(hipMemcpy(idx_xy_head_cell, s->h.head_cell, sizeof(int) * (s->h.n_columns + 1), hipMemcpyDeviceToHost));
    return OG_OK;
}

og_status og_get_cell_pairs(og_system *s, int32_t *cp5, int *idx_head) {
    if (!s) return OG_E_ARG;
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    std::vector<int>      head(c.n_cells + 1);
    std::vector<uint32_t> c2(c.n_cell_pairs);
    std::vector<uint64_t> mk(c.n_cell_pairs);
    OG_HIP(hipMemcpy(head.data(), s->h.cp_head, sizeof(int) * (c.n_cells + 1), hipMemcpyDeviceToHost));
    if (c.n_cell_pairs) {
        OG_HIP(hipMemcpy(c2.data(), s->h.cp_c2, sizeof(int) * c.n_cell_pairs, hipMemcpyDeviceToHost));
        OG_HIP(hipMemcpy(mk.data(), s->h.cp_mask, sizeof(uint64_t) * c.n_cell_pairs, hipMemcpyDeviceToHost));
    }
    if (idx_head) std::memcpy(idx_head, head.data(), sizeof(int) * (c.n_cells + 1));
    if (cp5)
        for (int c1 = 0; c1 < c.n_cells; c1++)
            for (int p = head[c1]; p < head[c1 + 1]; p++) {
                cp5[5 * p]     = c1;
                cp5[5 * p + 1] = (int32_t)(c2[p] & 0xffffffu);
                cp5[5 * p + 2] = (int32_t)(c2[p] >> 24);
                cp5[5 * p + 3] = (int32_t)(uint32_t)(mk[p] & 0xffffffffu);
                cp5[5 * p + 4] = (int32_t)(uint32_t)(mk[p] >> 32);
            }
    return OG_OK;
}

og_status og_get_work_grid(og_system *s, double *work_grid) {
    if (!s || !work_grid) return OG_E_ARG;
    og_dev_counters c;
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    const int na = c.n_cells * OG_NAC;
    std::vector<int>    ids(na);
    std::vector<double> w(3 * (size_t)s->h.n_atoms);
    OG_HIP(hipMemcpy(ids.data(), s->h.atomids, sizeof(int) * na, hipMemcpyDeviceToHost));
    OG_HIP(hipMemcpy(w.data(), s->h.work, sizeof(double) * w.size(), hipMemcpyDeviceToHost));
    for (int g = 0; g < na; g++)
        for (int k = 0; k < 3; k++) work_grid[3 * g + k] = ids[g] >= 0 ? w[3 * (size_t)ids[g] + k] : 0.0;
    return OG_OK;
}

og_status og_pair_kernel_time(og_system *s, double *ms, int *n) {
    if (!s) return OG_E_ARG;
    if (ms) *ms = s->pair_ms;
    if (n) *n = s->pair_launches;
    return OG_OK;
}

og_status og_system_device_bytes(og_system *s, uint64_t *bytes) {
    if (!s || !bytes) return OG_E_ARG;
    *bytes = s->dev_bytes;
    return OG_OK;
}

og_status og_step_fused(og_system *s, int *fused) {
    if (!s || !fused) return OG_E_ARG;
    *fused = s->h.integrator == 1 && s->h.n_water_blocks > 0;
    return OG_OK;
}

// timing switch for the pair kernel events inside captured graphs
og_status og_set_pair_timing(og_system *s, int on) {
    if (!s) return OG_E_ARG;
    s->time_pair = on != 0;
    return OG_OK;
}

// ---- batches ----
og_status og_batch_create(og_batch **out) {
    if (!out) return OG_E_ARG;
    *out = new og_batch();
    return OG_OK;
}
og_status og_batch_add(og_batch *b, og_system *s) {
    if (!b || !s) return OG_E_ARG;
    if (std::find(b->ctx.members.begin(), b->ctx.members.end(), s) != b->ctx.members.end()) return OG_E_ARG;
    b->ctx.members.push_back(s);
    b->ctx.dirty = true;
    return OG_OK;
}
og_status og_batch_remove(og_batch *b, og_system *s) {
    if (!b || !s) return OG_E_ARG;
    auto it = std::find(b->ctx.members.begin(), b->ctx.members.end(), s);
    if (it == b->ctx.members.end()) return OG_E_ARG;
    if (b->ctx.stream) OG_HIP(hipStreamSynchronize(b->ctx.stream));
    b->ctx.members.erase(it);
    b->ctx.dirty = true;
    return OG_OK;
}
og_status og_batch_run(og_batch *b, int n_steps) {
    if (!b) return OG_E_ARG;
    bool timing = false;
    for (og_system *s : b->ctx.members) timing |= s->time_pair;
    return exec_run(b->ctx, n_steps, timing);
}
og_status og_autotune(og_system *s, int trials, int *waves, double *pair_ms) {
    if (!s) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    if (st != OG_OK) return st;
    int shp = OG_PAIR_DEFAULT;
    st      = exec_autotune(*s->ctx, trials, &shp, pair_ms);
    if (waves) *waves = 1 << (shp & 3);
    return st;
}
og_status og_set_pair_shape(og_system *s, int waves) {
    if (!s) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    return st != OG_OK ? st : exec_set_pair_shape(*s->ctx, waves, 0);
}
og_status og_set_pair_shape_order(og_system *s, int waves, int order) {
    if (!s) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    return st != OG_OK ? st : exec_set_pair_shape(*s->ctx, waves, order);
}
og_status og_get_pair_shape(og_system *s, int *waves, int *order) {
    if (!s || !waves || !order) return OG_E_ARG;
    og_status st = ensure_ctx(s);
    if (st != OG_OK) return st;
    *waves = 1 << (s->ctx->pair_shape & 3);
    *order = s->ctx->pair_shape >> 2;
    return OG_OK;
}
og_status og_autotune_cell_edge(const og_system_desc *d, const og_params *p, int trials, double *best_scale,
                                double *pair_ms) {
    if (!d || !p || !best_scale) return OG_E_ARG;
    *best_scale = 1.0;
    if (p->gap_bug) return OG_OK;  // the reference list keeps the reference geometry
    static const double cand[] = {0.8, 0.9, 1.0, 1.12, 1.25};
    double best = 1e300;
    for (double sc : cand) {
        GridGeom g;
        if (grid_geom(g, d->n_atoms, d->box, p->nsgrid_cutoff, sc) != 0) continue;
        og_params q        = *p;
        q.cell_edge_scale  = sc;
        q.log_cap          = 1;
        og_system *s       = nullptr;
        og_status  st      = og_system_create(d, &q, &s);
        if (st != OG_OK) return st;
        int    waves = 0;
        double ms    = 0.0;
        st           = og_calc_energy_force(s, 1, nullptr, nullptr);
        if (st == OG_OK) st = og_autotune(s, trials, &waves, &ms);
        og_system_destroy(s);
        if (st != OG_OK) return st;
        if (ms < best) best = ms, *best_scale = sc;
    }
    if (pair_ms) *pair_ms = best;
    return OG_OK;
}

og_status og_batch_autotune(og_batch *b, int trials, int *waves, int *order, double *pair_ms) {
    if (!b) return OG_E_ARG;
    int       shp = OG_PAIR_DEFAULT;
    og_status st  = exec_autotune(b->ctx, trials, &shp, pair_ms);
    if (waves) *waves = 1 << (shp & 3);
    if (order) *order = shp >> 2;
    return st;
}
og_status og_batch_set_pair_shape(og_batch *b, int waves, int order) {
    if (!b) return OG_E_ARG;
    return exec_set_pair_shape(b->ctx, waves, order);
}
og_status og_batch_set_graph_steps(og_batch *b, int steps) {
    if (!b || steps < 0) return OG_E_ARG;
    if (b->ctx.graph_steps != steps) {
        if (b->ctx.stream) OG_HIP(hipStreamSynchronize(b->ctx.stream));
        free_graphs(b->ctx);
        b->ctx.graph_steps = steps;
    }
    return OG_OK;
}
og_status og_set_segment(og_system *s, long long n_steps) {
    if (!s) return OG_E_ARG;
    og_dev_counters c;
    if (s->ctx && s->ctx->stream) OG_HIP(hipStreamSynchronize(s->ctx->stream));
    OG_HIP(hipDeviceSynchronize());
    OG_HIP(hipMemcpy(&c, s->d_cnt, sizeof c, hipMemcpyDeviceToHost));
    s->host_step = c.step;
    s->step_end  = n_steps > 0 ? c.step + n_steps : OG_NO_STEP_END;
    OG_HIP(hipMemcpy(&s->d_cnt->step_end, &s->step_end, sizeof(long long), hipMemcpyHostToDevice));
    return OG_OK;
}
og_status og_batch_size(og_batch *b, int *n) {
    if (!b || !n) return OG_E_ARG;
    *n = (int)b->ctx.members.size();
    return OG_OK;
}
og_status og_batch_destroy(og_batch *b) {
    if (!b) return OG_E_ARG;
    if (b->ctx.stream) hipStreamSynchronize(b->ctx.stream);
    exec_release(b->ctx);
    delete b;
    return OG_OK;
}

}  // extern "C"
