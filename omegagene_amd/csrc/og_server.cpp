// og_server.cpp -- continuous-batching service for simulation segments.
//
// The reference runs walkers as separate OS processes launched by shell
// scripts (samples/cg_q8_vcmd/1/c2_exe.bash:11-19; one omegagene process per
// walker and segment).  Here a single process keeps every walker resident
// on the GPU and serves segment requests:
//  * a request names a new walker (its system is copied at submission and
//    created on the device at admission) or a resident one (its state stays
//    on the device between segments);
//  * admission is FIFO against an HBM budget, using each system's footprint;
//  * every pump advances all active segments together by one chunk of steps
//    in one og_batch (one launch sequence for all walkers);
//  * a walker whose segment ends mid-chunk stands still for the rest of the
//    chunk (og_set_segment), so segments of any length share the batch;
//  * finished segments are retired with their final state and energy rows,
//    and new requests take their place at the next pump.
// Forces are 64-bit fixed point and the Langevin noise is keyed by each
// walker's seed and step, so a segment's result does not depend on which
// other walkers shared the batch.
#include "../../include/og_engine.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

void og_set_error(const char *msg);  // og_engine.cpp

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Deep copy of an og_system_desc: the caller's arrays may go away after submission.
struct DescCopy {
    og_system_desc       d{};
    std::vector<double>  box, crd, vel, charge, mass, lj6, lj12, bond_k, bond_r0, spd2, strd2, vc_min, vc_max, vc_q;
    std::vector<int>     type, nb15off, excess, bond_ids, spids, strids, settle, com_off, com_atoms, vc_a, vc_b;

    template <class T> static const T *cp(std::vector<T> &dst, const T *src, size_t n) {
        if (!src || n == 0) return nullptr;
        dst.assign(src, src + n);
        return dst.data();
    }
    explicit DescCopy(const og_system_desc &s) {
        d                 = s;
        const size_t n    = (size_t)s.n_atoms;
        const size_t nt2  = (size_t)s.n_lj_types * s.n_lj_types;
        d.box             = cp(box, s.box, 12);
        d.crd             = cp(crd, s.crd, 3 * n);
        d.vel             = cp(vel, s.vel, 3 * n);
        d.charge          = cp(charge, s.charge, n);
        d.mass            = cp(mass, s.mass, n);
        d.atom_type       = cp(type, s.atom_type, n);
        d.lj6             = cp(lj6, s.lj6, nt2);
        d.lj12            = cp(lj12, s.lj12, nt2);
        d.nb15off         = cp(nb15off, s.nb15off, 32 * n);
        d.excess          = cp(excess, s.excess, 2 * (size_t)s.n_excess);
        d.bond_ids        = cp(bond_ids, s.bond_ids, 2 * (size_t)s.n_bonds);
        d.bond_k          = cp(bond_k, s.bond_k, (size_t)s.n_bonds);
        d.bond_r0         = cp(bond_r0, s.bond_r0, (size_t)s.n_bonds);
        d.shake_pair_ids  = cp(spids, s.shake_pair_ids, 2 * (size_t)s.n_shake_pairs);
        d.shake_pair_d2   = cp(spd2, s.shake_pair_d2, (size_t)s.n_shake_pairs);
        d.shake_trio_ids  = cp(strids, s.shake_trio_ids, 3 * (size_t)s.n_shake_trios);
        d.shake_trio_d2   = cp(strd2, s.shake_trio_d2, 3 * (size_t)s.n_shake_trios);
        d.settle_ids      = cp(settle, s.settle_ids, 3 * (size_t)s.n_settle);
        const size_t ng   = s.n_com_groups > 0 && s.com_group_off ? (size_t)s.com_group_off[s.n_com_groups] : 0;
        d.com_group_off   = cp(com_off, s.com_group_off, s.n_com_groups > 0 ? (size_t)s.n_com_groups + 1 : 0);
        d.com_group_atoms = cp(com_atoms, s.com_group_atoms, ng);
        d.vc_min          = cp(vc_min, s.vc_min, (size_t)s.vc_n_states);
        d.vc_max          = cp(vc_max, s.vc_max, (size_t)s.vc_n_states);
        d.vc_q_cano       = cp(vc_q, s.vc_q_cano, (size_t)s.vc_n_states);
        d.vc_atoms_a      = cp(vc_a, s.vc_atoms_a, (size_t)(s.vc_n_states ? s.vc_n_a : 0));
        d.vc_atoms_b      = cp(vc_b, s.vc_atoms_b, (size_t)(s.vc_n_states ? s.vc_n_b : 0));
    }
};

struct Walker {
    long long                 id = 0;
    std::unique_ptr<DescCopy> desc;      // until the walker is created on the device
    og_params                 prm{};
    og_system                *sys = nullptr;
    uint64_t                  bytes = 0;
    int                       n_atoms = 0;
    bool                      busy = false;  // a segment is queued or running
    bool                      failed = false;  // its last segment faulted; the walker left the device
};

struct Request {
    long long id = 0, walker = 0;
    int       n_steps = 0;
    long long end = 0;   // walker step at which the segment is complete
    double    t_submit = 0, t_admit = 0, t_done = 0;
    bool      done = false;
    og_status status = OG_OK;
    std::vector<double> crd, vel, rows;
    int       n_rows = 0;
};

}  // namespace

struct og_server {
    uint64_t                       budget = 0, used = 0;
    int                            chunk = 10;
    og_batch                      *batch = nullptr;
    std::map<long long, Walker>    walkers;
    std::map<long long, Request>   requests;
    std::deque<long long>          queue;    // request ids waiting for admission
    std::vector<long long>         active;   // request ids on the device
    long long                      next_walker = 1, next_request = 1;
    double                         t0 = 0;
    long long                      steps_done = 0;
};

namespace {

og_status fail(og_status st, const std::string &msg) {
    og_set_error(msg.c_str());
    return st;
}

// retire a request whose segment has completed: copy its result to the host
og_status retire(og_server *srv, Request &r) {
    Walker &w = srv->walkers.at(r.walker);
    r.crd.resize(3 * (size_t)w.n_atoms);
    r.vel.resize(3 * (size_t)w.n_atoms);
    og_status st = og_get_state(w.sys, r.crd.data(), r.vel.data(), nullptr);
    if (st != OG_OK) return st;
    r.rows.resize((size_t)r.n_steps * OG_NTERMS);
    st = og_get_energy_log(w.sys, r.rows.data(), r.n_steps, &r.n_rows);
    if (st != OG_OK) return st;
    r.rows.resize((size_t)r.n_rows * OG_NTERMS);
    r.t_done = now_s() - srv->t0;
    r.done   = true;
    w.busy   = false;
    srv->steps_done += r.n_steps;
    return OG_OK;
}

}  // namespace

extern "C" {

og_status og_server_create(uint64_t hbm_budget_bytes, int chunk_steps, og_server **out) {
    if (!out || hbm_budget_bytes == 0 || chunk_steps <= 0) return OG_E_ARG;
    auto srv    = std::make_unique<og_server>();
    srv->budget = hbm_budget_bytes;
    srv->chunk  = chunk_steps;
    srv->t0     = now_s();
    og_status st = og_batch_create(&srv->batch);
    if (st != OG_OK) return st;
    st = og_batch_set_graph_steps(srv->batch, chunk_steps);
    if (st != OG_OK) return st;
    *out = srv.release();
    return OG_OK;
}

og_status og_server_destroy(og_server *srv) {
    if (!srv) return OG_E_ARG;
    og_batch_destroy(srv->batch);
    for (auto &kv : srv->walkers)
        if (kv.second.sys) og_system_destroy(kv.second.sys);
    delete srv;
    return OG_OK;
}

og_status og_server_submit_new(og_server *srv, const og_system_desc *d, const og_params *p, int n_steps,
                               long long *request_id, long long *walker_id) {
    if (!srv || !d || !p || n_steps <= 0 || !request_id || !walker_id || d->n_atoms <= 0 || !d->nb15off) return OG_E_ARG;
    uint64_t bytes = 0;
    og_status st   = og_system_footprint(d, p, &bytes);
    if (st != OG_OK) return st;
    if (bytes > srv->budget) return fail(OG_E_CAPACITY, "og_server_submit_new: the system alone exceeds the HBM budget");
    Walker w;
    w.id      = srv->next_walker++;
    w.desc    = std::make_unique<DescCopy>(*d);
    w.prm     = *p;
    w.prm.log_stride = 1;
    w.prm.log_cap    = std::max(p->log_cap, n_steps);
    w.bytes   = bytes;
    w.n_atoms = d->n_atoms;
    w.busy    = true;
    Request r;
    r.id       = srv->next_request++;
    r.walker   = w.id;
    r.n_steps  = n_steps;
    r.t_submit = now_s() - srv->t0;
    *walker_id  = w.id;
    *request_id = r.id;
    srv->walkers.emplace(w.id, std::move(w));
    srv->queue.push_back(r.id);
    srv->requests.emplace(r.id, std::move(r));
    return OG_OK;
}

og_status og_server_submit_next(og_server *srv, long long walker_id, int n_steps, long long *request_id) {
    if (!srv || n_steps <= 0 || !request_id) return OG_E_ARG;
    auto it = srv->walkers.find(walker_id);
    if (it == srv->walkers.end()) return fail(OG_E_ARG, "og_server_submit_next: unknown walker");
    Walker &w = it->second;
    if (w.busy) return fail(OG_E_STATE, "og_server_submit_next: the walker has a segment in flight");
    if (w.failed) return fail(OG_E_STATE, "og_server_submit_next: the walker's last segment faulted");
    if (w.sys && n_steps > w.prm.log_cap)
        return fail(OG_E_ARG, "og_server_submit_next: segment longer than the walker's energy log");
    w.busy = true;
    Request r;
    r.id       = srv->next_request++;
    r.walker   = walker_id;
    r.n_steps  = n_steps;
    r.t_submit = now_s() - srv->t0;
    *request_id = r.id;
    srv->queue.push_back(r.id);
    srv->requests.emplace(r.id, std::move(r));
    return OG_OK;
}

og_status og_server_pump(og_server *srv, int *n_active, int *n_queued, int *n_done) {
    if (!srv) return OG_E_ARG;
    // admission, FIFO: stop at the first request whose new walker does not fit
    while (!srv->queue.empty()) {
        Request &r = srv->requests.at(srv->queue.front());
        Walker  &w = srv->walkers.at(r.walker);
        if (!w.sys) {
            if (srv->used + w.bytes > srv->budget) break;
            og_status st = og_system_create(&w.desc->d, &w.prm, &w.sys);
            if (st != OG_OK) {  // this request fails alone
                r.status = st;
                r.done   = true;
                r.t_done = now_s() - srv->t0;
                w.busy   = false;
                srv->queue.pop_front();
                continue;
            }
            w.desc.reset();
            srv->used += w.bytes;
            st = og_batch_add(srv->batch, w.sys);
            if (st != OG_OK) return st;
        }
        og_status st = og_set_segment(w.sys, r.n_steps);
        if (st != OG_OK) return st;
        long long step = 0;
        st = og_get_step(w.sys, &step);
        if (st != OG_OK) return st;
        r.end     = step + r.n_steps;
        r.t_admit = now_s() - srv->t0;
        srv->active.push_back(r.id);
        srv->queue.pop_front();
    }
    int done = 0;
    if (!srv->active.empty()) {
        og_status st = og_batch_run(srv->batch, srv->chunk);
        if (st != OG_OK && st != OG_E_CAPACITY && st != OG_E_NAN && st != OG_E_NOCONV) {
            // a runtime error (or a batch that could not run) leaves no per-walker outcome: every active
            // segment fails
            for (long long rid : srv->active) {
                Request &r = srv->requests.at(rid);
                r.status   = st;
                r.done     = true;
                r.t_done   = now_s() - srv->t0;
                srv->walkers.at(r.walker).busy = false;
            }
            srv->active.clear();
            return st;
        }
        // a walker whose list overflowed or whose forces went non-finite stood still on the device: its
        // segment fails and it leaves the batch; the other walkers went on and are unaffected
        std::vector<long long> still;
        for (long long rid : srv->active) {
            Request &r = srv->requests.at(rid);
            Walker  &w = srv->walkers.at(r.walker);
            int       ws = OG_OK;
            st           = og_system_status(w.sys, &ws);
            if (st != OG_OK) return st;
            if (ws != OG_OK) {
                r.status = (og_status)ws;
                r.done   = true;
                r.t_done = now_s() - srv->t0;
                w.busy   = false;
                w.failed = true;
                st       = og_batch_remove(srv->batch, w.sys);
                if (st != OG_OK) return st;
                og_system_destroy(w.sys);
                w.sys = nullptr;
                srv->used -= w.bytes;
                continue;
            }
            long long step = 0;
            st             = og_get_step(w.sys, &step);
            if (st != OG_OK) return st;
            if (step >= r.end) {
                st = retire(srv, r);
                if (st != OG_OK) return st;
                done++;
            } else {
                still.push_back(rid);
            }
        }
        srv->active.swap(still);
    }
    if (n_active) *n_active = (int)srv->active.size();
    if (n_queued) *n_queued = (int)srv->queue.size();
    if (n_done) *n_done = done;
    return OG_OK;
}

og_status og_server_result(og_server *srv, long long request_id, double *crd, double *vel, double *energy_rows,
                           int max_rows, int *n_rows, double *times) {
    if (!srv) return OG_E_ARG;
    auto it = srv->requests.find(request_id);
    if (it == srv->requests.end()) return fail(OG_E_ARG, "og_server_result: unknown request");
    Request &r = it->second;
    if (!r.done) return fail(OG_E_STATE, "og_server_result: the request has not finished");
    if (times) {
        times[0] = r.t_submit;
        times[1] = r.t_admit;
        times[2] = r.t_done;
    }
    if (r.status != OG_OK) return r.status;
    if (crd) std::memcpy(crd, r.crd.data(), sizeof(double) * r.crd.size());
    if (vel) std::memcpy(vel, r.vel.data(), sizeof(double) * r.vel.size());
    const int n = std::min(max_rows, r.n_rows);
    if (energy_rows && n > 0) std::memcpy(energy_rows, r.rows.data(), sizeof(double) * (size_t)n * OG_NTERMS);
    if (n_rows) *n_rows = n;
    return OG_OK;
}

og_status og_server_forget(og_server *srv, long long request_id) {
    if (!srv) return OG_E_ARG;
    auto it = srv->requests.find(request_id);
    if (it == srv->requests.end() || !it->second.done) return OG_E_ARG;
    srv->requests.erase(it);
    return OG_OK;
}

og_status og_server_release_walker(og_server *srv, long long walker_id) {
    if (!srv) return OG_E_ARG;
    auto it = srv->walkers.find(walker_id);
    if (it == srv->walkers.end()) return OG_E_ARG;
    if (it->second.busy) return fail(OG_E_STATE, "og_server_release_walker: a segment is in flight");
    if (it->second.sys) {
        og_status st = og_batch_remove(srv->batch, it->second.sys);
        if (st != OG_OK) return st;
        og_system_destroy(it->second.sys);
        srv->used -= it->second.bytes;
    }
    srv->walkers.erase(it);
    return OG_OK;
}

og_status og_server_stats(og_server *srv, double *t_now, long long *steps_done, uint64_t *bytes_used, int *n_walkers) {
    if (!srv) return OG_E_ARG;
    if (t_now) *t_now = now_s() - srv->t0;
    if (steps_done) *steps_done = srv->steps_done;
    if (bytes_used) *bytes_used = srv->used;
    if (n_walkers) *n_walkers = (int)srv->walkers.size();
    return OG_OK;
}

}  // extern "C"
