// bh_serve.cpp -- the serving loop: requests of different grid sizes, each a
// self-consistent-field sequence of dependent solves, admitted by device
// memory budget and run with continuous batching.
//
// Requests run in lanes by grid size (grid points below 2^20, below 2^22,
// and larger), each lane with its own non-blocking HIP stream.  A lane's
// iteration takes the next solve of every active request of that lane (its
// density generated on the device, then a cold first solve or a warm one
// from the potential the request's plan kept) and runs them together as
// grouped launches (bh_exec.h); the lanes' iterations run side by side on
// the GPU, so a small grid's solve does not wait for the large grids of an
// iteration it would otherwise share.  One worker thread admits requests,
// keeps up to two iterations in flight per lane and polls their completion.
// Requests that arrive while iterations are on the GPU are admitted as soon
// as a slot and device memory are free, so new work joins work already
// running.  Plans (with their density and potential buffers) of finished
// requests are kept per grid size and reused; idle ones are released when
// memory is needed.
#include "bh_density.h"
#include "bh_exec.h"
#include "bh_plan_impl.h"
#include "bh_mixed.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>

using namespace bh;

// ---------------------------------------------------------------------------
// bipolytrope profile (host)
// ---------------------------------------------------------------------------
namespace bh {
namespace {

// One RK4 step of the Lane-Emden equation theta'' = -theta^n - 2 theta'/xi
void le_step(double n, double xi, double h, double *th, double *dth)
{
    auto f = [n](double x, double t, double dt, double *a, double *b) {
        *a = dt;
        *b = -(t > 0.0 ? std::pow(t, n) : 0.0) - 2.0 * dt / x;
    };
    double k1a, k1b, k2a, k2b, k3a, k3b, k4a, k4b;
    f(xi, *th, *dth, &k1a, &k1b);
    f(xi + 0.5 * h, *th + 0.5 * h * k1a, *dth + 0.5 * h * k1b, &k2a, &k2b);
    f(xi + 0.5 * h, *th + 0.5 * h * k2a, *dth + 0.5 * h * k2b, &k3a, &k3b);
    f(xi + h, *th + h * k3a, *dth + h * k3b, &k4a, &k4b);
    *th += h / 6.0 * (k1a + 2.0 * k2a + 2.0 * k3a + k4a);
    *dth += h / 6.0 * (k1b + 2.0 * k2b + 2.0 * k3b + k4b);
}

// Integrates the composite star outwards in the core's Lane-Emden units
// (4 pi G = 1, rho_c = 1, K_core = 1 / (n_core + 1), so r = xi).  The core
// (index nc) runs to xi_i, the envelope (index ne, K_env from pressure
// continuity at continuous density) to its surface, which is returned.
// When rec is given, (r, rho) samples are appended to it.
double composite_surface(double nc, double ne, double xi_i, double h, std::vector<std::pair<double, double>> *rec)
{
    // series start at small xi: theta = 1 - xi^2/6 + n xi^4/120
    double xi = 1e-6, th = 1.0 - xi * xi / 6.0, dth = -xi / 3.0;
    if (rec) rec->push_back({0.0, 1.0});
    double prev_xi = xi, prev_th = th;
    while (xi < xi_i) {
        const double step = std::min(h, xi_i - xi);
        prev_xi = xi;
        prev_th = th;
        le_step(nc, xi, step, &th, &dth);
        xi += step;
        if (th <= 0.0) {  // the core reached its own surface first
            const double xs = prev_xi + (xi - prev_xi) * prev_th / (prev_th - th);
            if (rec) rec->push_back({xs, 0.0});
            return xs;
        }
        if (rec) rec->push_back({xi, std::pow(th, nc)});
    }
    // envelope: w = rho^(1/ne), w' = -G m / (K_e (ne+1) r^2), m' = 4 pi r^2 w^ne
    const double pi = std::acos(-1.0), G = 1.0 / (4.0 * pi);
    const double rho_i = std::pow(th, nc);
    const double Kc = 1.0 / (nc + 1.0);
    const double Ke = Kc * std::pow(rho_i, 1.0 + 1.0 / nc) / std::pow(rho_i, 1.0 + 1.0 / ne);
    double r = xi, w = std::pow(rho_i, 1.0 / ne), m = -4.0 * pi * xi * xi * dth;
    const double c = G / (Ke * (ne + 1.0));
    auto fw = [&](double rr, double mm) { return -c * mm / (rr * rr); };
    auto fm = [&](double rr, double ww) { return 4.0 * pi * rr * rr * (ww > 0.0 ? std::pow(ww, ne) : 0.0); };
    for (int it = 0; it < 100000000; ++it) {
        const double k1w = fw(r, m), k1m = fm(r, w);
        const double k2w = fw(r + 0.5 * h, m + 0.5 * h * k1m), k2m = fm(r + 0.5 * h, w + 0.5 * h * k1w);
        const double k3w = fw(r + 0.5 * h, m + 0.5 * h * k2m), k3m = fm(r + 0.5 * h, w + 0.5 * h * k2w);
        const double k4w = fw(r + h, m + h * k3m), k4m = fm(r + h, w + h * k3w);
        const double wn = w + h / 6.0 * (k1w + 2 * k2w + 2 * k3w + k4w);
        const double mn = m + h / 6.0 * (k1m + 2 * k2m + 2 * k3m + k4m);
        if (wn <= 0.0) {
            const double rs = r + h * w / (w - wn);
            if (rec) rec->push_back({rs, 0.0});
            return rs;
        }
        r += h;
        w = wn;
        m = mn;
        if (rec) rec->push_back({r, std::pow(w, ne)});
    }
    return r;
}

}  // namespace

int bipolytrope_profile(double n_core, double n_env, double core_frac, int ntab, std::vector<double> *table,
                        double info[2])
{
    if (!(n_core >= 0.0 && n_core < 5.0) || !(n_env >= 0.0 && n_env < 5.0) || ntab < 2 || !(core_frac >= 0.0) ||
        !(core_frac <= 1.0))
        return fail(BH_E_INVALID, "bipolytrope: need 0 <= n < 5, 0 <= core_frac <= 1, ntab >= 2");
    const double h = 2e-4;
    double xi_i, R;
    std::vector<std::pair<double, double>> rec;
    if (core_frac <= 0.0) {  // the envelope polytrope alone
        R = composite_surface(n_env, n_env, 1e9, h, &rec);
        xi_i = 0.0;
    } else {
        const double xi1 = composite_surface(n_core, n_core, 1e9, h, nullptr);  // the core polytrope's surface
        if (core_frac >= 1.0) {
            xi_i = R = xi1;
        } else {
            // interface radius with xi_i / R = core_frac, by bisection
            double lo = 1e-3, hi = xi1 * (1.0 - 1e-9);
            for (int it = 0; it < 48; ++it) {
                const double mid = 0.5 * (lo + hi);
                const double ratio = mid / composite_surface(n_core, n_env, mid, 5 * h, nullptr);
                if (ratio < core_frac) lo = mid;
                else hi = mid;
            }
            xi_i = 0.5 * (lo + hi);
        }
        R = composite_surface(n_core, n_env, xi_i, h, &rec);
    }
    table->assign(ntab, 0.0);
    size_t j = 0;
    for (int k = 0; k < ntab; ++k) {
        const double r = R * (double)k / (double)(ntab - 1);
        while (j + 1 < rec.size() && rec[j + 1].first < r) ++j;
        double v;
        if (j + 1 >= rec.size()) v = 0.0;
        else {
            const double r0 = rec[j].first, r1 = rec[j + 1].first;
            const double t = (r1 > r0) ? (r - r0) / (r1 - r0) : 0.0;
            v = rec[j].second + std::min(1.0, std::max(0.0, t)) * (rec[j + 1].second - rec[j].second);
        }
        (*table)[k] = v;
    }
    (*table)[ntab - 1] = 0.0;
    info[0] = R;
    info[1] = xi_i;
    return BH_OK;
}

}  // namespace bh

// ---------------------------------------------------------------------------
// server
// ---------------------------------------------------------------------------
namespace {

constexpr int NPROF = 2049;

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void fill_model_defaults(bh_binary_model *m)
{
    if (m->rho_c[0] == 0.0) m->rho_c[0] = 1.0;
    if (m->rho_c[1] == 0.0) m->rho_c[1] = 1.0;
    if (m->n_core == 0.0) m->n_core = 3.0;
    if (m->n_env == 0.0) m->n_env = 1.5;
    if (m->core_frac == 0.0) m->core_frac = 0.3;
}

// Density arguments of solve i of a request (separation drifting by
// `drift` per solve; centres about the centre of mass, masses ~ rho_c R^3).
DensityArgs model_args(const bh_binary_model &m, const Geo &g, int i, double *rho, const double *prof)
{
    DensityArgs a{};
    a.rho = rho;
    a.prof = prof;
    a.nprof = NPROF;
    a.numr = g.numr;
    a.numz = g.numz;
    a.numphi = g.numphi;
    a.dr = g.dr;
    a.dphi = g.dphi;
    a.isym = g.isym;
    const double sep = m.separation * (1.0 + m.drift * (double)i);
    const double m1 = m.rho_c[0] * m.radius[0] * m.radius[0] * m.radius[0];
    const double m2 = m.rho_c[1] * m.radius[1] * m.radius[1] * m.radius[1];
    a.cx[0] = sep * m2 / (m1 + m2) * g.dr;
    a.cx[1] = -sep * m1 / (m1 + m2) * g.dr;
    for (int s = 0; s < 2; ++s) {
        a.rho_c[s] = m.rho_c[s];
        a.R[s] = m.radius[s] * g.dr;
    }
    a.noise = m.noise;
    a.seed = m.seed;
    return a;
}

struct Shell {  // a plan with its density and potential buffers
    bh_plan *plan = nullptr;
    DevMem mem;
    double *rho = nullptr, *pot = nullptr;
    std::string key;
    size_t bytes = 0;  // own bytes (plan's own + rho + pot)
    ~Shell()
    {
        if (plan) bh_plan_destroy(plan);
    }
};

struct Req {
    long long id;
    bh_request r;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> lat_ms;
    int status = BH_OK;
    std::vector<double> start;  // the potential to resume from (copied at submission), or empty
};

struct Slot {
    Shell *sh = nullptr;
    Req *req = nullptr;
    int lane = 0;
    int enq = 0, done = 0;  // solves enqueued / completed
    bool resume = false;    // the first solve is warm (from the request's start potential)
    double t_ready = 0;     // completion time of the last completed solve
    std::shared_ptr<const DevMem> prof;
    const double *prof_ptr = nullptr;
};

struct Iteration {
    hipEvent_t ev;
    std::vector<std::pair<Slot *, int>> solves;  // (slot, solve index)
    size_t rec = 0;                              // its entry in bh_server::iter_log
};

// Host-side record of one iteration (seconds since the server started):
// the admission that preceded it, the enqueue, and the completion.
struct IterRec {
    double t_admit, t_enq0, t_enq1, t_done;
    int nsolves, new_shells;
    double points;  // grid points solved in the iteration
    int lane;
};

// lanes by grid points: [0, 2^20), [2^20, 2^22), [2^22, ...)
constexpr int NLANES = 3;
int lane_of(const bh_grid_desc &d)
{
    const double pts = (double)d.numr * d.numz * d.numphi;
    return pts < 1048576.0 ? 0 : pts < 4194304.0 ? 1 : 2;
}

// One lane: its stream, its iterations in flight (oldest first) and the
// grouped launches of the next one.
struct Lane {
    hipStream_t stream = nullptr;
    std::deque<Iteration> inflight;
    GroupExec ex;
    std::vector<std::vector<Op>> lists;
};

std::string grid_key(const bh_grid_desc &d)
{
    char b[200];
    std::snprintf(b, sizeof b, "%d|%d|%d|%.17g|%d|%d|%d|%d|%d", d.numr, d.numz, d.numphi, d.dr, d.mmax, d.nsteps_cold,
                  d.nsteps_warm, d.refine_passes, d.isym);
    return b;
}

}  // namespace

struct bh_server {
    bh_server_config cfg{};
    int device = 0;
    Lane lanes[NLANES];
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    bool stop = false;
    long long next_id = 1;
    std::deque<Req *> incoming;
    std::map<long long, Req *> finished;
    std::set<long long> waited;  // ids a bh_server_wait has taken (or is waiting for)
    // worker-owned
    std::vector<Slot *> active;
    std::map<std::string, std::vector<Shell *>> pool;
    std::vector<Shell *> shells;  // every shell (active and pooled)
    std::map<std::string, std::shared_ptr<const DevMem>> profiles;
    std::map<std::string, const double *> profile_ptr;
    std::vector<hipEvent_t> free_events;
    // statistics (under mu)
    long long submitted = 0, completed = 0, solves = 0, iterations = 0;
    int nactive = 0, peak_active = 0;
    size_t bytes_in_use = 0, peak_bytes = 0;
    std::vector<double> lat_ms;
    double t0 = 0, t_first = -1, t_last = 0;
    std::string worker_error;
    // iteration log (worker-written, read under mu)
    std::vector<IterRec> iter_log;
    int new_shells = 0;    // shells created since the last iteration
    double t_admit = 0;    // start of the last admission pass
};

namespace {

size_t resident_bytes(bh_server *S)
{
    size_t b = 0;
    std::set<const void *> seen;
    for (Shell *sh : S->shells) {
        b += sh->bytes;
        if (seen.insert(sh->plan->tab.get()).second) b += sh->plan->tab->mem.bytes;
        if (seen.insert(sh->plan->ops.get()).second) b += sh->plan->ops->mem.bytes;
    }
    for (auto &kv : S->profiles) b += kv.second->bytes;
    for (const Lane &L : S->lanes) b += L.ex.arena_bytes();
    return b;
}

size_t shell_new_bytes(const bh_grid_desc &d, int precision, int device)
{
    const size_t n = (size_t)d.numr * d.numz * d.numphi;
    return plan_private_bytes(d, precision) + 2 * n * sizeof(double) + plan_new_shared_bytes(d, precision, device);
}

void destroy_shell(bh_server *S, Shell *sh)
{
    S->shells.erase(std::remove(S->shells.begin(), S->shells.end(), sh), S->shells.end());
    delete sh;
}

// A pooled shell of the grid, or a new one (nullptr and rc on failure).
Shell *take_shell(bh_server *S, const bh_grid_desc &d, int *rc)
{
    const std::string key = grid_key(d);
    auto it = S->pool.find(key);
    if (it != S->pool.end() && !it->second.empty()) {
        Shell *sh = it->second.back();
        it->second.pop_back();
        *rc = BH_OK;
        return sh;
    }
    std::unique_ptr<Shell> sh(new Shell());
    sh->key = key;
    S->new_shells++;
    if ((*rc = plan_create(d, S->cfg.precision, 0, &sh->plan)) != BH_OK) return nullptr;
    const size_t n = (size_t)d.numr * d.numz * d.numphi;
    if ((*rc = sh->mem.alloc(&sh->rho, n)) != BH_OK || (*rc = sh->mem.alloc(&sh->pot, n)) != BH_OK) return nullptr;
    sh->bytes = sh->plan->mem.bytes + sh->mem.bytes;
    S->shells.push_back(sh.get());
    return sh.release();
}

// Device profile table for (n_core, n_env, core_frac), built once.
int get_profile(bh_server *S, const bh_binary_model &m, std::shared_ptr<const DevMem> *mem, const double **ptr)
{
    char b[128];
    std::snprintf(b, sizeof b, "%.17g|%.17g|%.17g", m.n_core, m.n_env, m.core_frac);
    auto it = S->profiles.find(b);
    if (it != S->profiles.end()) {
        *mem = it->second;
        *ptr = S->profile_ptr[b];
        return BH_OK;
    }
    std::vector<double> tab;
    double info[2];
    int rc = bipolytrope_profile(m.n_core, m.n_env, m.core_frac, NPROF, &tab, info);
    if (rc != BH_OK) return rc;
    auto dm = std::make_shared<DevMem>();
    double *p = nullptr;
    if ((rc = dm->alloc(&p, NPROF)) != BH_OK) return rc;
    HIPCHK(hipMemcpy(p, tab.data(), NPROF * sizeof(double), hipMemcpyHostToDevice));
    S->profiles[b] = dm;
    S->profile_ptr[b] = p;
    *mem = dm;
    *ptr = p;
    return BH_OK;
}

void finish_request(bh_server *S, Req *q, int status)
{
    q->status = status;
    q->t_done = now_s();
    if (q->r.solve_ms)
        for (size_t i = 0; i < q->lat_ms.size() && i < (size_t)q->r.nsolves; ++i) q->r.solve_ms[i] = q->lat_ms[i];
    std::lock_guard<std::mutex> lk(S->mu);
    S->finished[q->id] = q;
    S->completed++;
    S->t_last = q->t_done;
    S->done_cv.notify_all();
}

// Admit waiting requests (first come, first served) while slots and device
// memory allow; idle pooled plans are released to make room.
bool any_inflight(const bh_server *S)
{
    for (const Lane &L : S->lanes)
        if (!L.inflight.empty()) return true;
    return false;
}

void admit(bh_server *S)
{
    for (;;) {
        Req *q = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->incoming.empty() || (int)S->active.size() >= S->cfg.max_active) return;
            q = S->incoming.front();
        }
        bh_grid_desc d;
        normalize_desc(&q->r.grid, S->cfg.precision, &d);
        const std::string key = grid_key(d);
        auto pit = S->pool.find(key);
        const bool pooled = pit != S->pool.end() && !pit->second.empty();
        size_t need = pooled ? 0 : shell_new_bytes(d, S->cfg.precision, S->device);
        size_t used = resident_bytes(S);
        // release idle plans of other grid sizes until the request fits
        while (S->cfg.hbm_budget && used + need > S->cfg.hbm_budget) {
            Shell *victim = nullptr;
            for (auto &kv : S->pool)
                if (kv.first != key && !kv.second.empty()) {
                    victim = kv.second.back();
                    kv.second.pop_back();
                    break;
                }
            if (!victim) break;
            destroy_shell(S, victim);
            need = pooled ? 0 : shell_new_bytes(d, S->cfg.precision, S->device);
            used = resident_bytes(S);
        }
        if (S->cfg.hbm_budget && used + need > S->cfg.hbm_budget) {
            if (S->active.empty() && !any_inflight(S)) {
                // cannot fit even alone
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    S->incoming.pop_front();
                }
                finish_request(S, q, BH_E_NOMEM);
                continue;
            }
            return;  // wait for running requests to finish
        }
        int rc = BH_OK;
        Shell *sh = take_shell(S, d, &rc);
        std::shared_ptr<const DevMem> prof;
        const double *prof_ptr = nullptr;
        if (sh) rc = get_profile(S, q->r.model, &prof, &prof_ptr);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->incoming.pop_front();
        }
        if (!sh || rc != BH_OK) {
            if (sh) S->pool[sh->key].push_back(sh);
            finish_request(S, q, rc != BH_OK ? rc : BH_E_DEVICE);
            continue;
        }
        const int lane = lane_of(d);
        if (!q->start.empty()) {
            // resume: the potential into the shell's buffer, its transform
            // into the plan's warm-start state, ahead of the first solve on
            // the lane's stream
            const hipStream_t ls = S->lanes[lane].stream;
            hipError_t e = hipMemcpyAsync(sh->pot, q->start.data(), q->start.size() * sizeof(double),
                                          hipMemcpyHostToDevice, ls);
            if (e == hipSuccess) {
                bh_plan *p = sh->plan;
                const long M = (long)p->g.numr * p->g.numz;
                std::vector<Op> ops;
                dft_fwd64_ops(p, sh->pot, M, (int)M, p->XJ, M, false, &ops);
                p->lazy_state = false;  // XJ is rewritten whole
                p->faces_clean = false;
                if (p->lazy_flag) launch_fillf(p->lazy_flag, 1, 0.f, ls);
                for (const Op &op : ops) launch_op(op, ls);
                e = hipGetLastError();
            }
            // (q->start lives until the request is collected, after the copy)
            if (e != hipSuccess) {
                S->pool[sh->key].push_back(sh);
                finish_request(S, q, fail(BH_E_DEVICE, std::string("resume: ") + hipGetErrorString(e)));
                continue;
            }
        }
        Slot *sl = new Slot();
        sl->sh = sh;
        sl->req = q;
        sl->lane = lane;
        sl->resume = !q->start.empty();
        sl->prof = prof;
        sl->prof_ptr = prof_ptr;
        q->t_admit = now_s();
        sl->t_ready = q->t_submit;
        S->active.push_back(sl);
        std::lock_guard<std::mutex> lk(S->mu);
        S->nactive = (int)S->active.size();
        S->peak_active = std::max(S->peak_active, S->nactive);
        S->bytes_in_use = resident_bytes(S);
        S->peak_bytes = std::max(S->peak_bytes, S->bytes_in_use);
    }
}

// Enqueue the next solve of every active request of the lane that has one
// left.
int enqueue_iteration(bh_server *S, int lane, bool *any)
{
    *any = false;
    Lane &Ln = S->lanes[lane];
    Iteration it;
    Ln.lists.clear();
    for (Slot *sl : S->active) {
        if (sl->lane != lane || sl->enq >= sl->req->r.nsolves) continue;
        const int i = sl->enq;
        bh_plan *p = sl->sh->plan;
        std::vector<Op> ops;
        build_ops(p, sl->sh->rho, sl->sh->pot, (i == 0 && !sl->resume) ? 0 : 2, &ops);
        ops.insert(ops.begin(), density_op(model_args(sl->req->r.model, p->g, i, sl->sh->rho, sl->prof_ptr)));
        Ln.lists.push_back(std::move(ops));
        it.solves.push_back({sl, i});
        sl->enq++;
    }
    if (it.solves.empty()) return BH_OK;
    IterRec rec{S->t_admit - S->t0, now_s() - S->t0, 0, 0, (int)it.solves.size(), S->new_shells, 0, lane};
    S->new_shells = 0;
    for (auto &sv : it.solves) {
        const Geo &g = sv.first->sh->plan->g;
        rec.points += (double)g.numr * g.numz * g.numphi;
    }
    bool same;
    int rc;
    if ((rc = Ln.ex.prepare(Ln.lists, &same)) != BH_OK) return rc;
    if ((rc = Ln.ex.upload(Ln.stream)) != BH_OK) return rc;
    if ((rc = Ln.ex.launch(Ln.stream, nullptr)) != BH_OK) return rc;
    if (S->free_events.empty()) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        S->free_events.push_back(e);
    }
    it.ev = S->free_events.back();
    S->free_events.pop_back();
    HIPCHK(hipEventRecord(it.ev, Ln.stream));
    rec.t_enq1 = now_s() - S->t0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        it.rec = S->iter_log.size();
        S->iter_log.push_back(rec);
    }
    Ln.inflight.push_back(std::move(it));
    *any = true;
    return BH_OK;
}

// Wait for the lane's oldest iteration and account for its solves.
int complete_oldest(bh_server *S, int lane)
{
    Lane &Ln = S->lanes[lane];
    Iteration it = std::move(Ln.inflight.front());
    Ln.inflight.pop_front();
    hipError_t e = hipEventSynchronize(it.ev);
    const double t = now_s();
    S->free_events.push_back(it.ev);
    std::vector<Slot *> done_slots;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->iterations++;
        if (it.rec < S->iter_log.size()) S->iter_log[it.rec].t_done = t - S->t0;
        for (auto &sv : it.solves) {
            Slot *sl = sv.first;
            const double ms = (t - sl->t_ready) * 1e3;
            sl->req->lat_ms.push_back(ms);
            S->lat_ms.push_back(ms);
            S->solves++;
            sl->t_ready = t;
            sl->done++;
            if (sl->done == sl->req->r.nsolves) done_slots.push_back(sl);
        }
    }
    for (Slot *sl : done_slots) {
        Req *q = sl->req;
        int status = e == hipSuccess ? BH_OK : BH_E_DEVICE;
        if (status == BH_OK && q->r.pot_out) {
            const Geo &g = sl->sh->plan->g;
            const size_t n = (size_t)g.numr * g.numz * g.numphi;
            if (hipMemcpy(q->r.pot_out, sl->sh->pot, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                status = BH_E_DEVICE;
        }
        S->active.erase(std::remove(S->active.begin(), S->active.end(), sl), S->active.end());
        S->pool[sl->sh->key].push_back(sl->sh);
        delete sl;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->nactive = (int)S->active.size();
        }
        finish_request(S, q, status);
    }
    if (e != hipSuccess) return fail(BH_E_DEVICE, std::string("serving iteration: ") + hipGetErrorString(e));
    return BH_OK;
}

void worker_main(bh_server *S)
{
    (void)hipSetDevice(S->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] {
                return S->stop || !S->incoming.empty() || !S->active.empty() || any_inflight(S);
            });
            if (S->stop && S->incoming.empty() && S->active.empty() && !any_inflight(S)) return;
        }
        S->t_admit = now_s();
        admit(S);
        // per lane: account for the finished iterations, then keep two in
        // flight
        bool progress = false;
        int rc = BH_OK;
        for (int l = 0; l < NLANES && rc == BH_OK; ++l) {
            Lane &Ln = S->lanes[l];
            while (rc == BH_OK && !Ln.inflight.empty()) {
                const hipError_t q = hipEventQuery(Ln.inflight.front().ev);
                if (q == hipErrorNotReady) break;
                rc = complete_oldest(S, l);  // done, or failed: its error is reported there
                progress = true;
            }
            if (rc == BH_OK && Ln.inflight.size() < 2) {
                bool any = false;
                rc = enqueue_iteration(S, l, &any);
                progress = progress || any;
            }
        }
        if (rc == BH_OK && !progress && any_inflight(S))
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        if (rc != BH_OK) {
            // fail every request in the loop and keep serving new ones
            std::string msg = last_error();
            for (int l = 0; l < NLANES; ++l)
                while (!S->lanes[l].inflight.empty()) complete_oldest(S, l);
            for (Slot *sl : S->active) {
                S->pool[sl->sh->key].push_back(sl->sh);
                finish_request(S, sl->req, rc);
                delete sl;
            }
            S->active.clear();
            std::lock_guard<std::mutex> lk(S->mu);
            S->nactive = 0;
            S->worker_error = msg;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        S->bytes_in_use = resident_bytes(S);
        S->peak_bytes = std::max(S->peak_bytes, S->bytes_in_use);
    }
}

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" int bh_server_create(const bh_server_config *cfg, bh_server **out)
{
    if (!cfg || !out) return fail(BH_E_INVALID, "null argument");
    *out = nullptr;
    if (!valid_precision(cfg->precision))
        return fail(BH_E_INVALID, "precision must be BH_PREC_FP64, BH_PREC_MIXED, BH_PREC_MIXED_TABLE32 or "
                                  "BH_PREC_SPECTRAL64");
    std::unique_ptr<bh_server> S(new bh_server());
    S->cfg = *cfg;
    if (S->cfg.max_active <= 0) S->cfg.max_active = 64;
    HIPCHK(hipGetDevice(&S->device));
    if (!S->cfg.hbm_budget) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        S->cfg.hbm_budget = fr / 10 * 9;
    }
    for (Lane &L : S->lanes) HIPCHK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    S->t0 = now_s();
    bh_server *raw = S.get();
    S->worker = std::thread(worker_main, raw);
    *out = S.release();
    return BH_OK;
}

extern "C" int bh_server_submit(bh_server *S, const bh_request *req, long long *id)
{
    if (!S || !req) return fail(BH_E_INVALID, "null argument");
    bh_grid_desc d;
    int rc = normalize_desc(&req->grid, S->cfg.precision, &d);
    if (rc != BH_OK) return rc;
    if (req->nsolves < 1) return fail(BH_E_INVALID, "nsolves must be >= 1");
    const size_t alone = plan_private_bytes(d, S->cfg.precision) + plan_shared_bytes(d, S->cfg.precision) +
                         2 * (size_t)d.numr * d.numz * d.numphi * sizeof(double);
    if (alone > S->cfg.hbm_budget)
        return fail(BH_E_NOMEM, "request needs " + std::to_string(alone) + " bytes, over the server budget of " +
                                    std::to_string(S->cfg.hbm_budget));
    Req *q = new Req();
    q->r = *req;
    q->r.grid = d;
    if (req->start_pot) q->start.assign(req->start_pot, req->start_pot + (size_t)d.numr * d.numz * d.numphi);
    q->r.start_pot = nullptr;
    fill_model_defaults(&q->r.model);
    if (q->r.model.radius[0] <= 0.0 || q->r.model.radius[1] <= 0.0) {
        delete q;
        return fail(BH_E_INVALID, "model radii must be positive");
    }
    q->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete q;
            return fail(BH_E_STATE, "server is shutting down");
        }
        q->id = S->next_id++;
        if (S->t_first < 0) S->t_first = q->t_submit;
        S->incoming.push_back(q);
        S->submitted++;
    }
    S->cv.notify_all();
    if (id) *id = q->id;
    return BH_OK;
}

extern "C" int bh_server_wait(bh_server *S, long long id, bh_request_result *res)
{
    if (!S) return fail(BH_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    if (id <= 0 || id >= S->next_id) return fail(BH_E_INVALID, "unknown request id");
    // each id is waited for once: a second wait (or a concurrent one) would
    // never see it finish
    if (!S->waited.insert(id).second)
        return fail(BH_E_STATE, "request " + std::to_string(id) + " is already being or has been waited for");
    S->done_cv.wait(lk, [&] { return S->finished.count(id) > 0; });
    Req *q = S->finished[id];
    S->finished.erase(id);
    if (res) {
        res->id = q->id;
        res->status = q->status;
        res->nsolves = (int)q->lat_ms.size();
        res->submit_s = q->t_submit - S->t0;
        res->admit_s = q->t_admit > 0 ? q->t_admit - S->t0 : -1.0;
        res->done_s = q->t_done - S->t0;
    }
    const int st = q->status;
    delete q;
    return st == BH_OK ? BH_OK : fail(st, "request " + std::to_string(id) + " failed");
}

extern "C" int bh_server_drain(bh_server *S)
{
    if (!S) return fail(BH_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    S->done_cv.wait(lk, [&] { return S->completed == S->submitted; });
    return BH_OK;
}

extern "C" int bh_server_get_stats(bh_server *S, bh_server_stats *st)
{
    if (!S || !st) return fail(BH_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    st->submitted = S->submitted;
    st->completed = S->completed;
    st->solves = S->solves;
    st->iterations = S->iterations;
    st->active = S->nactive;
    st->waiting = (int)S->incoming.size();
    st->peak_active = S->peak_active;
    st->bytes_in_use = S->bytes_in_use;
    st->peak_bytes = S->peak_bytes;
    st->elapsed_s = S->t_first >= 0 ? S->t_last - S->t_first : 0.0;
    st->p50_ms = percentile(S->lat_ms, 50.0);
    st->p99_ms = percentile(S->lat_ms, 99.0);
    double sum = 0;
    for (double x : S->lat_ms) sum += x;
    st->mean_ms = S->lat_ms.empty() ? 0.0 : sum / (double)S->lat_ms.size();
    return BH_OK;
}

extern "C" int bh_server_destroy(bh_server *S)
{
    if (!S) return BH_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    for (Lane &L : S->lanes) (void)hipStreamSynchronize(L.stream);
    for (auto &kv : S->finished) delete kv.second;
    for (Shell *sh : std::vector<Shell *>(S->shells)) destroy_shell(S, sh);
    for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
    for (Lane &L : S->lanes) (void)hipStreamDestroy(L.stream);
    delete S;
    return BH_OK;
}

// ---------------------------------------------------------------------------
// Test hooks: the bipolytrope profile table, and one density on the device.
// ---------------------------------------------------------------------------
extern "C" int bhx_bipolytrope_profile(double n_core, double n_env, double core_frac, double *table, int ntab,
                                       double *info)
{
    if (!table || !info) return fail(BH_E_INVALID, "null argument");
    std::vector<double> t;
    int rc = bipolytrope_profile(n_core, n_env, core_frac, ntab, &t, info);
    if (rc != BH_OK) return rc;
    std::copy(t.begin(), t.end(), table);
    return BH_OK;
}

extern "C" int bhx_density(const bh_grid_desc *desc, const bh_binary_model *model, int solve_index, double *d_rho)
{
    if (!desc || !model || !d_rho) return fail(BH_E_INVALID, "null argument");
    bh_grid_desc d;
    int rc = normalize_desc(desc, BH_PREC_FP64, &d);
    if (rc != BH_OK) return rc;
    HostGeo h;
    host_geometry(d, &h);
    bh_binary_model m = *model;
    fill_model_defaults(&m);
    std::vector<double> tab;
    double info[2];
    if ((rc = bipolytrope_profile(m.n_core, m.n_env, m.core_frac, NPROF, &tab, info)) != BH_OK) return rc;
    DevMem mem;
    double *p;
    if ((rc = mem.alloc(&p, NPROF)) != BH_OK) return rc;
    HIPCHK(hipMemcpy(p, tab.data(), NPROF * sizeof(double), hipMemcpyHostToDevice));
    launch_op(density_op(model_args(m, h.g, solve_index, d_rho, p)), 0);
    HIPCHK(hipDeviceSynchronize());
    return BH_OK;
}

// Iteration log (test and bench hook, not part of the public ABI): up to n
// records of 8 doubles (t_admit, t_enq0, t_enq1, t_done, nsolves,
// new_shells, points, lane) into out; returns the number of iterations logged.
extern "C" int bhx_server_iterations(bh_server *S, double *out, int n)
{
    if (!S) return fail(BH_E_INVALID, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    const int m = (int)S->iter_log.size();
    for (int i = 0; i < m && i < n && out; ++i) {
        const IterRec &r = S->iter_log[i];
        const double v[8] = {r.t_admit, r.t_enq0, r.t_enq1, r.t_done, (double)r.nsolves, (double)r.new_shells,
                             r.points, (double)r.lane};
        for (int k = 0; k < 8; ++k) out[8 * i + k] = v[k];
    }
    return m;
}
