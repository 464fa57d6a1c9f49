// bh_batch.cpp -- batches of independent grids of different sizes, solved
// together by grouped persistent launches (bh_exec.h), optionally replayed
// from a captured hipGraph, with tile and pencil shapes picked by timing the
// candidates over the first solves.
#include "bh_exec.h"
#include "bh_plan_impl.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <set>

using namespace bh;

namespace {
// Candidates of each tuning axis (bh_exec.h), the default first, and the Op
// kind each applies to (-1: every kernel launch).
const std::vector<int> TUNE_CANDS[TUNE_AXES] = {{2, 0, 1, 3, 4}, {16, 8}, {4, 2, 8}, {0, 1, 2, 4}, {0, 1}};
constexpr int TUNE_KIND[TUNE_AXES] = {OP_GEMM, OP_SWEEP, OP_RESIDUAL, -1, -1};

bool axis_applies(int axis, int kind) { return kind != OP_MARK && (TUNE_KIND[axis] < 0 || TUNE_KIND[axis] == kind); }
}  // namespace

struct bh_batch {
    int precision = BH_PREC_MIXED, flags = 0, device = 0;
    size_t budget = 0, bytes = 0;
    std::vector<bh_plan *> plans;
    GroupExec ex;
    // autotuning: the axes in use, searched one after another, every
    // candidate of the current axis timed at every position it applies to
    // (one solve per candidate) with the other axes at their best so far
    Tuning tune;                    // best choices so far (the result, once done)
    bool tune_init = false;
    std::vector<int> tune_axes;     // axes with at least one position
    int tune_ax = 0, tune_cand = 0; // next round: axis tune_axes[tune_ax], candidate tune_cand
    int tune_total = 0, tune_done = 0;
    std::vector<std::vector<double>> tune_ms;  // [candidate][position] of the current axis
    std::vector<double> tune_gain;  // per axis searched: default ms - chosen ms (summed over positions)
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool timing = false;
    std::vector<hipEvent_t> ev;  // per-launch events (tuning, or timing without a graph)
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool timed_pending = false;
    double last_ms = 0.0;
    std::vector<std::vector<Op>> lists;

    ~bh_batch()
    {
        (void)hipDeviceSynchronize();
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        for (bh_plan *p : plans) bh_plan_destroy(p);
    }
};

namespace {

int ensure_events(bh_batch *b, int n)
{
    while ((int)b->ev.size() < n) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        b->ev.push_back(e);
    }
    if (!b->t0) HIPCHK(hipEventCreate(&b->t0));
    if (!b->t1) HIPCHK(hipEventCreate(&b->t1));
    return BH_OK;
}

size_t max_len(const bh_batch *b)
{
    size_t m = 0;
    for (const auto &l : b->lists) m = std::max(m, l.size());
    return m;
}

// positions where `axis` applies
std::vector<bool> axis_positions(const bh_batch *b, int axis)
{
    std::vector<bool> on(max_len(b), false);
    for (const auto &l : b->lists)
        for (size_t k = 0; k < l.size(); ++k)
            if (axis_applies(axis, l[k].kind)) on[k] = true;
    return on;
}

// First tuning solve: the axes in use and the number of rounds.
void init_tuning(bh_batch *b)
{
    const size_t m = max_len(b);
    for (int ax = 0; ax < TUNE_AXES; ++ax) b->tune.v[ax].assign(m, -1);
    b->tune_axes.clear();
    b->tune_total = 0;
    for (int ax = 0; ax < TUNE_AXES; ++ax) {
        const std::vector<bool> on = axis_positions(b, ax);
        if (std::find(on.begin(), on.end(), true) == on.end()) continue;
        b->tune_axes.push_back(ax);
        b->tune_total += (int)TUNE_CANDS[ax].size();
    }
    b->tune_ax = b->tune_cand = b->tune_done = 0;
    b->tune_ms.clear();
    b->tune_gain.assign(TUNE_AXES, 0.0);
    b->tune_init = true;
}

// Tuning of the next round: the best so far, the current axis at its
// current candidate wherever it applies.
Tuning round_tuning(const bh_batch *b)
{
    Tuning t = b->tune;
    const int ax = b->tune_axes[b->tune_ax];
    const std::vector<bool> on = axis_positions(b, ax);
    t.v[ax].assign(on.size(), -1);
    for (size_t k = 0; k < on.size(); ++k)
        if (on[k]) t.v[ax][k] = TUNE_CANDS[ax][b->tune_cand];
    return t;
}

// After the last candidate of an axis: the fastest per position.
void finish_axis(bh_batch *b)
{
    const int ax = b->tune_axes[b->tune_ax];
    const std::vector<bool> on = axis_positions(b, ax);
    b->tune.v[ax].assign(on.size(), -1);
    double gain = 0.0;
    for (size_t k = 0; k < on.size(); ++k) {
        if (!on[k]) continue;
        double best = 1e300;
        int bc = 0;
        for (size_t c = 0; c < b->tune_ms.size(); ++c) {
            const double t = k < b->tune_ms[c].size() ? b->tune_ms[c][k] : 0.0;
            if (t > 0.0 && t < best) {
                best = t;
                bc = (int)c;
            }
        }
        b->tune.v[ax][k] = TUNE_CANDS[ax][bc];
        if (k < b->tune_ms[0].size() && best < 1e300) gain += b->tune_ms[0][k] - best;
    }
    b->tune_gain[ax] = gain;
    b->tune_ms.clear();
}

int batch_enqueue(bh_batch *b, const double *const *d_rho, double *const *d_pot, const int *tstep, hipStream_t s)
{
    const int n = (int)b->plans.size();
    b->lists.resize(n);
    for (int i = 0; i < n; ++i) build_ops(b->plans[i], d_rho[i], d_pot[i], tstep ? tstep[i] : 0, &b->lists[i]);
    if ((b->flags & BH_BATCH_AUTOTUNE) && !b->tune_init) init_tuning(b);
    const bool tuning = (b->flags & BH_BATCH_AUTOTUNE) && b->tune_ax < (int)b->tune_axes.size();
    const Tuning t = tuning ? round_tuning(b) : b->tune;
    for (auto &l : b->lists) apply_tuning(t, &l);
    bool same = false;
    int rc = b->ex.prepare(b->lists, &same, &t);
    if (rc != BH_OK) return rc;
    if ((rc = b->ex.upload(s)) != BH_OK) return rc;
    const int nl = b->ex.nlaunches();
    if (tuning) {
        // time every launch of this candidate; wait for the result
        if ((rc = ensure_events(b, nl + 1)) != BH_OK) return rc;
        if ((rc = b->ex.launch(s, &b->ev)) != BH_OK) return rc;
        HIPCHK(hipEventSynchronize(b->ev[nl]));
        std::vector<double> ms;
        for (int i = 0; i < nl; ++i) {
            float d;
            HIPCHK(hipEventElapsedTime(&d, b->ev[i], b->ev[i + 1]));
            const size_t pos = (size_t)b->ex.position(i);
            if (ms.size() <= pos) ms.resize(pos + 1, 0.0);
            ms[pos] += d;
        }
        b->tune_ms.push_back(ms);
        float tot;
        HIPCHK(hipEventElapsedTime(&tot, b->ev[0], b->ev[nl]));
        b->last_ms = tot;
        b->tune_done++;
        if (++b->tune_cand == (int)TUNE_CANDS[b->tune_axes[b->tune_ax]].size()) {
            finish_axis(b);
            b->tune_ax++;
            b->tune_cand = 0;
        }
        return BH_OK;
    }
    if (b->flags & BH_BATCH_GRAPH) {
        if (!b->gexec || !same) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->gexec = nullptr;
            b->graph = nullptr;
            if (!b->cap_stream) HIPCHK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
            HIPCHK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
            rc = b->ex.launch(b->cap_stream, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
            if (rc != BH_OK) return rc;
            if (e != hipSuccess) return fail(BH_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            b->graph = g;
            HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        if (b->timing) {
            if ((rc = ensure_events(b, 1)) != BH_OK) return rc;
            HIPCHK(hipEventRecord(b->t0, s));
        }
        HIPCHK(hipGraphLaunch(b->gexec, s));
        if (b->timing) {
            HIPCHK(hipEventRecord(b->t1, s));
            b->timed_pending = true;
        }
        return BH_OK;
    }
    if (b->timing) {
        if ((rc = ensure_events(b, 1)) != BH_OK) return rc;
        HIPCHK(hipEventRecord(b->t0, s));
    }
    if ((rc = b->ex.launch(s, nullptr)) != BH_OK) return rc;
    if (b->timing) {
        HIPCHK(hipEventRecord(b->t1, s));
        b->timed_pending = true;
    }
    return BH_OK;
}

}  // namespace

extern "C" int bh_batch_create(const bh_grid_desc *descs, int ngrids, int precision, size_t hbm_budget, int flags,
                               bh_batch **out)
{
    if (!descs || !out || ngrids <= 0) return fail(BH_E_INVALID, "bh_batch_create: null argument or ngrids <= 0");
    *out = nullptr;
    std::vector<bh_grid_desc> d(ngrids);
    for (int i = 0; i < ngrids; ++i) {
        int rc = normalize_desc(&descs[i], precision, &d[i]);
        if (rc != BH_OK) return fail(rc, "grid " + std::to_string(i) + ": " + last_error());
    }
    int device = 0;
    HIPCHK(hipGetDevice(&device));
    // device memory: every grid's own buffers, plus each distinct geometry's
    // shared resources once (only those not already resident)
    size_t need = 0;
    std::set<std::string> seen;
    for (const auto &x : d) {
        need += plan_private_bytes(x, precision);
        char key[200];
        std::snprintf(key, sizeof key, "%d|%d|%d|%.17g|%d|%d|%d", x.numr, x.numz, x.numphi, x.dr, x.mmax,
                      x.nsteps_cold, x.nsteps_warm);
        if (seen.insert(key).second) need += plan_new_shared_bytes(x, precision, device);
    }
    if (hbm_budget && need > hbm_budget)
        return fail(BH_E_NOMEM, "batch needs about " + std::to_string(need) + " bytes of device memory, over the budget of " +
                                    std::to_string(hbm_budget));
    std::unique_ptr<bh_batch> b(new bh_batch());
    b->precision = precision;
    b->flags = flags;
    b->budget = hbm_budget;
    b->device = device;
    std::set<const void *> shared;
    for (int i = 0; i < ngrids; ++i) {
        bh_plan *p = nullptr;
        int rc = plan_create(d[i], precision, 0, &p);
        if (rc != BH_OK) return fail(rc, "grid " + std::to_string(i) + ": " + last_error());
        p->in_batch = true;  // its state is always kept whole (batches may replay captured launches)
        b->plans.push_back(p);
        b->bytes += p->mem.bytes;
        if (shared.insert(p->tab.get()).second) b->bytes += p->tab->mem.bytes;
        if (shared.insert(p->ops.get()).second) b->bytes += p->ops->mem.bytes;
    }
    *out = b.release();
    return BH_OK;
}

extern "C" int bh_batch_destroy(bh_batch *b)
{
    delete b;
    return BH_OK;
}

extern "C" size_t bh_batch_device_bytes(const bh_batch *b) { return b ? b->bytes + b->ex.arena_bytes() : 0; }

extern "C" int bh_batch_solve_device(bh_batch *b, const double *const *d_rho, double *const *d_pot, const int *tstep,
                                     void *stream)
{
    if (!b || !d_rho || !d_pot) return fail(BH_E_INVALID, "null argument");
    for (size_t i = 0; i < b->plans.size(); ++i)
        if (!d_rho[i] || !d_pot[i]) return fail(BH_E_INVALID, "null rho or pot pointer for grid " + std::to_string(i));
    return batch_enqueue(b, d_rho, d_pot, tstep, (hipStream_t)stream);
}

extern "C" int bh_batch_solve(bh_batch *b, const double *const *rho, double *const *pot, const int *tstep)
{
    if (!b || !rho || !pot) return fail(BH_E_INVALID, "null argument");
    const size_t n = b->plans.size();
    std::vector<const double *> dr(n);
    std::vector<double *> dp(n);
    for (size_t i = 0; i < n; ++i) {
        bh_plan *p = b->plans[i];
        const size_t cnt = (size_t)p->g.numr * p->g.numz * p->g.numphi;
        if (!rho[i] || !pot[i]) return fail(BH_E_INVALID, "null rho or pot pointer for grid " + std::to_string(i));
        for (size_t j = 0; j < cnt; ++j)
            if (!std::isfinite(rho[i][j]))
                return fail(BH_E_NONFINITE, "grid " + std::to_string(i) + ": rho holds a non-finite value at index " +
                                                std::to_string(j));
        if (!p->rho_in) {
            int rc;
            if ((rc = p->mem.alloc(&p->rho_in, cnt)) != BH_OK) return rc;
            if ((rc = p->mem.alloc(&p->pot_out, cnt)) != BH_OK) return rc;
            b->bytes += 2 * cnt * sizeof(double);
        }
        HIPCHK(hipMemcpy(p->rho_in, rho[i], cnt * sizeof(double), hipMemcpyHostToDevice));
        dr[i] = p->rho_in;
        dp[i] = p->pot_out;
    }
    int rc = batch_enqueue(b, dr.data(), dp.data(), tstep, 0);
    if (rc != BH_OK) return rc;
    HIPCHK(hipDeviceSynchronize());
    for (size_t i = 0; i < n; ++i) {
        bh_plan *p = b->plans[i];
        const size_t cnt = (size_t)p->g.numr * p->g.numz * p->g.numphi;
        HIPCHK(hipMemcpy(pot[i], p->pot_out, cnt * sizeof(double), hipMemcpyDeviceToHost));
    }
    return BH_OK;
}

extern "C" int bh_batch_set_timing(bh_batch *b, int on)
{
    if (!b) return fail(BH_E_INVALID, "null batch");
    b->timing = on != 0;
    return BH_OK;
}

extern "C" int bh_batch_get_stats(bh_batch *b, bh_batch_stats *st)
{
    if (!b || !st) return fail(BH_E_INVALID, "null argument");
    if (b->timed_pending) {
        HIPCHK(hipEventSynchronize(b->t1));
        float d;
        HIPCHK(hipEventElapsedTime(&d, b->t0, b->t1));
        b->last_ms = d;
        b->timed_pending = false;
    }
    st->ngrids = (int)b->plans.size();
    st->launches = b->ex.nlaunches();
    st->graph = b->gexec != nullptr;
    st->tuning_left = (b->flags & BH_BATCH_AUTOTUNE) ? (b->tune_init ? b->tune_total - b->tune_done : -1) : 0;
    st->last_ms = b->last_ms;
    return BH_OK;
}

extern "C" int bh_batch_tuning(const bh_batch *b, int axis, int *choice, int n, double *gain_ms)
{
    if (!b) return fail(BH_E_INVALID, "null batch");
    if (axis < 0 || axis >= TUNE_AXES) return fail(BH_E_INVALID, "bh_batch_tuning: axis out of range");
    const std::vector<int> &v = b->tune.v[axis];
    for (int k = 0; k < n && k < (int)v.size(); ++k)
        if (choice) choice[k] = v[k];
    if (gain_ms) *gain_ms = axis < (int)b->tune_gain.size() ? b->tune_gain[axis] : 0.0;
    return (int)v.size();
}

extern "C" int bh_batch_plan(bh_batch *b, int i, bh_plan **plan)
{
    if (!b || !plan || i < 0 || i >= (int)b->plans.size()) return fail(BH_E_INVALID, "bh_batch_plan: bad argument");
    *plan = b->plans[i];
    return BH_OK;
}
