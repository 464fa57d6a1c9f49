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
// candidates timed by the autotuner, one per solve: GEMM tile shape, and
// line-sweep lines per workgroup
constexpr int TUNE_TILES[] = {0, 1, 2, 3, 4};
constexpr int TUNE_NL[] = {16, 8, 16, 8, 16};
constexpr int TUNE_ROUNDS = 5;
}  // namespace

struct bh_batch {
    int precision = BH_PREC_MIXED, flags = 0, device = 0;
    size_t budget = 0, bytes = 0;
    std::vector<bh_plan *> plans;
    GroupExec ex;
    Tuning tune;
    int tune_round = 0;  // next autotuning round (TUNE_ROUNDS: done)
    std::vector<std::vector<double>> tune_ms;  // [round][position]
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

// Tuning of autotuning round r: candidate r at every tunable position.
Tuning round_tuning(const bh_batch *b, int r)
{
    Tuning t;
    size_t maxlen = 0;
    for (const auto &l : b->lists) maxlen = std::max(maxlen, l.size());
    t.tile.assign(maxlen, -1);
    t.nl.assign(maxlen, -1);
    for (const auto &l : b->lists)
        for (size_t k = 0; k < l.size(); ++k) {
            if (l[k].kind == OP_GEMM) t.tile[k] = TUNE_TILES[r];
            else if (l[k].kind == OP_SWEEP) t.nl[k] = TUNE_NL[r];
        }
    return t;
}

// After the last round: the fastest candidate per position.
void finish_tuning(bh_batch *b)
{
    size_t maxlen = 0;
    for (const auto &l : b->lists) maxlen = std::max(maxlen, l.size());
    Tuning best;
    best.tile.assign(maxlen, -1);
    best.nl.assign(maxlen, -1);
    for (size_t k = 0; k < maxlen; ++k) {
        double bt = 1e300;
        int br = -1;
        for (int r = 0; r < TUNE_ROUNDS; ++r) {
            if (k >= b->tune_ms[r].size()) continue;
            const double t = b->tune_ms[r][k];
            if (t > 0.0 && t < bt) {
                bt = t;
                br = r;
            }
        }
        if (br < 0) continue;
        bool gemm = false, sweep = false;
        for (const auto &l : b->lists)
            if (k < l.size()) {
                gemm |= l[k].kind == OP_GEMM;
                sweep |= l[k].kind == OP_SWEEP;
            }
        if (gemm) best.tile[k] = TUNE_TILES[br];
        if (sweep) best.nl[k] = TUNE_NL[br];
    }
    b->tune = best;
}

int batch_enqueue(bh_batch *b, const double *const *d_rho, double *const *d_pot, const int *tstep, hipStream_t s)
{
    const int n = (int)b->plans.size();
    b->lists.resize(n);
    for (int i = 0; i < n; ++i) build_ops(b->plans[i], d_rho[i], d_pot[i], tstep ? tstep[i] : 0, &b->lists[i]);
    const bool tuning = (b->flags & BH_BATCH_AUTOTUNE) && b->tune_round < TUNE_ROUNDS;
    const Tuning t = tuning ? round_tuning(b, b->tune_round) : b->tune;
    for (auto &l : b->lists) apply_tuning(t, &l);
    bool same = false;
    int rc = b->ex.prepare(b->lists, &same);
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
        if (++b->tune_round == TUNE_ROUNDS) finish_tuning(b);
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
    st->tuning_left = (b->flags & BH_BATCH_AUTOTUNE) ? std::max(0, TUNE_ROUNDS - b->tune_round) : 0;
    st->last_ms = b->last_ms;
    return BH_OK;
}

extern "C" int bh_batch_tuning(const bh_batch *b, int *tile, int *nl, int n)
{
    if (!b) return fail(BH_E_INVALID, "null batch");
    const int m = (int)std::max(b->tune.tile.size(), b->tune.nl.size());
    for (int k = 0; k < n && k < m; ++k) {
        if (tile) tile[k] = b->tune.at(b->tune.tile, k);
        if (nl) nl[k] = b->tune.at(b->tune.nl, k);
    }
    return m;
}

extern "C" int bh_batch_plan(bh_batch *b, int i, bh_plan **plan)
{
    if (!b || !plan || i < 0 || i >= (int)b->plans.size()) return fail(BH_E_INVALID, "bh_batch_plan: bad argument");
    *plan = b->plans[i];
    return BH_OK;
}
