// bh_exec.cpp -- grouped execution of several grids' Op lists (see bh_exec.h).
#include "bh_exec.h"
#include "bh_gemm.h"
#include "bh_mixed.h"
#include "bh_internal.h"
#include "bh_plan_impl.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

namespace bh {

void apply_tuning(const Tuning &t, std::vector<Op> *ops)
{
    for (size_t k = 0; k < ops->size(); ++k) {
        Op &op = (*ops)[k];
        if (op.kind == OP_GEMM) {
            const int tile = t.at(TUNE_TILE, k);
            if (tile >= 0) {
                GemmArgs a = op.as<GemmArgs>();
                a.tile = tile;
                op = gemm_op(a, op.variant / 8);
            }
        } else if (op.kind == OP_SWEEP) {
            const int nl = t.at(TUNE_NL, k);
            if (nl > 0) op = sweep_op(op.as<LineArgs>(), op.ny, nl);
        } else if (op.kind == OP_RESIDUAL) {
            const int rows = t.at(TUNE_ROWS, k);
            if (rows > 0) op = residual_op(op.as<ResidualArgs>(), op.nz, rows);
        }
    }
}

namespace {
std::mutex g_occ_mu;
std::map<long, int> g_occ;

int occupancy(int kind, int variant)
{
    const long key = (long)kind * 1000000L + variant;
    std::lock_guard<std::mutex> lk(g_occ_mu);
    auto it = g_occ.find(key);
    if (it != g_occ.end()) return it->second;
    int b = op_blocks_per_cu(kind, variant);
    g_occ[key] = b;
    return b;
}

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }
}  // namespace

GroupExec::~GroupExec()
{
    for (int i = 0; i < 2; ++i) {
        if (copied_[i]) {
            (void)hipEventSynchronize(copied_[i]);
            (void)hipEventDestroy(copied_[i]);
        }
        if (pinned_[i]) (void)hipHostFree(pinned_[i]);
    }
    if (dev_) (void)hipFree(dev_);
}

int GroupExec::prepare(const std::vector<std::vector<Op>> &lists, bool *same_layout, const Tuning *tune)
{
    if (!cus_) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
    }
    launches_.clear();
    image_.clear();
    for (int &m : marks_) m = -1;
    std::string layout;
    size_t maxlen = 0;
    for (const auto &l : lists) maxlen = std::max(maxlen, l.size());
    size_t off = 0;
    std::vector<std::pair<long, std::vector<int>>> buckets;
    for (size_t k = 0; k < maxlen; ++k) {
        // a stage mark counts when every list still running has it here
        int mark = -1;
        bool all_marks = true;
        for (const auto &l : lists) {
            if (k >= l.size()) continue;
            if (l[k].kind == OP_MARK) mark = l[k].variant;
            else all_marks = false;
        }
        if (mark >= 0 && all_marks && mark < 8) marks_[mark] = (int)launches_.size();
        buckets.clear();
        for (size_t g = 0; g < lists.size(); ++g) {
            if (k >= lists[g].size() || lists[g][k].kind == OP_MARK) continue;
            const long key = (long)lists[g][k].kind * 1000000L + lists[g][k].variant;
            auto it = std::find_if(buckets.begin(), buckets.end(), [&](const auto &b) { return b.first == key; });
            if (it == buckets.end()) buckets.push_back({key, {(int)g}});
            else it->second.push_back((int)g);
        }
        for (const auto &b : buckets) {
            const Op &first = lists[b.second[0]][k];
            Launch L;
            L.pos = (int)k;
            L.kind = first.kind;
            L.variant = first.variant;
            L.argsz = first.argsz;
            L.ngroups = (int)b.second.size();
            L.args_off = align16(off);
            off = L.args_off + (size_t)L.ngroups * L.argsz;
            L.sched_off = align16(off);
            off = L.sched_off + (size_t)L.ngroups * sizeof(GroupTile);
            image_.resize(off);
            long total = 0;
            char sig[96];
            std::snprintf(sig, sizeof sig, "|%d,%d,%d,%zu:", L.kind, L.variant, L.ngroups, L.args_off);
            layout += sig;
            for (int i = 0; i < L.ngroups; ++i) {
                const Op &op = lists[b.second[i]][k];
                std::memcpy(&image_[L.args_off + (size_t)i * L.argsz], op.args, L.argsz);
                GroupTile gt{(int)total, op.nx, op.ny, op.nz};
                std::memcpy(&image_[L.sched_off + (size_t)i * sizeof(GroupTile)], &gt, sizeof gt);
                total += op.tiles();
                std::snprintf(sig, sizeof sig, "%d.%d.%d;", op.nx, op.ny, op.nz);
                layout += sig;
            }
            if (total > 0x7fffffffL) return fail(BH_E_INVALID, "grouped launch with more than 2^31 tiles");
            L.total = (int)total;
            L.nblocks = persistent_blocks(total, occupancy(L.kind, L.variant));
            if (tune) {
                // a cap of g workgroups per CU (a multiple of 8 keeps the
                // XCD-chunked tile order), and one more than a multiple of
                // 8 for the linear order (BH_TILE_LOOP chunks only grids
                // that are multiples of 8)
                const int g = tune->at(TUNE_GRID, k), lin = tune->at(TUNE_ORDER, k);
                int nb = L.nblocks & ~7;
                if (g > 0) nb = std::max(8, std::min(nb, g * cus_ / 8 * 8));
                L.nblocks = (lin == 1 && nb < 0x7ffffff8) ? nb + 1 : nb;
            }
            char sigb[32];
            std::snprintf(sigb, sizeof sigb, "#%d", L.nblocks);
            layout += sigb;
            launches_.push_back(L);
        }
    }
    bytes_ = align16(off);
    image_.resize(bytes_);
    bool moved = false;
    if (bytes_ > dev_cap_) {
        if (dev_) {
            HIPCHK(hipDeviceSynchronize());
            (void)hipFree(dev_);
            dev_ = nullptr;
        }
        size_t cap = std::max(bytes_ + bytes_ / 4, (size_t)4096);
        HIPCHK(hipMalloc((void **)&dev_, cap));
        dev_cap_ = cap;
        moved = true;
    }
    *same_layout = !moved && layout == layout_;
    layout_ = layout;
    return BH_OK;
}

int GroupExec::upload(hipStream_t s)
{
    if (!bytes_) return BH_OK;
    cur_ ^= 1;
    if (copied_[cur_]) HIPCHK(hipEventSynchronize(copied_[cur_]));  // the copy two solves ago has read it
    else HIPCHK(hipEventCreateWithFlags(&copied_[cur_], hipEventDisableTiming));
    if (pinned_cap_[cur_] < bytes_) {
        if (pinned_[cur_]) (void)hipHostFree(pinned_[cur_]);
        pinned_[cur_] = nullptr;
        size_t cap = bytes_ + bytes_ / 4;
        HIPCHK(hipHostMalloc((void **)&pinned_[cur_], cap, hipHostMallocDefault));
        pinned_cap_[cur_] = cap;
    }
    std::memcpy(pinned_[cur_], image_.data(), bytes_);
    HIPCHK(hipMemcpyAsync(dev_, pinned_[cur_], bytes_, hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(copied_[cur_], s));
    return BH_OK;
}

int GroupExec::launch(hipStream_t s, const std::vector<hipEvent_t> *events)
{
    if (events) HIPCHK(hipEventRecord((*events)[0], s));
    for (size_t i = 0; i < launches_.size(); ++i) {
        const Launch &L = launches_[i];
        GroupSched gs{(const GroupTile *)(dev_ + L.sched_off), L.ngroups, L.total};
        launch_op_grouped(L.kind, L.variant, dev_ + L.args_off, gs, L.nblocks, s);
        if (events) HIPCHK(hipEventRecord((*events)[i + 1], s));
    }
    HIPCHK(hipGetLastError());
    return BH_OK;
}

}  // namespace bh
