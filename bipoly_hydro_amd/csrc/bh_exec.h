// bh_exec.h -- grouped execution of the solves of several grids.
//
// GroupExec takes one Op list per grid and runs position k of every list
// together: the Ops at position k are bucketed by (kind, variant) and each
// bucket becomes one persistent grouped launch over the tiles of all its
// grids.  Argument blocks and tile schedules live in one device arena that
// is rewritten (one host-to-device copy) before every solve, so a captured
// launch sequence stays valid while pointers and cold/warm choices change.
#pragma once
#include "bh_ops.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace bh {

// Per-position overrides chosen by the autotuner (-1: the default), one
// vector per axis, indexed by Op position:
//   tile   GEMM tile shape (0..4)
//   nl     line-sweep lines per workgroup (8, 16)
//   rows   residual K rows per tile (2, 4, 8)
//   grid   persistent workgroups per CU of the grouped launch (0: one per
//          tile, as many as the hardware keeps resident)
//   order  tile order of the grouped launch: 0 XCD-chunked, 1 linear
// None of them changes any result: every choice computes each value with
// the same operations in the same order.
enum { TUNE_TILE = 0, TUNE_NL, TUNE_ROWS, TUNE_GRID, TUNE_ORDER, TUNE_AXES };
struct Tuning {
    std::vector<int> v[TUNE_AXES];
    int at(int axis, size_t k) const { return k < v[axis].size() ? v[axis][k] : -1; }
};

// Apply the overrides of `t` to one grid's Op list.
void apply_tuning(const Tuning &t, std::vector<Op> *ops);

class GroupExec {
  public:
    GroupExec() = default;
    GroupExec(const GroupExec &) = delete;
    GroupExec &operator=(const GroupExec &) = delete;
    ~GroupExec();

    // Lay out the grouped launches for these Op lists.  Returns BH_OK; sets
    // *same_layout when the launch sequence and arena layout equal the
    // previous call's (a captured graph of them is still valid).
    int prepare(const std::vector<std::vector<Op>> &lists, bool *same_layout, const Tuning *tune = nullptr);
    // Copy the argument arena to the device, stream-ordered on s.
    int upload(hipStream_t s);
    // Enqueue the grouped launches on s.  With `events` (size nlaunches()+1)
    // an event is recorded before the first launch and after each one.
    int launch(hipStream_t s, const std::vector<hipEvent_t> *events);

    int nlaunches() const { return (int)launches_.size(); }
    int position(int i) const { return launches_[i].pos; }
    int kind(int i) const { return launches_[i].kind; }
    size_t arena_bytes() const { return bytes_; }
    // Stage marks: launch index after which mark m was reached by every grid
    // (-1: not aligned), for m = 0..6.
    int mark_after(int m) const { return m >= 0 && m < 8 ? marks_[m] : -1; }

  private:
    struct Launch {
        int pos, kind, variant, argsz;
        int ngroups, total, nblocks;
        size_t args_off, sched_off;
    };
    std::vector<Launch> launches_;
    std::string layout_;               // signature of the launch sequence
    std::vector<unsigned char> image_;  // host copy of the arena
    unsigned char *pinned_[2] = {nullptr, nullptr};
    size_t pinned_cap_[2] = {0, 0};
    hipEvent_t copied_[2] = {nullptr, nullptr};
    int cur_ = 0;
    unsigned char *dev_ = nullptr;
    size_t dev_cap_ = 0, bytes_ = 0;
    int cus_ = 0;
    int marks_[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    bool dev_moved_ = false;
};

}  // namespace bh
