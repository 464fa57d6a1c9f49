// bh_hostcopy.cpp -- pipelined host <-> device copies of whole grids for the
// host entry points (bh_hostcopy.h).
#include "bh_hostcopy.h"

#include "../../include/bh_potential.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace bh {

int host_copy_workers()
{
    if (const char *e = std::getenv("BH_COPY_THREADS")) {
        const int v = std::atoi(e);
        if (v > 0) return std::min(v, 16);
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1, std::min(4, (int)(hw / 4)));
}

HostStager::HostStager(int device, int workers) : device_(device), nw_(std::max(1, workers))
{
    for (int w = 0; w < nw_; ++w) threads_.emplace_back([this, w] { worker_main(w); });
}

HostStager::~HostStager()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    go_.notify_all();
    for (auto &t : threads_) t.join();
    for (auto &e : ev_) {
        (void)hipEventSynchronize(e);
        (void)hipEventDestroy(e);
    }
    if (ring_) (void)hipHostFree(ring_);
}

int HostStager::ensure_ring()
{
    if (ring_) return BH_OK;
    void *r = nullptr;
    if (hipHostMalloc(&r, (size_t)nw_ * SLOTS * SLOT_BYTES, hipHostMallocDefault) != hipSuccess) return BH_E_DEVICE;
    ring_ = (char *)r;
    ev_.assign((size_t)nw_ * SLOTS, nullptr);
    for (auto &e : ev_)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return BH_E_DEVICE;
    // every slot's event recorded once, so that a wait before a slot's
    // first use returns at once
    for (auto &e : ev_)
        if (hipEventRecord(e, 0) != hipSuccess) return BH_E_DEVICE;
    return BH_OK;
}

void HostStager::run(const std::function<void(int)> &fn)
{
    std::unique_lock<std::mutex> lk(mu_);
    job_ = &fn;
    active_ = nw_;
    ++gen_;
    go_.notify_all();
    done_.wait(lk, [&] { return active_ == 0; });
    job_ = nullptr;
}

void HostStager::worker_main(int w)
{
    (void)hipSetDevice(device_);
    unsigned seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        go_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        const std::function<void(int)> *fn = job_;
        lk.unlock();
        (*fn)(w);
        lk.lock();
        if (--active_ == 0) done_.notify_one();
    }
}

namespace {
// index of the first non-finite value of x[0, n), or -1
long first_nonfinite(const double *x, size_t n)
{
    const uint64_t EXP = 0x7ff0000000000000ull;
    const size_t B = 4096;
    for (size_t b = 0; b < n; b += B) {
        const size_t e = std::min(n, b + B);
        uint64_t any = 0;
        for (size_t i = b; i < e; ++i) {
            uint64_t u;
            std::memcpy(&u, x + i, 8);
            any |= (uint64_t)((u & EXP) == EXP);
        }
        if (any)
            for (size_t i = b; i < e; ++i) {
                uint64_t u;
                std::memcpy(&u, x + i, 8);
                if ((u & EXP) == EXP) return (long)i;
            }
    }
    return -1;
}
}  // namespace

int HostStager::to_device(double *dst, const double *src, size_t n, hipStream_t s, bool check, long *bad)
{
    if (bad) *bad = -1;
    int rc = ensure_ring();
    if (rc != BH_OK) return rc;
    const size_t per = SLOT_BYTES / sizeof(double);
    std::atomic<int> err{BH_OK};
    std::atomic<long> first{LONG_MAX};
    std::function<void(int)> job = [&](int w) {
        const size_t lo = n * w / nw_, hi = n * (w + 1) / nw_;
        int c = 0;
        for (size_t off = lo; off < hi; off += per, ++c) {
            const size_t cnt = std::min(per, hi - off);
            const int k = w * SLOTS + c % SLOTS;
            double *buf = (double *)(ring_ + (size_t)k * SLOT_BYTES);
            // the slot's previous transfer (this call's or an earlier one's)
            if (hipEventSynchronize(ev_[k]) != hipSuccess) { err = BH_E_DEVICE; return; }
            std::memcpy(buf, src + off, cnt * sizeof(double));
            if (check && first.load(std::memory_order_relaxed) > (long)off) {
                const long i = first_nonfinite(buf, cnt);
                if (i >= 0) {
                    long cur = first.load();
                    while ((long)off + i < cur && !first.compare_exchange_weak(cur, (long)off + i)) {
                    }
                }
            }
            if (hipMemcpyAsync(dst + off, buf, cnt * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipEventRecord(ev_[k], s) != hipSuccess) {
                err = BH_E_DEVICE;
                return;
            }
        }
    };
    run(job);
    if (bad && first.load() != LONG_MAX) *bad = first.load();
    return err.load();
}

int HostStager::to_host(double *dst, const double *src, size_t n, hipStream_t s)
{
    int rc = ensure_ring();
    if (rc != BH_OK) return rc;
    const size_t per = SLOT_BYTES / sizeof(double);
    std::atomic<int> err{BH_OK};
    std::function<void(int)> job = [&](int w) {
        const size_t lo = n * w / nw_, hi = n * (w + 1) / nw_;
        if (hi <= lo) return;
        const int nc = (int)((hi - lo + per - 1) / per);
        auto slot = [&](int c) { return w * SLOTS + c % SLOTS; };
        auto enqueue = [&](int c) {
            const size_t off = lo + (size_t)c * per, cnt = std::min(per, hi - off);
            const int k = slot(c);
            return hipMemcpyAsync(ring_ + (size_t)k * SLOT_BYTES, src + off, cnt * sizeof(double),
                                  hipMemcpyDeviceToHost, s) == hipSuccess &&
                   hipEventRecord(ev_[k], s) == hipSuccess;
        };
        for (int c = 0; c < std::min(nc, SLOTS); ++c)
            if (hipEventSynchronize(ev_[slot(c)]) != hipSuccess || !enqueue(c)) { err = BH_E_DEVICE; return; }
        for (int c = 0; c < nc; ++c) {
            const size_t off = lo + (size_t)c * per, cnt = std::min(per, hi - off);
            const int k = slot(c);
            if (hipEventSynchronize(ev_[k]) != hipSuccess) { err = BH_E_DEVICE; return; }
            std::memcpy(dst + off, ring_ + (size_t)k * SLOT_BYTES, cnt * sizeof(double));
            if (c + SLOTS < nc && !enqueue(c + SLOTS)) { err = BH_E_DEVICE; return; }
        }
    };
    run(job);
    return err.load();
}

}  // namespace bh
