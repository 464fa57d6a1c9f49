// bh_hostcopy.h -- whole-grid copies between pageable host arrays and the
// device for the host entry points (bh_solve, bh_set_potential), through a
// small pinned ring.
//
// A copy from pageable memory is staged by the runtime at about half the
// PCIe rate of pinned memory (137 MB: 4.8 ms in, 8.0 ms out, against 2.4 ms
// each way pinned, on MI355X).  Here a few worker threads each own a
// contiguous stripe of the array and two slots of the ring: a worker copies
// a chunk of its stripe into one slot with the CPU while the DMA engine moves
// the chunk in its other slot, so the CPU copies and the transfers overlap.
// The upload also checks every value for finiteness as it copies (the host
// entry point's input check, in the same pass over the array).
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace bh {

class HostStager {
public:
    // workers copy on device `device`; the ring is allocated on first use
    HostStager(int device, int workers);
    ~HostStager();
    HostStager(const HostStager &) = delete;
    HostStager &operator=(const HostStager &) = delete;

    // n doubles from host src to device dst, enqueued on stream s (the copies
    // are in flight on return: work enqueued on s after them sees dst).  With
    // check, *bad receives the index of the first non-finite value, or -1.
    // Returns a BH_* status.
    int to_device(double *dst, const double *src, size_t n, hipStream_t s, bool check, long *bad);
    // n doubles from device src (ready once the work enqueued on s so far
    // has run) to host dst; returns when dst holds them.
    int to_host(double *dst, const double *src, size_t n, hipStream_t s);

    // measured at 258 x 130 x 512 (137 MB each way): 2 or 3 slots of 4 or
    // 8 MB and 4 to 8 workers are alike (7.0-7.7 ms per bh_solve, against
    // 16.4 ms with the runtime's own staging); 1 worker 13 ms, 2 workers 8.4
    static constexpr size_t SLOT_BYTES = size_t(4) << 20;
    static constexpr int SLOTS = 2;  // per worker

private:
    int ensure_ring();
    // fn(w) on every worker w, the calling thread waiting
    void run(const std::function<void(int)> &fn);
    void worker_main(int w);

    int device_, nw_;
    char *ring_ = nullptr;         // pinned [worker][slot][SLOT_BYTES]
    std::vector<hipEvent_t> ev_;   // [worker][slot]: the slot's last transfer
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    const std::function<void(int)> *job_ = nullptr;
    unsigned gen_ = 0;
    int active_ = 0;
    bool stop_ = false;
};

// Worker threads for a stager: BH_COPY_THREADS, else a quarter of the host's
// hardware threads, 1..4.
int host_copy_workers();

}  // namespace bh
