// ============================================================================
// host_io.hip -- host-array transfers of the drop-in entry points (afmm_eval,
// afmm_eval_batch): the reference's seam takes and returns host vectors
// (ForceSolver3D.hpp:20-28), so every call moves the bodies over PCIe.
//
// A copy from pageable memory is staged by the runtime through its own small
// buffers with a single-threaded host memcpy. Here each context owns two
// pinned staging buffers (hipHostMalloc, grow-only): the host copies chunk k
// of the caller's array into one of them with several threads while the DMA
// of chunk k-1 runs from the other, and downloads run the same pipeline in
// reverse. The copies are byte-exact; only the time changes.
// ============================================================================
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include <unistd.h>

#include "afmm_internal.h"

namespace afmm {
namespace {

constexpr size_t kStageChunk = 8u << 20;    // bytes per pipeline stage
constexpr size_t kParallelMin = 1u << 20;   // below this a plain memcpy
constexpr int kCopyThreads = 8;             // host threads per copy (including the caller)

// A fixed pool of copy threads shared by all contexts; one copy at a time.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool pool;
        return pool;
    }
    void copy(void* dst, const void* src, size_t bytes) {
        // (a forked child has none of the parent's threads: it copies alone)
        if (bytes < kParallelMin || workers_.empty() || getpid() != owner_) {
            std::memcpy(dst, src, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(job_mu_);
        const int parts = (int)workers_.size() + 1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            part_ = ((bytes + parts - 1) / parts + 63) / 64 * 64;
            next_.store(0);
            pending_ = parts;
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }

private:
    CopyPool() : owner_(getpid()) {
        const unsigned hw = std::thread::hardware_concurrency();
        const int n = std::max(0, std::min<int>(kCopyThreads, hw ? (int)hw : 1) - 1);
        for (int i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
    }
    // Takes parts until none are left; the last finisher wakes the caller.
    void work() {
        for (;;) {
            const size_t k = next_.fetch_add(1);
            const size_t a = k * part_;
            if (a >= bytes_ || k >= (size_t)workers_.size() + 1) break;
            std::memcpy(dst_ + a, src_ + a, std::min(part_, bytes_ - a));
        }
        std::lock_guard<std::mutex> lk(mu_);
        if (--pending_ == 0) done_.notify_all();
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
        }
    }
    pid_t owner_;
    std::vector<std::thread> workers_;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, part_ = 0;
    std::atomic<size_t> next_{0};
    int pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace

HostStager::~HostStager() {
    for (int b = 0; b < 2; ++b) {
        if (ev_[b]) {
            (void)hipEventSynchronize(ev_[b]);
            (void)hipEventDestroy(ev_[b]);
        }
        if (pin_[b]) (void)hipHostFree(pin_[b]);
    }
}

void HostStager::reserve() {
    if (pin_[0]) return;
    for (int b = 0; b < 2; ++b) {
        AFMM_CHECK(hipHostMalloc(&pin_[b], kStageChunk, hipHostMallocDefault));
        AFMM_CHECK(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming));
    }
}

void HostStager::upload(void* d_dst, const void* h_src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    reserve();
    const char* src = static_cast<const char*>(h_src);
    char* dst = static_cast<char*>(d_dst);
    int b = 0;
    for (size_t off = 0; off < bytes; off += kStageChunk, b ^= 1) {
        const size_t len = std::min(kStageChunk, bytes - off);
        AFMM_CHECK(hipEventSynchronize(ev_[b]));  // the DMA out of this buffer two chunks ago is done
        CopyPool::get().copy(pin_[b], src + off, len);
        AFMM_CHECK(hipMemcpyAsync(dst + off, pin_[b], len, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipEventRecord(ev_[b], s));
    }
}

void HostStager::download(void* h_dst, const void* d_src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    reserve();
    const char* src = static_cast<const char*>(d_src);
    char* dst = static_cast<char*>(h_dst);
    const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
    auto enqueue = [&](size_t k) {
        const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
        AFMM_CHECK(hipMemcpyAsync(pin_[k & 1], src + off, len, hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipEventRecord(ev_[k & 1], s));
    };
    enqueue(0);
    for (size_t k = 0; k < nchunks; ++k) {
        // buffer (k+1)&1 last held chunk k-1, whose host copy finished last iteration
        if (k + 1 < nchunks) enqueue(k + 1);
        AFMM_CHECK(hipEventSynchronize(ev_[k & 1]));
        const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
        CopyPool::get().copy(dst + off, pin_[k & 1], len);
    }
}

}  // namespace afmm
