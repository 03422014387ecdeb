// Host threads for the staging copies of streamed batches.  Plain C++ (no HIP), so that
// it also builds with g++ and the sanitizers on its own.
#pragma once

#include <stddef.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace mr {

// A chunked memcpy shared by `workers` threads and the calling thread.  copy() hands
// chunks 0, 1, ... in order to a callback on the calling thread as soon as each is
// complete, so that the caller can send a chunk on while the later ones are still being
// copied.  One copy() at a time (the caller serialises them).
class HostCopyPool {
public:
    explicit HostCopyPool(int workers);
    ~HostCopyPool();
    HostCopyPool(const HostCopyPool&) = delete;
    HostCopyPool& operator=(const HostCopyPool&) = delete;

    // Copies src[0, bytes) to dst in nch chunks (at least 1); on_chunk(offset, length) runs
    // for every chunk in order.  Returns the first non-zero on_chunk result (chunks after
    // it are still copied, not handed on), else 0.  Every worker has left the job when it
    // returns.
    int copy(void* dst, const void* src, size_t bytes, int nch,
             const std::function<int(size_t, size_t)>& on_chunk);

    int workers() const { return static_cast<int>(threads_.size()); }

private:
    void run();                       // worker loop
    void take_chunks();               // copies chunks until none is left

    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable wake_, idle_;
    bool stop_ = false;
    unsigned long gen_ = 0;           // job generation
    int busy_ = 0;                    // workers inside the current job
    int acked_ = 0;                   // workers that have seen the current generation
    // the current job
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, step_ = 0;
    int nch_ = 0;
    std::atomic<int> next_{0};
    std::unique_ptr<std::atomic<int>[]> ready_;
    int ready_len_ = 0;
};

}  // namespace mr
