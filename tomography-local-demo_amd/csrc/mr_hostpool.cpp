// Host threads for the staging copies of streamed batches (see mr_hostpool.h).
#include "mr_hostpool.h"

#include <algorithm>
#include <cstring>

namespace mr {

HostCopyPool::HostCopyPool(int workers) {
    for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { run(); });
}

HostCopyPool::~HostCopyPool() {
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
}

void HostCopyPool::take_chunks() {
    for (;;) {
        const int i = next_.fetch_add(1, std::memory_order_relaxed);
        if (i >= nch_) return;
        const size_t o = static_cast<size_t>(i) * step_;
        std::memcpy(dst_ + o, src_ + o, std::min(step_, bytes_ - o));
        ready_[i].store(1, std::memory_order_release);
    }
}

void HostCopyPool::run() {
    unsigned long seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(m_);
            wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            ++acked_;
            ++busy_;
        }
        take_chunks();
        {
            std::lock_guard<std::mutex> lk(m_);
            --busy_;
        }
        idle_.notify_all();
    }
}

int HostCopyPool::copy(void* dst, const void* src, size_t bytes, int nch,
                       const std::function<int(size_t, size_t)>& on_chunk) {
    if (bytes == 0) return 0;
    nch = std::max(1, nch);
    const size_t step = std::max<size_t>(1, (bytes + nch - 1) / nch);
    nch = static_cast<int>((bytes + step - 1) / step);
    {
        std::lock_guard<std::mutex> lk(m_);
        dst_ = static_cast<char*>(dst);
        src_ = static_cast<const char*>(src);
        bytes_ = bytes;
        step_ = step;
        nch_ = nch;
        if (ready_len_ < nch) {
            ready_.reset(new std::atomic<int>[nch]);
            ready_len_ = nch;
        }
        for (int i = 0; i < nch; ++i) ready_[i].store(0, std::memory_order_relaxed);
        next_.store(0, std::memory_order_relaxed);
        acked_ = 0;
        ++gen_;
    }
    wake_.notify_all();
    int rc = 0;
    for (int i = 0; i < nch; ++i) {
        // help with the copying while chunk i is not complete
        while (!ready_[i].load(std::memory_order_acquire)) {
            const int j = next_.fetch_add(1, std::memory_order_relaxed);
            if (j < nch) {
                const size_t o = static_cast<size_t>(j) * step;
                std::memcpy(dst_ + o, src_ + o, std::min(step, bytes - o));
                ready_[j].store(1, std::memory_order_release);
            } else {
                std::this_thread::yield();
            }
        }
        if (!rc) rc = on_chunk(static_cast<size_t>(i) * step, std::min(step, bytes - static_cast<size_t>(i) * step));
    }
    // the job's fields stay until every worker has seen this generation and left it
    std::unique_lock<std::mutex> lk(m_);
    idle_.wait(lk, [&] { return acked_ == static_cast<int>(threads_.size()) && busy_ == 0; });
    return rc;
}

}  // namespace mr
