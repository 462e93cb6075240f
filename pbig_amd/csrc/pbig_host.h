// Host-side helpers of the library's host entry points: a fixed pool of
// threads for the record packing and the copies of pairs into the caller's
// memory (page faults and memory bandwidth of a single thread bound those:
// about 10 GB/s into fresh pageable memory).
#pragma once

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace pbig {

// Runs one job at a time over all its threads: fn(t) for t in [0, size()),
// the calling thread taking t = 0; run() returns when every part is done.
class HostPool {
 public:
  explicit HostPool(unsigned n) : n_(n ? n : 1u) {
    for (unsigned t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    go_.notify_all();
    for (auto& t : th_) t.join();
  }
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

  unsigned size() const { return n_; }

  void run(const std::function<void(unsigned)>& fn) {
    if (n_ == 1) {
      fn(0);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      fn_ = &fn;
      left_ = n_ - 1;
      ++gen_;
    }
    go_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return left_ == 0; });
    fn_ = nullptr;
  }

  // Split [0, n) into size() contiguous parts; part t is [lo, hi).
  static void part(uint64_t n, unsigned parts, unsigned t, uint64_t* lo, uint64_t* hi) {
    const uint64_t q = n / parts, r = n % parts;
    *lo = t * q + (t < r ? t : r);
    *hi = *lo + q + (t < r ? 1u : 0u);
  }

 private:
  void loop(unsigned t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(unsigned)>* f;
      {
        std::unique_lock<std::mutex> lk(m_);
        go_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        f = fn_;
      }
      (*f)(t);
      std::lock_guard<std::mutex> lk(m_);
      if (--left_ == 0) done_.notify_one();
    }
  }

  unsigned n_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable go_, done_;
  const std::function<void(unsigned)>* fn_ = nullptr;
  uint64_t gen_ = 0;
  unsigned left_ = 0;
  bool stop_ = false;
};

}  // namespace pbig
