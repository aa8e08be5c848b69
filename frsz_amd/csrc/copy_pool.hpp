// Host-side memcpy split over a fixed set of worker threads plus the caller
// (the CPU staging copies of the pipelined host-memory entry points).  Plain
// C++ with no HIP dependency, so it can be built and checked on its own.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace frsz2 {

// copies below this size run on the caller alone
constexpr std::uint64_t PARALLEL_COPY_MIN = std::uint64_t{ 4 } << 20;

// One copy at a time (the owner serialises its calls); the parts of a copy
// run on the workers and the caller concurrently.
class copy_pool
{
public:
  explicit copy_pool(int workers)
  {
    for (int i = 0; i < workers; ++i) {
      threads_.emplace_back([this] { work(); });
    }
  }
  ~copy_pool()
  {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) {
      t.join();
    }
  }

  void copy(void* dst, const void* src, std::uint64_t bytes)
  {
    if (bytes < PARALLEL_COPY_MIN || threads_.empty()) {
      std::memcpy(dst, src, bytes);
      return;
    }
    const std::uint64_t parts = threads_.size() + 1;
    const std::uint64_t part = ((bytes + parts - 1) / parts + 4095) & ~std::uint64_t{ 4095 };
    {
      std::lock_guard<std::mutex> g(mu_);
      dst_ = static_cast<std::uint8_t*>(dst);
      src_ = static_cast<const std::uint8_t*>(src);
      bytes_ = bytes;
      part_ = part;
      next_ = 1; // part 0 is the caller's
      pending_ = (bytes + part - 1) / part;
    }
    cv_.notify_all();
    run_part(0);
    std::unique_lock<std::mutex> lk(mu_);
    --pending_;
    while (true) {
      if (next_ * part_ < bytes_) {
        const std::uint64_t p = next_++;
        lk.unlock();
        run_part(p);
        lk.lock();
        --pending_;
        continue;
      }
      if (pending_ == 0) {
        break;
      }
      done_.wait(lk);
    }
    bytes_ = 0;
  }

private:
  void run_part(std::uint64_t p)
  {
    const std::uint64_t lo = p * part_;
    const std::uint64_t hi = std::min(bytes_, lo + part_);
    std::memcpy(dst_ + lo, src_ + lo, hi - lo);
  }

  void work()
  {
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
      cv_.wait(lk, [this] { return stop_ || next_ * part_ < bytes_; });
      if (stop_) {
        return;
      }
      const std::uint64_t p = next_++;
      lk.unlock();
      run_part(p);
      lk.lock();
      if (--pending_ == 0) {
        done_.notify_all();
      }
    }
  }

  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::uint8_t* dst_ = nullptr;
  const std::uint8_t* src_ = nullptr;
  std::uint64_t bytes_ = 0, part_ = 1, next_ = 0, pending_ = 0;
  bool stop_ = false;
};


} // namespace frsz2
