// Host-to-host copies between the caller's pageable arrays and the pinned
// staging of the host-array pipeline, split over a few worker threads: one
// thread moves roughly 5-10 GB/s, well short of what the PCIe DMA engines
// move to and from pinned memory.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include <unistd.h>

#include "mol.h"

namespace ct {
namespace {

// Below this many bytes the caller copies alone.
constexpr size_t kParallelBytes = size_t(1) << 21;

class CopyPool {
 public:
  explicit CopyPool(int workers) : workers_(workers), pid_(getpid()) {
    for (int i = 0; i < workers_; ++i) std::thread([this, i] { loop(i + 1); }).detach();
  }
  pid_t pid() const { return pid_; }

  // Copies the jobs: part 0 on the calling thread, parts 1..workers on the
  // pool; returns when every part is done.  One run at a time.
  void run(const HostCopy* jobs, int njobs, size_t total) {
    std::lock_guard<std::mutex> one(run_mu_);
    {
      std::lock_guard<std::mutex> g(mu_);
      jobs_ = jobs;
      njobs_ = njobs;
      total_ = total;
      pending_ = workers_;
      ++gen_;
    }
    cv_.notify_all();
    copy_part(0);
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [&] { return pending_ == 0; });
  }

 private:
  void loop(int part) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return gen_ != seen; });
        seen = gen_;
      }
      copy_part(part);
      std::lock_guard<std::mutex> g(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  // start of part p in the concatenation of the jobs (4 KiB-aligned cuts)
  size_t cut(int p) const {
    const int parts = workers_ + 1;
    return p >= parts ? total_ : (total_ / parts * p) & ~size_t(4095);
  }
  void copy_part(int p) const {
    const size_t a = cut(p), b = cut(p + 1);
    size_t off = 0;
    for (int j = 0; j < njobs_ && off < b; ++j) {
      const size_t lo = off, hi = off + jobs_[j].bytes;
      const size_t x = std::max(a, lo), y = std::min(b, hi);
      if (x < y)
        std::memcpy(static_cast<char*>(jobs_[j].dst) + (x - lo), static_cast<const char*>(jobs_[j].src) + (x - lo),
                    y - x);
      off = hi;
    }
  }

  const int workers_;
  const pid_t pid_;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  int pending_ = 0;
  const HostCopy* jobs_ = nullptr;
  int njobs_ = 0;
  size_t total_ = 0;
};

CopyPool* pool() {
  // process-wide and never destroyed (its detached workers outlive static
  // destruction); a forked child, which has none of the parent's threads,
  // starts its own
  static std::mutex mu;
  static CopyPool* p = nullptr;
  std::lock_guard<std::mutex> g(mu);
  if (p == nullptr || p->pid() != getpid()) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    p = new CopyPool((int)std::min(7u, std::max(1u, hw / 2)));
  }
  return p;
}

}  // namespace

void host_copy(const HostCopy* jobs, int njobs) {
  size_t total = 0;
  for (int j = 0; j < njobs; ++j) total += jobs[j].bytes;
  if (total < kParallelBytes) {
    for (int j = 0; j < njobs; ++j) std::memcpy(jobs[j].dst, jobs[j].src, jobs[j].bytes);
    return;
  }
  pool()->run(jobs, njobs, total);
}

}  // namespace ct
