// Device-to-host copies into pageable memory for the host-array entry points.
//
// The host-array API (the reference's own boundary: numpy arrays in and out)
// ends with the cc rows crossing PCIe into memory the caller has just
// allocated.  A plain hipMemcpy into such memory takes every first-touch page
// fault of the destination inside the runtime's copy thread.  Measured on an
// MI355X host with 16 usable CPUs, 4 GiB copies: hipMemcpy into fresh
// pageable memory 3.9-4.4 GB/s, into already-touched pageable memory 18 GB/s;
// first touch of fresh pages by 16 threads 69 GB/s; DMA into pinned memory
// 57 GB/s; pinned -> touched pageable memcpy by 16 threads 105 GB/s.
//
// copy_to_host therefore
//   1. faults the destination's pages in with a team of host threads
//      (MADV_POPULATE_WRITE where the kernel has it, else a read-write touch of
//      one byte per page that keeps the contents), while the GPU still runs
//      the work queued before the copy;
//   2. streams the data through two pinned staging buffers: the DMA of chunk
//      i + 1 (on the caller's stream) runs while the team copies chunk i into
//      the destination.
// Copies below kStagedMin bytes, or when pinned memory or threads are not
// available, take a plain hipMemcpyAsync and a stream synchronisation.
#include <sched.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "fmf_internal.h"

#ifndef MADV_POPULATE_WRITE
#define MADV_POPULATE_WRITE 23
#endif

namespace fmf {

namespace {

constexpr size_t kStage = 64u << 20;      // bytes per staging buffer
constexpr size_t kStagedMin = 32u << 20;  // smaller copies go straight
constexpr int kMaxTeam = 16;

// Pinned staging pairs, reused across calls (each concurrent copy takes its
// own pair); the library keeps them until the process ends.
std::mutex g_stage_mu;
std::vector<char *> g_stage_free;

char *take_stage() {
  {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!g_stage_free.empty()) {
      char *p = g_stage_free.back();
      g_stage_free.pop_back();
      return p;
    }
  }
  void *p = nullptr;
  if (hipHostMalloc(&p, 2 * kStage, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return static_cast<char *>(p);
}

void give_stage(char *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_stage_mu);
  g_stage_free.push_back(p);
}

int team_size() {
  cpu_set_t set;
  int n = 1;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
  return std::max(1, std::min(n, kMaxTeam));
}

// A fork-join team for the duration of one copy: run(f) calls f(j, n) on
// every member j in [0, n) (member 0 is the calling thread) and returns when
// all have finished.  The caller waits for every worker to finish one job
// before it posts the next, so each worker runs each job exactly once.
class Team {
 public:
  explicit Team(int n) {
    for (int j = 1; j < n; j++) {
      try {
        th_.emplace_back([this, j] { work(j); });
      } catch (...) {  // fewer threads than asked for: run with those
        break;
      }
    }
    n_ = (int)th_.size() + 1;
  }
  ~Team() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  int size() const { return n_; }
  void run(const std::function<void(int, int)> &f) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = f;
      done_ = 0;
      ++gen_;
    }
    cv_.notify_all();
    f(0, n_);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_ == n_ - 1; });
  }

 private:
  void work(int j) {
    long seen = 0;
    for (;;) {
      std::function<void(int, int)> job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      job(j, n_);
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (++done_ == n_ - 1) done_cv_.notify_one();
      }
    }
  }
  std::vector<std::thread> th_;
  int n_ = 1;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::function<void(int, int)> job_;
  long gen_ = 0;
  int done_ = 0;
  bool stop_ = false;
};

// Member j's share [lo, hi) of n bytes split over k members, on 4 KiB
// boundaries where possible.
void share(size_t n, int j, int k, size_t &lo, size_t &hi) {
  const size_t per = ((n + k - 1) / k + 4095) & ~(size_t)4095;
  lo = std::min(n, per * j);
  hi = std::min(n, lo + per);
}

void populate(char *p, size_t n) {
  if (n == 0) return;
  const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
  const uintptr_t a = (uintptr_t)p & ~(uintptr_t)(pg - 1);
  const uintptr_t b = ((uintptr_t)p + n + pg - 1) & ~(uintptr_t)(pg - 1);
  if (madvise((void *)a, b - a, MADV_POPULATE_WRITE) == 0) return;
  // Older kernels: touch one byte per page, writing back what it holds.
  volatile char *v = p;
  const size_t first = (pg - ((uintptr_t)p & (pg - 1))) & (pg - 1);
  if (first) v[0] = v[0];
  for (size_t k = first; k < n; k += pg) v[k] = v[k];
}

}  // namespace

int copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return kOk;
  char *stage = bytes >= kStagedMin ? take_stage() : nullptr;
  if (!stage) {
    FMF_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    FMF_HIP_CHECK(hipStreamSynchronize(s));
    return kOk;
  }
  struct StageHolder {
    char *p;
    ~StageHolder() { give_stage(p); }
  } hold{stage};
  hipEvent_t ev[2] = {nullptr, nullptr};
  struct EvHolder {
    hipEvent_t *e;
    ~EvHolder() {
      for (int b = 0; b < 2; b++)
        if (e[b]) (void)hipEventDestroy(e[b]);
    }
  } evh{ev};
  for (int b = 0; b < 2; b++)
    FMF_HIP_CHECK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));

  char *out = static_cast<char *>(dst);
  const char *in = static_cast<const char *>(src);
  const size_t n_chunks = (bytes + kStage - 1) / kStage;
  auto chunk_len = [&](size_t i) { return std::min(kStage, bytes - i * kStage); };
  auto issue = [&](size_t i) -> int {
    char *buf = stage + (i & 1) * kStage;
    FMF_HIP_CHECK(hipMemcpyAsync(buf, in + i * kStage, chunk_len(i), hipMemcpyDeviceToHost, s));
    FMF_HIP_CHECK(hipEventRecord(ev[i & 1], s));
    return kOk;
  };
  Team team(team_size());
  auto body = [&]() -> int {
    // The first DMA is queued behind the caller's work; the pages are
    // faulted in meanwhile.
    int r = issue(0);
    if (r) return r;
    team.run([&](int j, int k) {
      size_t lo, hi;
      share(bytes, j, k, lo, hi);
      populate(out + lo, hi - lo);
    });
    for (size_t i = 0; i < n_chunks; i++) {
      if (i + 1 < n_chunks && (r = issue(i + 1))) return r;
      FMF_HIP_CHECK(hipEventSynchronize(ev[i & 1]));
      const char *buf = stage + (i & 1) * kStage;
      char *to = out + i * kStage;
      const size_t len = chunk_len(i);
      team.run([&](int j, int k) {
        size_t lo, hi;
        share(len, j, k, lo, hi);
        if (hi > lo) memcpy(to + lo, buf + lo, hi - lo);
      });
    }
    return kOk;
  };
  const int rc = body();
  // On a failure, DMAs already queued may still write the staging buffers:
  // let them finish before the buffers go back to the pool.
  if (rc) (void)hipStreamSynchronize(s);
  return rc;
}

}  // namespace fmf
