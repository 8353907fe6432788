// host_pipeline.hpp -- host-pointer calls staged through page-locked memory.
//
// The reference's host-pointer entries (fft_cffi.cpp:713-804) are called with ordinary
// (pageable) buffers. DMA engines need page-locked memory, and page-locking the caller's
// own buffers is not available everywhere, so these calls stage through bounce buffers
// allocated with hipHostMalloc: the caller's samples are copied into a pinned slot and
// uploaded, the planes are downloaded into a pinned slot and copied out. The copies on the
// host run on a small pool of persistent threads; the uploads, kernels and downloads run
// on three streams of the default priority, so chunk c's download overlaps chunk c+1's
// kernel and chunk c-2's copy out.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace fftfree {

void hip_check(hipError_t e, const char* what);

// Page-locked host allocation that only grows.
struct PinnedBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t want) {
    if (want <= bytes) return;
    release();
    hip_check(hipHostMalloc(&ptr, want, hipHostMallocDefault), "hipHostMalloc(staging)");
    bytes = want;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
  ~PinnedBuffer() { release(); }
};

// Persistent worker threads for large host copies. run() cuts the copies into pieces of
// at most kPiece bytes, which the workers and the calling thread take from a shared
// counter; it returns when every piece is done.
class CopyPool {
 public:
  struct Copy {
    void* dst;
    const void* src;
    size_t bytes;
  };
  static constexpr size_t kPiece = size_t(2) << 20;

  explicit CopyPool(int workers) {
    for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { worker(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  CopyPool(const CopyPool&) = delete;
  CopyPool& operator=(const CopyPool&) = delete;

  void run(const Copy* copies, size_t n) {
    std::vector<Copy> pieces;
    for (size_t i = 0; i < n; ++i)
      for (size_t off = 0; off < copies[i].bytes; off += kPiece)
        pieces.push_back({static_cast<char*>(copies[i].dst) + off, static_cast<const char*>(copies[i].src) + off,
                          std::min(kPiece, copies[i].bytes - off)});
    if (pieces.empty()) return;
    if (pieces.size() == 1 || threads_.empty()) {
      for (const Copy& p : pieces) std::memcpy(p.dst, p.src, p.bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      pieces_.swap(pieces);
      next_.store(0);
      busy_ = threads_.size();
      ++generation_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return busy_ == 0; });
  }

 private:
  void drain() {
    for (size_t i; (i = next_.fetch_add(1)) < pieces_.size();)
      std::memcpy(pieces_[i].dst, pieces_[i].src, pieces_[i].bytes);
  }
  void worker() {
    unsigned long long seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
      if (stop_) return;
      seen = generation_;
      lk.unlock();
      drain();
      lk.lock();
      if (--busy_ == 0) done_cv_.notify_all();
    }
  }

  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::vector<Copy> pieces_;
  std::atomic<size_t> next_{0};
  size_t busy_ = 0;
  unsigned long long generation_ = 0;
  bool stop_ = false;
};

}  // namespace fftfree
