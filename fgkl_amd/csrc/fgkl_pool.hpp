// Process-wide pools of reusable device batches for the per-call entry points.
//
// A per-call entry checks a batch out for the duration of one call and returns it afterwards, so
// consecutive calls reuse its device memory, stream and events (the reference keeps one runner
// per thread for the same reason, kernel.rs:144-191).  Nothing is tied to the calling thread: a
// thread that exits holds no batch, and no destructor makes HIP calls at thread exit, when the
// HIP runtime's per-thread state may already be gone.  The pools themselves are never destroyed
// (a HIP call from a static destructor at process exit has the same problem); their idle batches
// can be released explicitly, and the driver reclaims the rest when the process ends.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace fgkl {

template <typename T>
class Pool {
 public:
  // An idle object, or null when none is idle (the caller creates one).
  std::unique_ptr<T> take() {
    std::lock_guard<std::mutex> lk(mu_);
    ++out_;
    if (idle_.empty()) return nullptr;
    std::unique_ptr<T> p = std::move(idle_.back());
    idle_.pop_back();
    ++reused_;
    return p;
  }
  // Returns an object checked out with take() (or created because none was idle); null only
  // ends the checkout (the object was destroyed after a device error).
  void give(std::unique_ptr<T> p) {
    std::lock_guard<std::mutex> lk(mu_);
    --out_;
    if (p) idle_.push_back(std::move(p));
  }
  // Destroys every idle object on the calling thread; returns how many.
  int64_t release_idle() {
    std::vector<std::unique_ptr<T>> drop;
    {
      std::lock_guard<std::mutex> lk(mu_);
      drop.swap(idle_);
    }
    return int64_t(drop.size());  // destroyed here, outside the lock
  }
  // Idle objects, checkouts not yet returned, and checkouts served by an idle object.
  void stats(int64_t* idle, int64_t* out, int64_t* reused) {
    std::lock_guard<std::mutex> lk(mu_);
    if (idle) *idle = int64_t(idle_.size());
    if (out) *out = out_;
    if (reused) *reused = reused_;
  }

 private:
  std::mutex mu_;
  std::vector<std::unique_ptr<T>> idle_;
  int64_t out_ = 0;
  int64_t reused_ = 0;
};

// Checks an object out of a pool for one scope and returns it when the scope ends.  `drop()`
// destroys it instead (after a device error, so a later call does not inherit its state).
template <typename T>
class Checkout {
 public:
  explicit Checkout(Pool<T>& pool) : pool_(pool), obj_(pool.take()) {}
  ~Checkout() { pool_.give(std::move(obj_)); }
  Checkout(const Checkout&) = delete;
  Checkout& operator=(const Checkout&) = delete;
  std::unique_ptr<T>& get() { return obj_; }
  void drop() { obj_.reset(); }

 private:
  Pool<T>& pool_;
  std::unique_ptr<T> obj_;
};

}  // namespace fgkl
