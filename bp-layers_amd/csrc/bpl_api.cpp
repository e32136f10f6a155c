// Host-side plumbing of libbpl_amd: thread-local error strings, version and the optional
// per-kernel timing instrumentation.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "bpl_common.h"

namespace bpl {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return BPL_OK;
  set_error("%s: %s", what, hipGetErrorString(e));
  return BPL_E_HIP;
}

// ---------------------------------------------------------------- kernel timing
namespace {
struct Pending {
  int id;
  hipEvent_t start, stop;
};
struct TimingState {
  std::atomic<bool> on{false};
  std::mutex mu;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[BPL_KERNEL_COUNT] = {};
  int64_t launches[BPL_KERNEL_COUNT] = {};
};
TimingState &timing() {
  static TimingState *s = new TimingState();  // never destroyed: events may outlive exit order
  return *s;
}
hipEvent_t take_event(TimingState &t) {  // called with t.mu held
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
// Waits for and accumulates every pending pair (called with t.mu held).
void drain(TimingState &t) {
  for (const Pending &p : t.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      t.total_ms[p.id] += ms;
      t.launches[p.id] += 1;
    }
    t.pool.push_back(p.start);
    t.pool.push_back(p.stop);
  }
  t.pending.clear();
}
}  // namespace

KernelTimer::KernelTimer(int id, hipStream_t stream) : id_(id), stream_(stream) {
  TimingState &t = timing();
  if (!t.on.load(std::memory_order_relaxed)) return;
  std::lock_guard<std::mutex> lk(t.mu);
  start_ = take_event(t);
  stop_ = take_event(t);
  if (!start_ || !stop_ || hipEventRecord(start_, stream_) != hipSuccess) {
    if (start_) t.pool.push_back(start_);
    if (stop_) t.pool.push_back(stop_);
    start_ = stop_ = nullptr;
  }
}

KernelTimer::~KernelTimer() {
  if (!start_) return;
  TimingState &t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  if (hipEventRecord(stop_, stream_) == hipSuccess) {
    t.pending.push_back({id_, start_, stop_});
  } else {
    t.pool.push_back(start_);
    t.pool.push_back(stop_);
  }
}

}  // namespace bpl

using namespace bpl;

extern "C" int bpl_version(void) { return 1; }
extern "C" const char *bpl_last_error(void) { return bpl::g_err; }

extern "C" int bpl_timing_enable(int on) {
  timing().on.store(on != 0);
  return BPL_OK;
}

extern "C" int bpl_timing_reset(void) {
  TimingState &t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  drain(t);
  for (int i = 0; i < BPL_KERNEL_COUNT; ++i) {
    t.total_ms[i] = 0.0;
    t.launches[i] = 0;
  }
  return BPL_OK;
}

extern "C" int bpl_timing_read(int kernel, double *total_ms, int64_t *launches) {
  BPL_REQUIRE(kernel >= 0 && kernel < BPL_KERNEL_COUNT, BPL_E_ARG, "bpl_timing_read: unknown kernel %d", kernel);
  TimingState &t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  drain(t);
  if (total_ms) *total_ms = t.total_ms[kernel];
  if (launches) *launches = t.launches[kernel];
  return BPL_OK;
}

extern "C" const char *bpl_timing_name(int kernel) {
  static const char *names[BPL_KERNEL_COUNT] = {
      "lbp_fwd_horizontal", "lbp_fwd_vertical", "lbp_bwd_vertical", "lbp_bwd_horizontal",
      "lbp_bwd_combine",    "lbp_bwd_reduce",   "sad_stereo_fwd",   "sad_stereo_bwd",
      "sad_flow_fwd",       "sad_flow_bwd",     "softmax_labels",   "lbp_beliefs",
      "sad_stereo_softmax"};
  return kernel >= 0 && kernel < BPL_KERNEL_COUNT ? names[kernel] : "";
}
