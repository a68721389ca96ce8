// Sampled kernel timing of the Lanczos loop, for the live roofline of a
// running workload.
//
// tnx_profile_begin(every) arms sampling: in every `every`-th Lanczos call
// made afterwards (on any host thread), the launches of the first Krylov
// step (every solve still active, so the algorithmic work of each launch is
// the full work its descriptors describe) are timed by the kernels
// themselves: each block stamps its start and end with the device's
// constant wall clock into the launch's slot (tnx_stamp_begin / _end), so
// a launch's time is the span from its first block's start to its last
// block's end, the kernel duration a profiler's trace reports.
// tnx_profile_end returns, per kernel class, the launches made while armed
// with their summed algorithmic work (flops or bytes), and the number of
// timed launches, their summed duration and their summed algorithmic work.
// Nothing is timed while sampling is not armed; when armed, a timed launch
// costs two atomics per block.
#include <string.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "tnx_common.h"

namespace {

struct Sample {
  int cls;
  double work;
};

const char* const kClassName[TNX_PROF_NCLASS] = {
    "gemm f64 matvec T = F x",      "gemm f64 matvec w = U R^T",
    "gemm f32 matvec T = F x",      "gemm f32 matvec w = U R^T",
    "gemm bf16 matvec T = F x",     "gemm bf16 matvec w = U R^T",
    "mixer matvec",                 "k_dots",
    "k_update_dots",                "k_update_norm",
};
const int kClassUnit[TNX_PROF_NCLASS] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
constexpr int kSlotWords = 2 * TNX_TS_PAIRS;

std::mutex g_mu;
std::atomic<int> g_armed{0};    // launches are counted while armed
std::atomic<long long> g_count[TNX_PROF_NCLASS];
std::atomic<double> g_work_all[TNX_PROF_NCLASS];   // algorithmic work of every counted launch
int g_every = 0;                // 0: not armed
long long g_calls = 0;          // Lanczos calls seen while armed
std::vector<Sample> g_samples;  // launches timed
int g_cap = 0;                  // stamp slots
unsigned long long* g_ts = nullptr;   // device: g_cap slots of kSlotWords
thread_local bool t_on = false; // the current thread's launches are sampled

}  // namespace

bool tnx_prof_call_begin() {
  t_on = false;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_every <= 0) return false;
  return (g_calls++ % g_every) == 0;
}

void tnx_prof_set(bool on) { t_on = on; }

bool tnx_prof_on() { return t_on; }

void tnx_prof_count(int cls, double work) {
  if (!g_armed.load(std::memory_order_relaxed)) return;
  g_count[cls].fetch_add(1, std::memory_order_relaxed);
  double cur = g_work_all[cls].load(std::memory_order_relaxed);
  while (!g_work_all[cls].compare_exchange_weak(cur, cur + work, std::memory_order_relaxed)) {
  }
}

int tnx_prof_mark(int cls, hipStream_t s) {
  (void)s;
  if (!t_on) return -1;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ts == nullptr || (int)g_samples.size() >= g_cap) return -1;   // slots spent
  g_samples.push_back(Sample{cls, 0.0});
  return (int)g_samples.size() - 1;
}

unsigned long long* tnx_prof_ts(int mark) {
  if (mark < 0) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_ts == nullptr ? nullptr : g_ts + (size_t)mark * kSlotWords;
}

void tnx_prof_done(int mark, hipStream_t s, double work) {
  (void)s;
  if (mark < 0) return;
  std::lock_guard<std::mutex> lock(g_mu);
  g_samples[mark].work = work;
}

extern "C" int tnx_profile_begin(int every, int max_launches) {
  if (every < 1 || max_launches < 1) {
    tnx_set_error("tnx_profile_begin: bad arguments (every=%d max_launches=%d)", every,
                  max_launches);
    return TNX_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_every > 0) {
    tnx_set_error("tnx_profile_begin: already armed");
    return TNX_EINVAL;
  }
  // stamp slots: starts at the largest clock value, ends at zero
  std::vector<unsigned long long> init((size_t)max_launches * kSlotWords);
  for (size_t i = 0; i < init.size(); i += 2) {
    init[i] = ~0ull;
    init[i + 1] = 0ull;
  }
  TNX_CHECK_HIP(hipMalloc((void**)&g_ts, init.size() * sizeof(unsigned long long)));
  if (hipMemcpy(g_ts, init.data(), init.size() * sizeof(unsigned long long),
                hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(g_ts);
    g_ts = nullptr;
    tnx_set_error("tnx_profile_begin: slot upload failed");
    return TNX_EHIP;
  }
  g_cap = max_launches;
  g_samples.clear();
  g_calls = 0;
  g_every = every;
  for (int c = 0; c < TNX_PROF_NCLASS; ++c) {
    g_count[c].store(0);
    g_work_all[c].store(0.0);
  }
  g_armed.store(1);
  return TNX_OK;
}

extern "C" int tnx_profile_end(tnx_prof_stat* out, int cap) {
  std::vector<Sample> samples;
  unsigned long long* ts = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_every <= 0) {
      tnx_set_error("tnx_profile_end: not armed");
      return TNX_EINVAL;
    }
    g_every = 0;
    g_armed.store(0);
    samples.swap(g_samples);
    ts = g_ts;
    g_ts = nullptr;
    g_cap = 0;
  }
  TNX_CHECK_HIP(hipDeviceSynchronize());
  std::vector<unsigned long long> stamps(samples.size() * kSlotWords);
  int dev = 0, khz = 0;
  const bool ok = hipMemcpy(stamps.data(), ts, stamps.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost) == hipSuccess &&
                  hipGetDevice(&dev) == hipSuccess &&
                  hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess &&
                  khz > 0;
  (void)hipFree(ts);
  tnx_prof_stat st[TNX_PROF_NCLASS];
  for (int c = 0; c < TNX_PROF_NCLASS; ++c) {
    memset(&st[c], 0, sizeof(st[c]));
    snprintf(st[c].name, sizeof(st[c].name), "%s", kClassName[c]);
    st[c].unit = kClassUnit[c];
    st[c].launches = g_count[c].load();
    st[c].work_all = g_work_all[c].load();
  }
  if (!ok) {
    tnx_set_error("tnx_profile_end: stamps not read back");
    return TNX_EHIP;
  }
  for (size_t k = 0; k < samples.size(); ++k) {
    const unsigned long long* w = stamps.data() + k * kSlotWords;
    unsigned long long a = ~0ull, b = 0ull;
    for (int p = 0; p < TNX_TS_PAIRS; ++p) {
      if (w[2 * p] < a) a = w[2 * p];
      if (w[2 * p + 1] > b) b = w[2 * p + 1];
    }
    if (a == ~0ull || b <= a) continue;   // no block stamped (every solve masked)
    const Sample& smp = samples[k];
    st[smp.cls].samples += 1;
    st[smp.cls].seconds += (double)(b - a) / (1e3 * (double)khz);
    st[smp.cls].work += smp.work;
  }
  const int n = cap < TNX_PROF_NCLASS ? cap : TNX_PROF_NCLASS;
  if (out != nullptr)
    for (int c = 0; c < n; ++c) out[c] = st[c];
  return TNX_PROF_NCLASS;
}
