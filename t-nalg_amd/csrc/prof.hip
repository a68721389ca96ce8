// Sampled kernel timing of the Lanczos loop, for the live roofline of a
// running workload.
//
// tnx_profile_begin(every) arms sampling: in every `every`-th Lanczos call
// made afterwards (on any host thread), the launches of the first Krylov
// step (every solve still active, so the algorithmic work of each launch is
// the full work its descriptors describe) are each bracketed by two HIP
// events on the launch's stream.  tnx_profile_end resolves the events and
// returns, per kernel class, the launches made while armed with their summed
// algorithmic work (flops or bytes), and the number of timed launches, their
// summed duration and their summed algorithmic work.  Nothing is
// timed while sampling is not armed; the cost when armed is two event
// records per timed launch.
#include <string.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "tnx_common.h"

namespace {

struct Sample {
  int cls;
  hipEvent_t a, b;
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

std::mutex g_mu;
std::atomic<int> g_armed{0};    // launches are counted while armed
std::atomic<long long> g_count[TNX_PROF_NCLASS];
std::atomic<double> g_work_all[TNX_PROF_NCLASS];   // algorithmic work of every counted launch
int g_every = 0;                // 0: not armed
long long g_calls = 0;          // Lanczos calls seen while armed
std::vector<Sample> g_samples;  // launches timed
std::vector<hipEvent_t> g_pool; // events not yet used
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
  if (!t_on) return -1;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_pool.size() < 2) return -1;   // pool spent: stop timing
  Sample smp;
  smp.cls = cls;
  smp.b = g_pool.back();
  g_pool.pop_back();
  smp.a = g_pool.back();
  g_pool.pop_back();
  smp.work = 0.0;
  if (hipEventRecord(smp.a, s) != hipSuccess) return -1;
  g_samples.push_back(smp);
  return (int)g_samples.size() - 1;
}

void tnx_prof_done(int mark, hipStream_t s, double work) {
  if (mark < 0) return;
  std::lock_guard<std::mutex> lock(g_mu);
  g_samples[mark].work = work;
  (void)hipEventRecord(g_samples[mark].b, s);
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
  for (int i = 0; i < 2 * max_launches; ++i) {
    hipEvent_t e;
    TNX_CHECK_HIP(hipEventCreate(&e));
    g_pool.push_back(e);
  }
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
  std::vector<hipEvent_t> pool;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_every <= 0) {
      tnx_set_error("tnx_profile_end: not armed");
      return TNX_EINVAL;
    }
    g_every = 0;
    g_armed.store(0);
    samples.swap(g_samples);
    pool.swap(g_pool);
  }
  TNX_CHECK_HIP(hipDeviceSynchronize());
  tnx_prof_stat st[TNX_PROF_NCLASS];
  for (int c = 0; c < TNX_PROF_NCLASS; ++c) {
    memset(&st[c], 0, sizeof(st[c]));
    snprintf(st[c].name, sizeof(st[c].name), "%s", kClassName[c]);
    st[c].unit = kClassUnit[c];
    st[c].launches = g_count[c].load();
    st[c].work_all = g_work_all[c].load();
  }
  int rc = TNX_OK;
  for (const Sample& smp : samples) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, smp.a, smp.b) == hipSuccess) {
      st[smp.cls].samples += 1;
      st[smp.cls].seconds += 1e-3 * (double)ms;
      st[smp.cls].work += smp.work;
    } else {
      rc = TNX_EHIP;
    }
    (void)hipEventDestroy(smp.a);
    (void)hipEventDestroy(smp.b);
  }
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  const int n = cap < TNX_PROF_NCLASS ? cap : TNX_PROF_NCLASS;
  if (out != nullptr)
    for (int c = 0; c < n; ++c) out[c] = st[c];
  if (rc) {
    tnx_set_error("tnx_profile_end: an event was not resolved");
    return rc;
  }
  return TNX_PROF_NCLASS;
}
