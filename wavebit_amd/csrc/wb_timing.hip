// Per-kernel HIP-event timing registry (see wb_timing.h).
#include <mutex>
#include <vector>

#include "wb_timing.h"

bool wb_timing_on = false;
bool wb_timing_coarse = false;
int wb_timing_stage = 0;

namespace {
struct Pending {
  hipEvent_t a, b;
};
std::mutex g_mu;
std::vector<Pending> g_pending[WB_T_COUNT];
hipEvent_t g_open[WB_T_COUNT];
double g_total_ms[WB_T_COUNT];
double g_flops[WB_T_COUNT];
long long g_count[WB_T_COUNT];

void drain(int id) {
  for (auto& p : g_pending[id]) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      g_total_ms[id] += ms;
      g_count[id] += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending[id].clear();
}
}  // namespace

void wb_timer_begin(int id, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  g_open[id] = e;
}

void wb_timer_end(int id, hipStream_t s, double flops) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_flops[id] += flops;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  g_pending[id].push_back({g_open[id], e});
  if (g_pending[id].size() > 4096) drain(id);
}

extern "C" {

// on: 0 off, 1 per kernel, 2 per stage (eigensolve, S-matrices, assembly gathers)
int wb_timing_enable(int on) {
  if (on < 0 || on > 2) return 1;
  std::lock_guard<std::mutex> lk(g_mu);
  wb_timing_on = on != 0;
  wb_timing_coarse = on == 2;
  return 0;
}

// Total milliseconds and launch count of kernel `id` since the last reset (waits on events).
int wb_timing_read(int id, double* total_ms, long long* count, double* flops, int reset) {
  if (id < 0 || id >= WB_T_COUNT || !total_ms || !count) return 1;
  std::lock_guard<std::mutex> lk(g_mu);
  drain(id);
  *total_ms = g_total_ms[id];
  *count = g_count[id];
  if (flops) *flops = g_flops[id];
  if (reset) {
    g_total_ms[id] = 0.0;
    g_count[id] = 0;
    g_flops[id] = 0.0;
  }
  return 0;
}

}  // extern "C"
