// Optional per-kernel timing: when enabled (wb_timing_enable), selected launches are bracketed
// by hipEvent records on the stream they are launched on; wb_timing_read sums the elapsed
// times per kernel id. Disabled (the default), the hooks cost one branch.
#pragma once
#include <hip/hip_runtime.h>

enum {
  WB_T_ZGEMM = 0, WB_T_GETRF_PANEL, WB_T_LASWP_TRSM, WB_T_TRSM_DIAG, WB_T_GEHRD, WB_T_UNGHR,
  WB_T_HQR, WB_T_TREVC, WB_T_COLNORM, WB_T_ASSEMBLE, WB_T_COUNT
};

extern bool wb_timing_on;
extern int wb_timing_stage;   // > 0 while a multi-launch stage is being timed as one entry
void wb_timer_begin(int id, hipStream_t s);
void wb_timer_end(int id, hipStream_t s, double flops);

// `flops` is the launch's algorithmic real FP64 flop count (0 where none is defined). Inside a
// stage (WbStageTimer scope) the launches are not timed on their own.
#define WB_TIMED_F(id, stream, flops, ...)                                           \
  do {                                                                               \
    const bool timed__ = wb_timing_on && wb_timing_stage == 0;                       \
    if (timed__) wb_timer_begin((id), (stream));                                     \
    __VA_ARGS__;                                                                     \
    if (timed__) wb_timer_end((id), (stream), (flops));                              \
  } while (0)
// A stage of several launches timed as one entry of `id` with `flops` algorithmic flops, for
// the scope of the object (early returns included).
struct WbStageTimer {
  int id;
  hipStream_t stream;
  double flops;
  bool timed;
  WbStageTimer(int id_, hipStream_t s_, double flops_) : id(id_), stream(s_), flops(flops_) {
    timed = wb_timing_on && wb_timing_stage == 0;
    if (timed) wb_timer_begin(id, stream);
    ++wb_timing_stage;
  }
  ~WbStageTimer() {
    --wb_timing_stage;
    if (timed) wb_timer_end(id, stream, flops);
  }
};
#define WB_TIMED(id, stream, ...) WB_TIMED_F(id, stream, 0.0, __VA_ARGS__)
