// Optional per-kernel timing: when enabled (wb_timing_enable), selected launches are bracketed
// by hipEvent records on the stream they are launched on; wb_timing_read sums the elapsed
// times per kernel id. Disabled (the default), the hooks cost one branch.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/wavebit_amd.h"   // the timer ids WB_T_*

// Two timing modes: per-kernel (wb_timing_enable(1): WB_TIMED_F launches and fine stages) and
// per-stage (wb_timing_enable(2): the coarse stages of a solve, eigensolve vs S-matrices).
// A stage timer of the active mode suppresses the timers of the launches inside it.
enum { WB_STAGE_FINE = 0, WB_STAGE_COARSE = 1, WB_STAGE_BOTH = 2 };

extern bool wb_timing_on;
extern bool wb_timing_coarse;
extern int wb_timing_stage;   // > 0 while a multi-launch stage is being timed as one entry
void wb_timer_begin(int id, hipStream_t s);
void wb_timer_end(int id, hipStream_t s, double flops);

// `flops` is the launch's algorithmic real FP64 flop count (0 where none is defined). Inside a
// stage (WbStageTimer scope) the launches are not timed on their own.
#define WB_TIMED_F(id, stream, flops, ...)                                           \
  do {                                                                               \
    const bool timed__ = wb_timing_on && !wb_timing_coarse && wb_timing_stage == 0;  \
    if (timed__) wb_timer_begin((id), (stream));                                     \
    __VA_ARGS__;                                                                     \
    if (timed__) wb_timer_end((id), (stream), (flops));                              \
  } while (0)
// A stage of several launches timed as one entry of `id` with `flops` algorithmic flops, for
// the scope of the object (early returns included), in the timing mode(s) `level` names.
struct WbStageTimer {
  int id, level;
  hipStream_t stream;
  double flops;
  bool active = false, timed = false;
  WbStageTimer(int id_, hipStream_t s_, double flops_, int level_ = WB_STAGE_FINE)
      : id(id_), level(level_), stream(s_), flops(flops_) {
    start();
  }
  ~WbStageTimer() { stop(); }
  void start() {
    active = level == WB_STAGE_BOTH || (level == WB_STAGE_COARSE) == wb_timing_coarse;
    timed = active && wb_timing_on && wb_timing_stage == 0;
    if (timed) wb_timer_begin(id, stream);
    if (active) ++wb_timing_stage;
  }
  // End the stage early (the destructor then does nothing); start() opens it again.
  void stop() {
    if (active) --wb_timing_stage;
    if (timed) wb_timer_end(id, stream, flops);
    active = timed = false;
  }
};
#define WB_TIMED(id, stream, ...) WB_TIMED_F(id, stream, 0.0, __VA_ARGS__)
