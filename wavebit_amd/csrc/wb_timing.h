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
void wb_timer_begin(int id, hipStream_t s);
void wb_timer_end(int id, hipStream_t s, double flops);

// `flops` is the launch's algorithmic real FP64 flop count (0 where none is defined).
#define WB_TIMED_F(id, stream, flops, ...)                     \
  do {                                                         \
    if (wb_timing_on) wb_timer_begin((id), (stream));          \
    __VA_ARGS__;                                               \
    if (wb_timing_on) wb_timer_end((id), (stream), (flops));   \
  } while (0)
#define WB_TIMED(id, stream, ...) WB_TIMED_F(id, stream, 0.0, __VA_ARGS__)
