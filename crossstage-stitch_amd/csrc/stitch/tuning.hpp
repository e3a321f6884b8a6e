// Host-side launch-shape knobs shared by the stage launchers and the autotuner.
#pragma once

namespace stitch::jobs {
// Train rows per LDS segment of the S4 fast path (64..512, multiple of 64).
void s4_set_segment(int rows);
int s4_segment();
// Workgroups per CU of the persistent launches, as multiples of 1/8 of the
// per-kernel default (8: the defaults; 4: half as many; 16: twice as many).
void set_workgroups_per_cu(int eighths);
int workgroups_per_cu();
}  // namespace stitch::jobs
