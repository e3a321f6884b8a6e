// Host-side launch-shape knobs shared by the stage launchers and the autotuner.
#pragma once

namespace stitch::jobs {
// Train rows per LDS segment of the S4 fast path (64..512, multiple of 64).
void s4_set_segment(int rows);
int s4_segment();
}  // namespace stitch::jobs
