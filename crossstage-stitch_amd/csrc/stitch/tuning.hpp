// Host-side launch-shape knobs shared by the stage launchers and the autotuner.
#pragma once

namespace stitch::jobs {
// Train rows per LDS segment of the S4 fast path (64..512, multiple of 64).
void s4_set_segment(int rows);
int s4_segment();

// Stages whose persistent launches size their grids separately.
enum class Stage : int { S1 = 0, S2, S3, S4, S5, S6 };
constexpr int kStages = 6;

// Workgroups per CU of the persistent launches, as multiples of 1/8 of the
// per-kernel default (8: the defaults; 4: half as many; 16: twice as many),
// for every stage at once or for one stage.
void set_workgroups_per_cu(int eighths);
void set_workgroups_per_cu(Stage stage, int eighths);
int workgroups_per_cu(Stage stage);

// Output rows per S1 tile (16, 32 or 64; tiles are 64 pixels wide).
void s1_set_tile_rows(int rows);
int s1_tile_rows();
}  // namespace stitch::jobs
