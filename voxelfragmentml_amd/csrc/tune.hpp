// Launch shapes chosen by the autotuner (tune.cpp), per bucket of the largest
// grid edge in a launch.  Launchers read the active shape of the calling
// thread; the engine activates a table's bucket before launching its stages.
#pragma once

#include <cstdint>

namespace vf {

struct LaunchShape {
    uint32_t cell_blocks;   // streaming cell stages (2048-cell tiles per block step)
    uint32_t flood_blocks;  // flood compaction / relaxation workgroups (one wave each)
    uint32_t pair_blocks;   // voxelizer SAT / parity pair chunks
    uint32_t brick_blocks;  // tiled boundary bricks
    uint32_t naive_run;     // tiles per run of the naive walker, occupancy loads issued together (4, 8, 16)
    uint32_t brick_run;     // bricks per step of the tiled boundary kernel, interiors loaded together (2, 4, 8)
    uint32_t round_chunk;   // flood rounds per captured chunk of the batched form (2 .. 32)
    uint32_t lead_chunks;   // chunks the batched form runs before its first drain check (1 .. 8)
    uint32_t mc_blocks;     // marching cubes' count pass workgroups (a multiple of 8, for the XCD split)
};

constexpr int kTuneBuckets = 4;               // largest edge <= 64, <= 128, <= 256, larger
int tune_bucket(uint32_t max_edge);
bool shape_ok(const LaunchShape& s);
LaunchShape active_shape();                   // shape for the calling thread's launches
void set_active_bucket(int bucket);           // per thread
LaunchShape get_shape(int bucket);
void set_shape(int bucket, const LaunchShape& s);
uint64_t tune_generation();                   // bumped whenever a shape changes (graphs recapture)

}  // namespace vf
