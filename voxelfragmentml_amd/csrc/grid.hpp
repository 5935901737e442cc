// vf_grid: the device-resident RegularGrid (MF/Source/DataStructures/RegularGrid.h:15-367).
#pragma once

#include "jobs.hpp"

struct vf_grid {
    uint32_t dims[3] = { 0, 0, 0 };
    float mn[3] = { 0, 0, 0 }, mx[3] = { 0, 0, 0 }, cs[3] = { 0, 0, 0 };
    uint64_t V = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    vf::DevBuf labels, occ;          // uint16 cells, uint8 occupancy
    vf::DevBuf jobdev, prefix, stats, hist, seeds;
    vf::DevBuf bits;                 // shell + toggle bitmasks
    vf::DevBuf vox_setup, vox_counts, vox_shell_off, scan_tmp, occ_tmp;
    vf::DevBuf mesh_xyz, mesh_tri;
    vf::DevBuf seed_tile_cnt, seed_tile_off;
    vf::DevBuf keys, buckets, fctl, ctl, flood_lists, flood_sub, near_in, near_out, erode_tmp, erode_mask;
    vf::DevBuf ri_a, ri_b, ri_cnt, ri_vis, ri_out;
    vf::McScratch mc;              // marching-cubes workspace and last result

    bool labels_from_occ = false;  // labels equal (occ ? FREE : EMPTY) but are not written yet
    bool clean = true;             // every cell EMPTY
    vf::JobDev job{};              // host copy of the descriptor
    bool job_dirty = true;
    uint32_t* pinned = nullptr;    // 4 KB pinned staging for small readbacks
    uint32_t last_flood_levels = 0;
};

namespace vf {

uint32_t grid_tiles(const vf_grid* g);
Launch grid_launch(vf_grid* g);     // uploads the descriptor when dirty
void grid_materialize(vf_grid* g);  // writes labels from occupancy if still pending
uint32_t key_bits_for(uint32_t S);
uint64_t read_stat(vf_grid* g, int slot);
void check_device_error(vf_grid* g, const char* stage);

void grid_marching_cubes(vf_grid* g, uint32_t label, vf_mesh_info* info);
void grid_mesh_read(vf_grid* g, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* voff,
                    uint32_t* toff);
// vertices [v0, v0 + nv) and triangles [t0, t0 + nt) of the last mesh
void grid_mesh_slice(vf_grid* g, uint32_t v0, uint32_t nv, uint32_t t0, uint32_t nt, float* xyz, uint32_t* tri);

}  // namespace vf
