// Marching cubes (placeholder until the all-labels kernel lands).
#include "grid.hpp"

namespace vf {

void grid_marching_cubes(vf_grid* g, uint32_t label, vf_mesh_info* info)
{
    (void)g; (void)label; (void)info;
    fail(VF_ERR_INVALID_ARG, "marching cubes not built yet");
}

void grid_mesh_read(vf_grid* g, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* voff, uint32_t* toff)
{
    (void)g; (void)xyz; (void)tri; (void)tri_flag; (void)labels; (void)voff; (void)toff;
    fail(VF_ERR_INVALID_ARG, "marching cubes not built yet");
}

}  // namespace vf
