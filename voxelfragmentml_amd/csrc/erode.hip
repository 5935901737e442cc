// Grid erosion: RegularGrid::erode (MF/Source/DataStructures/RegularGrid.cpp:82-159),
// erodeGrid-comp.glsl:26-59, copyGrid-comp.glsl, and the final
// removeIsolatedRegions (RegularGrid.cpp:1006-1015,
// removeIsolatedRegionsGrid-comp.glsl:16-38).
//
// One erosion iteration is detectBoundaries(1) (in place, launch_boundaries_tiled)
// followed by one stencil step read from the grid and written to a second
// buffer, then copied back, as the reference's erode / copyGrid pair does.
// The step's noise gate is the reference's uniform float noise buffer as a
// counter-based draw per cell: (hash32(seed, cell) >> 8) / 2^24.
//
// Both stencils read a clamped window around each cell from HBM (through L1 /
// L2); a window of side s costs s^3 loads per tested cell, so the stencil is
// bound by cache bandwidth for s >= 5 and by HBM for s = 3.
#include "cellops.hpp"

#include <cmath>
#include <limits>
#include <vector>

namespace vf {

namespace {

constexpr uint32_t kErodeThreads = 256;

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// erodeGrid-comp.glsl: the shader's isBoundary is unmaskedBit(value, 15) != 0,
// which holds for every non-EMPTY cell, so every cell above FREE that passes
// the noise gate is tested.  Neighbours count when their raw value (boundary
// bit included) equals the cell's and the mask is set; the cell empties when
// count / (in-bounds window cells) < limit, all in FP32.
__global__ __launch_bounds__(kErodeThreads) void k_erode_step(const JobDev* Jg, uint16_t* out, const uint8_t* mask,
                                                              uint32_t side, float limit, float probability,
                                                              uint64_t noise_seed)
{
    __shared__ uint8_t smask[kMaxErosionSide * kMaxErosionSide * kMaxErosionSide];
    const uint32_t n3 = side * side * side;
    for (uint32_t i = threadIdx.x; i < n3; i += blockDim.x) smask[i] = mask[i];
    __syncthreads();
    const uint16_t* in = Jg->labels;
    const int nx = (int)Jg->nx, ny = (int)Jg->ny, nz = (int)Jg->nz, h = (int)side / 2;
    const uint64_t V = Jg->V;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < V; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint16_t v = in[c];
        uint16_t r = v;
        if (v > kFree && (v & 0x7FFFu) != 0 &&
            (float)(hash32(noise_seed, (uint32_t)c) >> 8) * (1.0f / 16777216.0f) < probability) {
            const int z = (int)(c % (uint64_t)nz), y = (int)((c / (uint64_t)nz) % (uint64_t)ny),
                      x = (int)(c / ((uint64_t)ny * nz));
            const int x0 = clampi(x - h, 0, nx - 1), x1 = clampi(x + h, 0, nx - 1);
            const int y0 = clampi(y - h, 0, ny - 1), y1 = clampi(y + h, 0, ny - 1);
            const int z0 = clampi(z - h, 0, nz - 1), z1 = clampi(z + h, 0, nz - 1);
            uint32_t count = 0;
            for (int xx = x0; xx <= x1; ++xx)
                for (int yy = y0; yy <= y1; ++yy) {
                    const uint16_t* row = in + ((size_t)xx * ny + yy) * nz;
                    // mask index of (xx, yy, zz) relative to the unclamped window corner (x - h, y - h, z - h)
                    const int mbase = ((xx - (x - h)) * (int)side + (yy - (y - h))) * (int)side - (z - h);
                    for (int zz = z0; zz <= z1; ++zz) count += (row[zz] == v && smask[mbase + zz]) ? 1u : 0u;
                }
            const uint32_t global = (uint32_t)((x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1));
            if ((float)count / (float)global < limit) r = kEmpty;
        }
        out[c] = r;
    }
}

// removeIsolatedRegionsGrid-comp.glsl: a cell with fewer than 6 other cells of
// its raw value in the clamped 3x3x3 window becomes EMPTY.  The shader
// updates in place (thread-order dependent); this reads in and writes out.
__global__ __launch_bounds__(kErodeThreads) void k_isolated_cells(const JobDev* Jg, uint16_t* out)
{
    const uint16_t* in = Jg->labels;
    const int nx = (int)Jg->nx, ny = (int)Jg->ny, nz = (int)Jg->nz;
    const uint64_t V = Jg->V;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < V; c += (uint64_t)gridDim.x * blockDim.x) {
        const uint16_t v = in[c];
        const int z = (int)(c % (uint64_t)nz), y = (int)((c / (uint64_t)nz) % (uint64_t)ny),
                  x = (int)(c / ((uint64_t)ny * nz));
        int count = -1;
        for (int xx = clampi(x - 1, 0, nx - 1); xx <= clampi(x + 1, 0, nx - 1); ++xx)
            for (int yy = clampi(y - 1, 0, ny - 1); yy <= clampi(y + 1, 0, ny - 1); ++yy) {
                const uint16_t* row = in + ((size_t)xx * ny + yy) * nz;
                for (int zz = clampi(z - 1, 0, nz - 1); zz <= clampi(z + 1, 0, nz - 1); ++zz) count += row[zz] == v;
            }
        out[c] = count < 6 ? kEmpty : v;
    }
}

uint32_t stencil_blocks(uint64_t V)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((V + kErodeThreads - 1) / kErodeThreads, 8192));
}

}  // namespace

int erosion_mask(int type, uint32_t size, std::vector<uint8_t>& mask, float& activations)
{
    if (size % 2 == 0) ++size;
    const uint32_t n = size * size * size, c = size / 2;
    mask.assign(n, 0);
    float act = 0.0f;
    if (type == VF_EROSION_SQUARE) {
        std::fill(mask.begin(), mask.end(), 1);
        act = (float)n;
    } else if (type == VF_EROSION_CROSS) {
        for (uint32_t i = 0; i < size; ++i) {
            mask[i * size * size + c * size + c] = 1;
            mask[c * size * size + i * size + c] = 1;
            mask[c * size * size + c * size + i] = 1;
        }
        act = 1.0f / 3.0f * (float)n;  // as RegularGrid.cpp:104 sets it, not the cell count
    } else if (type == VF_EROSION_ELLIPSE) {
        for (uint32_t x = 0; x < size; ++x)
            for (uint32_t y = 0; y < size; ++y)
                for (uint32_t z = 0; z < size; ++z) {
                    const float dx = (float)x - (float)c, dy = (float)y - (float)c, dz = (float)z - (float)c;
                    if (std::sqrt(dx * dx + dy * dy + dz * dz) < (float)c + std::numeric_limits<float>::epsilon()) {
                        mask[x * size * size + y * size + z] = 1;
                        act += 1.0f;
                    }
                }
    } else {
        return -1;
    }
    activations = act / (float)n;
    return (int)size;
}

void launch_erode_step(const Launch& L, uint16_t* out, const uint8_t* mask, uint32_t side, float limit,
                       float probability, uint64_t noise_seed, uint64_t V)
{
    if (side > kMaxErosionSide || side % 2 == 0) fail(VF_ERR_INVALID_ARG, "erosion window side %u", side);
    TimedLaunch tl("erode_step", L.s);
    k_erode_step<<<stencil_blocks(V), kErodeThreads, 0, L.s>>>(L.jobs, out, mask, side, limit, probability,
                                                                noise_seed);
    VF_CHECK_LAUNCH();
}

void launch_isolated_cells(const Launch& L, uint16_t* out, uint64_t V)
{
    TimedLaunch tl("isolated_cells", L.s);
    k_isolated_cells<<<stencil_blocks(V), kErodeThreads, 0, L.s>>>(L.jobs, out);
    VF_CHECK_LAUNCH();
}

}  // namespace vf
