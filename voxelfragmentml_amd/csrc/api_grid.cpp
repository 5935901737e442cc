// C ABI for single grids: RegularGrid lifetime, copies, voxelization, seeding,
// fracture and the post-fracture stages.  Each entry point validates its
// arguments on the host (so no kernel is ever launched on shapes it does not
// support), launches the multi-job kernels with a one-job table, and reports
// failures through vf_last_error().
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "grid.hpp"
#include "tune.hpp"

namespace vf {

constexpr uint32_t kMaxEdge = 2048;

static void validate_dims(const uint32_t d[3])
{
    if (!d) fail(VF_ERR_INVALID_ARG, "dims is NULL");
    for (int q = 0; q < 3; ++q)
        if (d[q] == 0 || d[q] > kMaxEdge) fail(VF_ERR_INVALID_ARG, "grid edge %u out of range [1, %u]", d[q], kMaxEdge);
    if ((uint64_t)d[0] * d[1] * d[2] >= (1ull << 31)) fail(VF_ERR_INVALID_ARG, "grid has 2^31 cells or more");
}

static void validate_box(const float mn[3], const float mx[3])
{
    if (!mn || !mx) fail(VF_ERR_INVALID_ARG, "aabb pointer is NULL");
    for (int q = 0; q < 3; ++q)
        if (!std::isfinite(mn[q]) || !std::isfinite(mx[q]) || !(mx[q] > mn[q]))
            fail(VF_ERR_INVALID_ARG, "aabb axis %d is empty or not finite", q);
}

uint32_t grid_tiles(const vf_grid* g) { return ceil_div(g->V, kTileCells); }

uint32_t key_bits_for(uint32_t S)
{
    uint32_t b = 1;
    while ((1ull << b) < S) ++b;
    return b;
}

NaiveKeyForm naive_key_form(const uint32_t dims[3], int distance, uint32_t S)
{
    const uint32_t b = key_bits_for(S);
    const uint64_t ex = dims[0] - 1, ey = dims[1] - 1, ez = dims[2] - 1;
    const uint64_t dmax = distance == VF_MANHATTAN ? ex + ey + ez
                          : distance == VF_CHEBYSHEV ? std::max(ex, std::max(ey, ez))
                                                     : ex * ex + ey * ey + ez * ez;
    NaiveKeyForm f{ distance, 0 };
    if (distance == VF_EUCLIDEAN && dmax > kSqrtInjectiveMax) {
        f.dist = kDistEuclidF32;  // float root bits (below 2^31) << b: 64-bit keys
        f.key64 = 1;
        return f;
    }
    // 32-bit keys pack d * 2^b + i with a 24-bit multiply: 2^b must fit 24 bits
    f.key64 = b > 23 || ((dmax + 1) << b) > 0xFFFFFFFFull ? 1 : 0;
    return f;
}

static void grid_setup_shape(vf_grid* g, const uint32_t dims[3], const float mn[3], const float mx[3])
{
    for (int q = 0; q < 3; ++q) {
        g->dims[q] = dims[q];
        g->mn[q] = mn[q];
        g->mx[q] = mx[q];
        g->cs[q] = (mx[q] - mn[q]) / (float)dims[q];  // RegularGrid::setAABB, RegularGrid.cpp:426-441
    }
    g->V = (uint64_t)dims[0] * dims[1] * dims[2];
    g->labels.ensure(g->V * 2 + 64);
    g->occ.ensure(g->V + 64);
    VF_HIP(hipMemsetAsync(g->labels.p, 0, g->V * 2, g->stream));
    VF_HIP(hipMemsetAsync(g->occ.p, 0, g->V, g->stream));
    g->labels_from_occ = false;
    g->clean = true;
    g->mc.valid = false;
    JobDev& J = g->job;
    J = JobDev{};
    J.nx = dims[0];
    J.ny = dims[1];
    J.nz = dims[2];
    J.wp = next_pow2(ceil_div(dims[2], 32));
    J.V = g->V;
    for (int q = 0; q < 3; ++q) {
        J.mn[q] = g->mn[q];
        J.mx[q] = g->mx[q];
        J.cs[q] = g->cs[q];
    }
    J.labels = g->labels.as<uint16_t>();
    J.occ = g->occ.as<uint8_t>();
    J.stats = g->stats.as<unsigned long long>();
    J.hist = g->hist.as<uint32_t>();
    J.active = 1;
    g->job_dirty = true;
    uint32_t pre[4] = { 0, grid_tiles(g), 0, brick_count(dims[0], dims[1], dims[2]) };
    VF_HIP(hipMemcpyAsync(g->prefix.p, pre, sizeof pre, hipMemcpyHostToDevice, g->stream));
    VF_HIP(hipStreamSynchronize(g->stream));
}

Launch grid_launch(vf_grid* g)
{
    if (g->job_dirty) {
        std::memcpy(g->pinned + 256, &g->job, sizeof(JobDev));
        VF_HIP(hipMemcpyAsync(g->jobdev.p, g->pinned + 256, sizeof(JobDev), hipMemcpyHostToDevice, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
        g->job_dirty = false;
    }
    set_active_bucket(tune_bucket(std::max(g->dims[0], std::max(g->dims[1], g->dims[2]))));
    Launch L;
    L.jobs = g->jobdev.as<JobDev>();
    L.njobs = 1;
    L.tile_prefix = g->prefix.as<uint32_t>();
    L.total_tiles = grid_tiles(g);
    L.s = g->stream;
    return L;
}

void grid_materialize(vf_grid* g)
{
    if (!g->labels_from_occ) return;
    launch_occ_to_labels(grid_launch(g));
    g->labels_from_occ = false;
}

uint64_t read_stat(vf_grid* g, int slot)
{
    VF_HIP(hipMemcpyAsync(g->pinned, g->stats.as<unsigned long long>() + slot, 8, hipMemcpyDeviceToHost, g->stream));
    VF_HIP(hipStreamSynchronize(g->stream));
    uint64_t v;
    std::memcpy(&v, g->pinned, 8);
    return v;
}

void check_device_error(vf_grid* g, const char* stage)
{
    uint64_t e = read_stat(g, kStatError);
    if (!e) return;
    VF_HIP(hipMemsetAsync(g->stats.as<unsigned long long>() + kStatError, 0, 8, g->stream));
    if (e & kDevErrSeeding) fail(VF_ERR_SEEDING, "%s: fewer candidate cells than seeds", stage);
    if (e & kDevErrKeyOverflow) fail(VF_ERR_OVERFLOW, "%s: flood distance exceeds the packed key range", stage);
    if (e & kDevErrRelax) fail(VF_ERR_OVERFLOW, "%s: flood brick did not settle", stage);
    if (e & kDevErrMesh) fail(VF_ERR_INVALID_ARG, "%s: a triangle references a vertex index >= nv", stage);
    fail(VF_ERR_OVERFLOW, "%s: device error flag 0x%llx", stage, (unsigned long long)e);
}

static void validate_seeds(const vf_grid* g, const vf_seed* seeds, uint32_t S)
{
    if (S && !seeds) fail(VF_ERR_INVALID_ARG, "seeds is NULL");
    for (uint32_t s = 0; s < S; ++s)
        if (seeds[s].x >= g->dims[0] || seeds[s].y >= g->dims[1] || seeds[s].z >= g->dims[2])
            fail(VF_ERR_INVALID_ARG, "seed %u at (%u,%u,%u) lies outside the %ux%ux%u grid", s, seeds[s].x, seeds[s].y,
                 seeds[s].z, g->dims[0], g->dims[1], g->dims[2]);
}

static void upload_seeds(vf_grid* g, const vf_seed* seeds, uint32_t S)
{
    g->seeds.ensure((size_t)std::max<uint32_t>(S, 1) * sizeof(uint4));
    if (S) VF_HIP(hipMemcpyAsync(g->seeds.p, seeds, (size_t)S * sizeof(uint4), hipMemcpyHostToDevice, g->stream));
    if (g->job.seeds != g->seeds.as<uint4>() || g->job.S != S) {
        g->job.seeds = g->seeds.as<uint4>();
        g->job.S = S;
        g->job_dirty = true;
    }
    g->job.S_draw = S;
    VF_HIP(hipStreamSynchronize(g->stream));
}

static vf_grid* checked(vf_grid* g)
{
    if (!g) fail(VF_ERR_INVALID_ARG, "grid is NULL");
    VF_HIP(hipSetDevice(g->device));
    return g;
}

}  // namespace vf

using namespace vf;

extern "C" {

int vf_grid_create(const uint32_t dims[3], const float aabb_min[3], const float aabb_max[3], vf_grid** out)
{
    return guard([&] {
        if (!out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        *out = nullptr;
        validate_dims(dims);
        validate_box(aabb_min, aabb_max);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            (void)hipGetLastError();
            fail(VF_ERR_DEVICE, "no HIP device available");
        }
        auto* g = new vf_grid();
        try {
            VF_HIP(hipGetDevice(&g->device));
            VF_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
            VF_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->pinned), 4096));
            g->jobdev.ensure(sizeof(JobDev));
            g->prefix.ensure(64);
            g->stats.ensure(kStatCount * 8);
            g->hist.ensure(65536 * 4);
            g->ctl.ensure(256);
            VF_HIP(hipMemsetAsync(g->stats.p, 0, kStatCount * 8, g->stream));
            grid_setup_shape(g, dims, aabb_min, aabb_max);
        } catch (...) {
            vf_grid_destroy(g);
            throw;
        }
        *out = g;
    });
}

int vf_grid_destroy(vf_grid* g)
{
    return guard([&] {
        if (!g) return;
        (void)hipSetDevice(g->device);
        if (g->stream) (void)hipStreamSynchronize(g->stream);
        if (g->pinned) (void)hipHostFree(g->pinned);
        if (g->stream) (void)hipStreamDestroy(g->stream);
        delete g;
    });
}

int vf_grid_set_aabb(vf_grid* g, const float aabb_min[3], const float aabb_max[3], const uint32_t dims[3])
{
    return guard([&] {
        checked(g);
        validate_dims(dims);
        validate_box(aabb_min, aabb_max);
        grid_setup_shape(g, dims, aabb_min, aabb_max);
    });
}

int vf_grid_dims(const vf_grid* g, uint32_t dims[3])
{
    return guard([&] {
        if (!g || !dims) fail(VF_ERR_INVALID_ARG, "NULL argument");
        for (int q = 0; q < 3; ++q) dims[q] = g->dims[q];
    });
}

int vf_grid_read(vf_grid* g, uint16_t* out)
{
    return guard([&] {
        checked(g);
        if (!out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        grid_materialize(g);
        VF_HIP(hipMemcpyAsync(out, g->labels.p, g->V * 2, hipMemcpyDeviceToHost, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_grid_write(vf_grid* g, const uint16_t* in)
{
    return guard([&] {
        checked(g);
        if (!in) fail(VF_ERR_INVALID_ARG, "in is NULL");
        VF_HIP(hipMemcpyAsync(g->labels.p, in, g->V * 2, hipMemcpyHostToDevice, g->stream));
        g->labels_from_occ = false;
        g->clean = false;
        g->mc.valid = false;
        launch_labels_to_occ(grid_launch(g));
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_occupancy_read(vf_grid* g, uint8_t* out)
{
    return guard([&] {
        checked(g);
        if (!out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        VF_HIP(hipMemcpyAsync(out, g->occ.p, g->V, hipMemcpyDeviceToHost, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_grid_count_occupied(vf_grid* g, uint64_t* out)
{
    return guard([&] {
        checked(g);
        if (!out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        VF_HIP(hipMemsetAsync(g->stats.as<unsigned long long>() + kStatOccupied, 0, 8, g->stream));
        launch_count_occupied(grid_launch(g));
        *out = read_stat(g, kStatOccupied);
    });
}

int vf_grid_device_ptr(vf_grid* g, void** ptr)
{
    return guard([&] {
        checked(g);
        if (!ptr) fail(VF_ERR_INVALID_ARG, "ptr is NULL");
        grid_materialize(g);
        VF_HIP(hipStreamSynchronize(g->stream));
        *ptr = g->labels.p;
    });
}

int vf_grid_synchronize(vf_grid* g)
{
    return guard([&] {
        checked(g);
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_voxelize_device(vf_grid* g, const float* d_xyz, uint32_t nv, const uint32_t* d_tri, uint32_t nt)
{
    return guard([&] {
        checked(g);
        if (nt && (!d_xyz || !d_tri)) fail(VF_ERR_INVALID_ARG, "mesh pointer is NULL");
        JobDev& J = g->job;
        const uint64_t words = (uint64_t)J.nx * J.ny * J.wp;
        g->bits.ensure(words * 8 + 64);
        J.shell_bits = g->bits.as<uint32_t>();
        J.toggle_bits = g->bits.as<uint32_t>() + words;
        J.xyz = d_xyz;
        J.tri = d_tri;
        J.nt = nt;
        J.nv = nv;
        // RegularGrid::fill only sets occupied cells to FREE; on a grid that
        // already holds values the new occupancy is merged afterwards.
        const bool merge = !g->clean;
        if (merge) {
            g->occ_tmp.ensure(g->V + 64);
            J.occ = g->occ_tmp.as<uint8_t>();
        }
        g->job_dirty = true;
        const uint32_t groups = ceil_div(words, 256);
        uint32_t pre[4] = { 0, nt, 0, groups };
        g->vox_setup.ensure((size_t)std::max<uint32_t>(nt, 1) * sizeof(TriSetup));
        g->vox_counts.ensure((size_t)std::max<uint32_t>(nt, 1) * 8);
        g->vox_shell_off.ensure((2 * (size_t)nt + 1) * 8);
        g->scan_tmp.ensure(scan_tmp_bytes(2 * (uint64_t)nt) + 64);
        g->ctl.ensure(256);
        std::memcpy(g->pinned + 512, pre, sizeof pre);
        VF_HIP(hipMemcpyAsync(g->ctl.p, g->pinned + 512, sizeof pre, hipMemcpyHostToDevice, g->stream));
        Launch L = grid_launch(g);
        VoxScratch X;
        X.setup = g->vox_setup.as<TriSetup>();
        X.counts = g->vox_counts.as<uint32_t>();
        X.shell_off = g->vox_shell_off.as<unsigned long long>();
        X.tri_prefix = g->ctl.as<uint32_t>();
        X.total_tris = nt;
        X.scan_tmp = g->scan_tmp.p;
        X.scan_tmp_bytes = g->scan_tmp.bytes;
        launch_voxelize(L, X, g->ctl.as<uint32_t>() + 2, groups, nullptr);
        check_device_error(g, "voxelize");
        if (merge) {
            // labels = new ? FREE : old, occupancy = old | new
            std::vector<uint16_t> old(g->V);
            std::vector<uint8_t> nw(g->V);
            grid_materialize(g);
            VF_HIP(hipMemcpyAsync(old.data(), g->labels.p, g->V * 2, hipMemcpyDeviceToHost, g->stream));
            VF_HIP(hipMemcpyAsync(nw.data(), g->occ_tmp.p, g->V, hipMemcpyDeviceToHost, g->stream));
            VF_HIP(hipStreamSynchronize(g->stream));
            for (uint64_t i = 0; i < g->V; ++i)
                if (nw[i]) old[i] = kFree;
            J.occ = g->occ.as<uint8_t>();
            g->job_dirty = true;
            VF_HIP(hipMemcpyAsync(g->labels.p, old.data(), g->V * 2, hipMemcpyHostToDevice, g->stream));
            launch_labels_to_occ(grid_launch(g));
            g->labels_from_occ = false;
        } else {
            g->labels_from_occ = true;
        }
        g->clean = false;
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_voxelize(vf_grid* g, const float* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt)
{
    return guard([&] {
        checked(g);
        if (nt && (!xyz || !tri)) fail(VF_ERR_INVALID_ARG, "mesh pointer is NULL");
        for (uint64_t i = 0; i < 3ull * nt; ++i)
            if (tri[i] >= nv) fail(VF_ERR_INVALID_ARG, "triangle %llu references vertex %u of %u", (unsigned long long)(i / 3), tri[i], nv);
        g->mesh_xyz.ensure((size_t)std::max<uint32_t>(nv, 1) * 12);
        g->mesh_tri.ensure((size_t)std::max<uint32_t>(nt, 1) * 12);
        if (nv) VF_HIP(hipMemcpyAsync(g->mesh_xyz.p, xyz, (size_t)nv * 12, hipMemcpyHostToDevice, g->stream));
        if (nt) VF_HIP(hipMemcpyAsync(g->mesh_tri.p, tri, (size_t)nt * 12, hipMemcpyHostToDevice, g->stream));
        int rc = vf_voxelize_device(g, g->mesh_xyz.as<float>(), nv, g->mesh_tri.as<uint32_t>(), nt);
        if (rc != VF_OK) throw Exception{ rc };
    });
}

int vf_seed_uniform(vf_grid* g, uint32_t S, uint64_t seed, int location, vf_seed* out)
{
    return guard([&] {
        checked(g);
        if (S && !out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        if (location != VF_SEED_BOTH && location != VF_SEED_OUTER && location != VF_SEED_INNER)
            fail(VF_ERR_INVALID_ARG, "unknown seed location %d", location);
        if (S > kMaxDeviceSeeds) fail(VF_ERR_INVALID_ARG, "at most %u seeds can be drawn", kMaxDeviceSeeds);
        if (S == 0) return;
        g->seeds.ensure((size_t)S * sizeof(uint4));
        g->job.seeds = g->seeds.as<uint4>();
        g->job.S = S;
        g->job.S_draw = S;
        g->job.seed_key = seed;
        g->job.seed_location = location;
        g->job_dirty = true;
        const uint32_t tiles = grid_tiles(g);
        g->seed_tile_cnt.ensure((size_t)tiles * 4 + 64);
        g->seed_tile_off.ensure(((size_t)tiles + 1) * 8 + 64);
        g->scan_tmp.ensure(scan_tmp_bytes(tiles) + 64);
        Launch L = grid_launch(g);
        launch_seed_count(L, g->seed_tile_cnt.as<uint32_t>(), false);
        scan_exclusive_u32_to_u64(g->seed_tile_cnt.as<uint32_t>(), g->seed_tile_off.as<unsigned long long>(), tiles,
                                  g->scan_tmp.p, g->stream);
        launch_seed_pick(L, g->seed_tile_off.as<unsigned long long>());
        check_device_error(g, "seed_uniform");
        VF_HIP(hipMemcpyAsync(out, g->seeds.p, (size_t)S * sizeof(uint4), hipMemcpyDeviceToHost, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_seed_near(vf_grid* g, const vf_seed* frags, uint32_t F, uint32_t impacts, uint32_t num_seeds,
                 uint32_t spreading, uint64_t seed, vf_seed* out, uint32_t* count)
{
    return guard([&] {
        checked(g);
        if (!frags || !F) fail(VF_ERR_INVALID_ARG, "near seeding needs at least one fragment seed");
        if (!out || !count) fail(VF_ERR_INVALID_ARG, "out or count is NULL");
        if (spreading == 0 || spreading > 64) fail(VF_ERR_INVALID_ARG, "spreading %u outside 1..64", spreading);
        if (spreading > std::min(g->dims[0], std::min(g->dims[1], g->dims[2])))
            fail(VF_ERR_INVALID_ARG, "spreading %u above the smallest grid edge", spreading);
        if (num_seeds > kMaxDeviceSeeds || F > kMaxDeviceSeeds)
            fail(VF_ERR_INVALID_ARG, "at most %u seeds", kMaxDeviceSeeds);
        validate_seeds(g, frags, F);
        grid_materialize(g);
        g->near_in.ensure((size_t)F * sizeof(uint4) + 64);
        g->near_out.ensure(((size_t)F + num_seeds) * sizeof(uint4) + 64);
        VF_HIP(hipMemcpyAsync(g->near_in.p, frags, (size_t)F * sizeof(uint4), hipMemcpyHostToDevice, g->stream));
        Launch L = grid_launch(g);
        g->fctl.ensure(64 * 4);
        uint32_t* dres = g->fctl.as<uint32_t>() + 40;  // words 40..41: count, draw-cap flag
        launch_seed_near(L, g->near_in.as<uint4>(), F, impacts, num_seeds, spreading, seed, g->near_out.as<uint4>(), dres);
        VF_HIP(hipMemcpyAsync(g->pinned, dres, 8, hipMemcpyDeviceToHost, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
        if (g->pinned[1]) fail(VF_ERR_INVALID_ARG, "an impact found too few free surface cells in 10^6 draws");
        const uint32_t n = g->pinned[0];
        VF_HIP(hipMemcpyAsync(out, g->near_out.p, (size_t)n * sizeof(uint4), hipMemcpyDeviceToHost, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
        *count = n;
    });
}

int vf_seed_merge(const vf_seed* frags, uint32_t F, vf_seed* seeds, uint32_t N, int distance)
{
    return guard([&] {
        if (N == 0) return;
        if (!seeds || !frags || F == 0) fail(VF_ERR_INVALID_ARG, "merging needs fragment seeds");
        if (distance < VF_EUCLIDEAN || distance > VF_CHEBYSHEV) fail(VF_ERR_INVALID_ARG, "unknown distance %d", distance);
        for (uint32_t f = 0; f < F; ++f)
            if (frags[f].w > 255) fail(VF_ERR_INVALID_ARG, "fragment label %u does not fit 8 bits", frags[f].w);
        // the reference's float distances on vec3 coordinates (glm::distance for Euclidean), compiled
        // without contraction: integer coordinates below 2^24 convert exactly, sums are rounded in
        // the order written
        auto d = [&](const vf_seed& a, const vf_seed& b) -> float {
            const float dx = (float)a.x - (float)b.x, dy = (float)a.y - (float)b.y, dz = (float)a.z - (float)b.z;
            if (distance == VF_EUCLIDEAN) return std::sqrt(dx * dx + dy * dy + dz * dz);
            if (distance == VF_MANHATTAN) return std::fabs(dx) + std::fabs(dy) + std::fabs(dz);
            return std::max(std::fabs(dx), std::max(std::fabs(dy), std::fabs(dz)));
        };
        uint32_t next[256] = {};
        std::vector<uint32_t> label(N);
        for (uint32_t i = 0; i < N; ++i) {
            uint32_t nearest = 0;
            float best = FLT_MAX;
            for (uint32_t f = 0; f < F; ++f) {
                const float v = d(seeds[i], frags[f]);
                if (v < best) {
                    best = v;
                    nearest = f;
                }
            }
            const uint32_t w = frags[nearest].w;
            if (++next[w] > 255) fail(VF_ERR_TOO_MANY_SEEDS, "fragment %u would get more than 255 sub-seeds", w);
            label[i] = w | (next[w] << 8);
        }
        for (uint32_t i = 0; i < N; ++i) seeds[i].w = label[i];
    });
}

int vf_fracture_naive(vf_grid* g, const vf_seed* seeds, uint32_t S, int distance)
{
    return guard([&] {
        checked(g);
        if (distance < VF_EUCLIDEAN || distance > VF_CHEBYSHEV) fail(VF_ERR_INVALID_ARG, "unknown distance %d", distance);
        validate_seeds(g, seeds, S);
        if (S == 0) return;  // buildCPU leaves every cell untouched without seeds
        upload_seeds(g, seeds, S);
        const NaiveKeyForm kf = naive_key_form(g->dims, distance, S);
        const bool key64 = kf.key64 != 0;
        g->job.key_bits = key_bits_for(S);
        g->job.dist = kf.dist;
        g->job.algo = VF_NAIVE;
        g->job.naive_key64 = kf.key64;
        g->job_dirty = true;
        launch_naive(grid_launch(g), key64);
        g->labels_from_occ = false;
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_remove_isolated_regions(vf_grid* g, const vf_seed* seeds, uint32_t S)
{
    return guard([&] {
        checked(g);
        validate_seeds(g, seeds, S);
        if (S > g->V) fail(VF_ERR_INVALID_ARG, "more seeds than cells");
        upload_seeds(g, seeds, S);
        grid_materialize(g);
        g->ri_a.ensure(g->V * 8 + 64);
        g->ri_b.ensure(g->V * 8 + 64);
        g->ri_cnt.ensure(64);
        g->ri_vis.ensure((g->V + 31) / 32 * 4 + 64);
        g->ri_out.ensure(g->V * 2 + 64);
        launch_remove_isolated(g->stream, g->labels.as<uint16_t>(), g->ri_out.as<uint16_t>(), g->ri_a.as<uint32_t>(),
                               g->ri_b.as<uint32_t>(), g->ri_cnt.as<uint32_t>(), g->ri_vis.as<uint32_t>(), g->dims,
                               g->seeds.as<uint4>(), S, g->pinned);
        VF_HIP(hipMemcpyAsync(g->labels.p, g->ri_out.p, g->V * 2, hipMemcpyDeviceToDevice, g->stream));
        launch_labels_to_occ(grid_launch(g));
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_fracture_flood(vf_grid* g, const vf_seed* seeds, uint32_t S, int distance, int neighbourhood, uint64_t noise_seed,
                      uint32_t noise_amp)
{
    return guard([&] {
        checked(g);
        if (distance < VF_EUCLIDEAN || distance > VF_CHEBYSHEV) fail(VF_ERR_INVALID_ARG, "unknown distance %d", distance);
        int nb = neighbourhood;
        if (nb == VF_NEIGH_FROM_DISTANCE) nb = distance == VF_MANHATTAN ? 6 : 26;  // FloodFracturer.cpp:114
        if (nb != 6 && nb != 26) fail(VF_ERR_INVALID_ARG, "neighbourhood must be 0, 6 or 26");
        if (noise_amp > kMaxNoiseAmp) fail(VF_ERR_INVALID_ARG, "noise amplitude %u above %u", noise_amp, kMaxNoiseAmp);
        validate_seeds(g, seeds, S);
        for (uint32_t s = 0; s < S; ++s)
            if ((seeds[s].w & 0xFFu) < 2 || seeds[s].w > 0xFFFF)
                fail(VF_ERR_TOO_MANY_SEEDS, "flood label 0x%x: the fragment id (low 8 bits, Seeder.h:12) must be 2..255 "
                     "and a sub-seed id fits bits 8..15", seeds[s].w);
        if (S > kMaxDeviceSeeds) fail(VF_ERR_TOO_MANY_SEEDS, "%u flood seeds (max %u)", S, kMaxDeviceSeeds);
        // the sub-seed pass matters only when some fragment's live seeds carry different ids
        bool sub = false;
        {
            std::map<uint64_t, uint32_t> last;  // cell -> last seed on it
            for (uint32_t s = 0; s < S; ++s)
                last[((uint64_t)seeds[s].x * g->dims[1] + seeds[s].y) * g->dims[2] + seeds[s].z] = s;
            std::map<uint32_t, std::set<uint32_t>> ids;
            for (auto& kv : last) ids[seeds[kv.second].w & 0xFFu].insert(seeds[kv.second].w >> 8);
            for (auto& kv : ids) sub |= kv.second.size() > 1;
        }
        upload_seeds(g, seeds, S);
        JobDev& J = g->job;
        if (sub) g->flood_sub.ensure(8ull * S + 64);
        J.sub = sub ? g->flood_sub.as<uint32_t>() : nullptr;
        const uint32_t nbricks = relax_bricks(g->dims[0], g->dims[1], g->dims[2]);
        g->keys.ensure(g->V * 4 + 64);
        g->buckets.ensure(relax_flag_bytes(g->dims[0], g->dims[1], g->dims[2]));  // two flag parities, brick liveness
        g->fctl.ensure(64 * 4);
        J.keys = g->keys.as<uint32_t>();
        J.bflags = g->buckets.as<unsigned long long>();
        J.nbricks = nbricks;
        J.fctl = g->fctl.as<uint32_t>();
        J.key_bits = key_bits_for(S);
        J.neighbourhood = nb;
        J.noise_seed = noise_seed;
        J.noise_amp = noise_amp;
        J.algo = VF_FLOOD;
        g->job_dirty = true;
        Launch L = grid_launch(g);
        // fctl words 32..33: task prefix of the one job; 48..63: worklist counters (two list sets)
        const uint32_t tasks = relax_tasks(g->dims[0], g->dims[1], g->dims[2]);
        FloodRoundScratch W{};
        W.task_prefix = g->fctl.as<uint32_t>() + 32;
        W.total_tasks = tasks;
        W.list_cap = flood_list_cap(tasks);
        g->flood_lists.ensure(2ull * kFloodXcds * W.list_cap * 8);
        W.lists = g->flood_lists.as<unsigned long long>();
        W.counts = g->fctl.as<uint32_t>() + 48;
        std::memset(g->pinned, 0, 128);
        g->pinned[1] = tasks;
        VF_HIP(hipMemcpyAsync(g->fctl.as<uint32_t>() + 32, g->pinned, 128, hipMemcpyHostToDevice, g->stream));
        VF_HIP(hipMemsetAsync(J.bflags, 0, 16ull * relax_words_per_parity(nbricks), g->stream));  // both flag parities
        VF_HIP(hipStreamSynchronize(g->stream));  // the pinned words are reused below
        launch_flood_init(L, W);
        launch_flood_seed(L, W);
        auto drain = [&]() {
            uint32_t chunk = 4, rounds = 0;
            for (;;) {
                launch_flood_rounds(L, W, chunk, active_shape().flood_blocks);
                rounds += chunk;
                VF_HIP(hipMemcpyAsync(g->pinned, g->fctl.p, 8, hipMemcpyDeviceToHost, g->stream));
                VF_HIP(hipStreamSynchronize(g->stream));
                if (g->pinned[1] != 0) return g->pinned[1];
                chunk = std::min<uint32_t>(chunk * 2, 32);
                // every round lowers a cell's key at least one step along a shortest path, so the
                // worklist drains within (cells + 1) rounds
                if ((uint64_t)rounds > g->V + S + 16) fail(VF_ERR_OVERFLOW, "flood did not drain");
            }
        };
        g->last_flood_levels = drain();
        if (sub) {
            launch_flood_sub(L);
            VF_HIP(hipMemcpyAsync(g->pinned, g->fctl.p, 8, hipMemcpyDeviceToHost, g->stream));
            VF_HIP(hipStreamSynchronize(g->stream));
            if (g->pinned[1] == 0) g->last_flood_levels += drain();  // regions were freed: second growth
        }
        check_device_error(g, "flood");
        launch_flood_finish(L);
        g->labels_from_occ = false;
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_detect_boundaries(vf_grid* g, int size)
{
    return guard([&] {
        checked(g);
        if (size < 0 || size > 64) fail(VF_ERR_INVALID_ARG, "boundary size %d out of range", size);
        grid_materialize(g);
        if (g->job.bsize != size) {
            g->job.bsize = size;
            g->job_dirty = true;
        }
        Launch L = grid_launch(g);
        if (size <= 1) launch_boundaries_tiled(L, g->prefix.as<uint32_t>() + 2, brick_count(g->dims[0], g->dims[1], g->dims[2]), false,
                                              false);
        else launch_boundaries(L, size, 0);
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_undo_mask(vf_grid* g)
{
    return guard([&] {
        checked(g);
        grid_materialize(g);
        launch_undo_mask(grid_launch(g), 0x7FFF);
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_erode(vf_grid* g, int convolution, uint32_t size, uint32_t iterations, float probability, float threshold,
             uint64_t noise_seed)
{
    return guard([&] {
        checked(g);
        std::vector<uint8_t> mask;
        float act = 0.0f;
        if (size == 0 || size > kMaxErosionSide) fail(VF_ERR_INVALID_ARG, "erosion size %u outside 1..%u", size, kMaxErosionSide);
        const int side = erosion_mask(convolution, size, mask, act);
        if (side < 0) fail(VF_ERR_INVALID_ARG, "unknown erosion convolution %d", convolution);
        const float limit = act * threshold;
        grid_materialize(g);
        g->erode_tmp.ensure(g->V * 2 + 64);
        g->erode_mask.ensure(mask.size() + 64);
        VF_HIP(hipMemcpyAsync(g->erode_mask.p, mask.data(), mask.size(), hipMemcpyHostToDevice, g->stream));
        if (g->job.bsize != 1) {
            g->job.bsize = 1;
            g->job_dirty = true;
        }
        Launch L = grid_launch(g);
        const uint32_t bricks = brick_count(g->dims[0], g->dims[1], g->dims[2]);
        for (uint32_t it = 0; it < iterations; ++it) {
            launch_boundaries_tiled(L, g->prefix.as<uint32_t>() + 2, bricks, false, false);
            launch_erode_step(L, g->erode_tmp.as<uint16_t>(), g->erode_mask.as<uint8_t>(), (uint32_t)side, limit,
                              probability, noise_seed, g->V);
            VF_HIP(hipMemcpyAsync(g->labels.p, g->erode_tmp.p, g->V * 2, hipMemcpyDeviceToDevice, g->stream));
        }
        launch_isolated_cells(L, g->erode_tmp.as<uint16_t>(), g->V);
        VF_HIP(hipMemcpyAsync(g->labels.p, g->erode_tmp.p, g->V * 2, hipMemcpyDeviceToDevice, g->stream));
        launch_labels_to_occ(L);
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_remove_isolated_cells(vf_grid* g)
{
    return guard([&] {
        checked(g);
        grid_materialize(g);
        g->erode_tmp.ensure(g->V * 2 + 64);
        Launch L = grid_launch(g);
        launch_isolated_cells(L, g->erode_tmp.as<uint16_t>(), g->V);
        VF_HIP(hipMemcpyAsync(g->labels.p, g->erode_tmp.p, g->V * 2, hipMemcpyDeviceToDevice, g->stream));
        launch_labels_to_occ(L);
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_homogenize(vf_grid* g)
{
    return guard([&] {
        checked(g);
        grid_materialize(g);
        launch_homogenize(grid_launch(g));
        g->mc.valid = false;
        VF_HIP(hipStreamSynchronize(g->stream));
    });
}

int vf_histogram(vf_grid* g, uint32_t* counts, uint32_t ncounts)
{
    return guard([&] {
        checked(g);
        if (ncounts && !counts) fail(VF_ERR_INVALID_ARG, "counts is NULL");
        grid_materialize(g);
        Launch L = grid_launch(g);
        launch_clear_hist(L);
        launch_histogram(L);
        const uint32_t n = std::min<uint32_t>(ncounts, 65536);
        if (n) VF_HIP(hipMemcpyAsync(counts, g->hist.p, (size_t)n * 4, hipMemcpyDeviceToHost, g->stream));
        VF_HIP(hipStreamSynchronize(g->stream));
        for (uint32_t i = n; i < ncounts; ++i) counts[i] = 0;
    });
}

int vf_fragment_metadata(const uint32_t* counts, const uint32_t dims[3], vf_fragment_meta* out, uint32_t cap,
                         uint32_t* n)
{
    return guard([&] {
        if (!counts || !dims || !n) fail(VF_ERR_INVALID_ARG, "counts, dims or n is NULL");
        if (cap && !out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        // labels above FREE only, as countValues reports them; the reference sums
        // into an unsigned
        uint32_t total = 0, m = 0;
        for (uint32_t v = 2; v < 65536; ++v) {
            total += counts[v];
            m += counts[v] != 0;
        }
        uint32_t k = 0;
        for (uint32_t v = 2; v < 65536 && k < cap; ++v) {
            if (!counts[v]) continue;
            vf_fragment_meta& f = out[k];
            f.id = k;
            f.value = v;
            f.voxels = counts[v];
            f.occupied_voxels = total;
            f.percentage = (float)counts[v] / (float)total;
            for (int a = 0; a < 3; ++a) f.voxelization_size[a] = dims[a];
            ++k;
        }
        *n = m;
    });
}

int vf_marching_cubes(vf_grid* g, uint32_t label, vf_mesh_info* info)
{
    return guard([&] {
        checked(g);
        if (!info) fail(VF_ERR_INVALID_ARG, "info is NULL");
        grid_materialize(g);
        grid_marching_cubes(g, label, info);
    });
}

int vf_mesh_read(vf_grid* g, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* vert_offsets,
                 uint32_t* tri_offsets)
{
    return guard([&] {
        checked(g);
        if (!g->mc.valid) fail(VF_ERR_INVALID_ARG, "no marching-cubes result: call vf_marching_cubes first");
        grid_mesh_read(g, xyz, tri, tri_flag, labels, vert_offsets, tri_offsets);
    });
}

int vf_mesh_smooth(vf_grid* g, uint32_t nonboundary_iterations, float nonboundary_weight, uint32_t boundary_iterations,
                   float boundary_weight)
{
    return guard([&] {
        checked(g);
        mc_smooth(g->mc, g->stream, nonboundary_iterations, nonboundary_weight, boundary_iterations, boundary_weight);
    });
}

int vf_mesh_export_binm(vf_grid* g, uint32_t label_index, uint8_t* out, size_t cap, size_t* size)
{
    return guard([&] {
        checked(g);
        if (!g->mc.valid) fail(VF_ERR_INVALID_ARG, "no marching-cubes result: call vf_marching_cubes first");
        const uint32_t nl = g->mc.nl;
        if (label_index >= nl) fail(VF_ERR_INVALID_ARG, "label index %u out of range (%u labels)", label_index, nl);
        std::vector<uint32_t> voff(nl + 1), toff(nl + 1);
        grid_mesh_read(g, nullptr, nullptr, nullptr, nullptr, voff.data(), toff.data());
        const uint32_t nv = voff[label_index + 1] - voff[label_index], nt = toff[label_index + 1] - toff[label_index];
        const size_t total = 4 + (size_t)nv * 64 + 4 + (size_t)nt * 16;
        if (size) *size = total;
        if (!out || cap < total) return;
        std::vector<float> xyz(3ull * nv);
        std::vector<uint32_t> tri(3ull * nt);
        grid_mesh_slice(g, voff[label_index], nv, toff[label_index], nt, xyz.data(), tri.data());
        std::memset(out, 0, total);
        std::memcpy(out, &nv, 4);
        for (uint32_t v = 0; v < nv; ++v) std::memcpy(out + 4 + (size_t)v * 64, &xyz[3ull * v], 12);
        uint8_t* f = out + 4 + (size_t)nv * 64;
        std::memcpy(f, &nt, 4);
        for (uint32_t t = 0; t < nt; ++t) std::memcpy(f + 4 + (size_t)t * 16, &tri[3ull * t], 12);
    });
}

int vf_grid_export_rle(vf_grid* g, uint8_t* out, size_t cap, size_t* size)
{
    // RegularGrid::exportRLE (RegularGrid.cpp:672-714): uvec3 dims, then maximal runs
    return guard([&] {
        checked(g);
        std::vector<uint16_t> h(g->V);
        int rc = vf_grid_read(g, h.data());
        if (rc != VF_OK) throw Exception{ rc };
        size_t pos = 0;
        auto put = [&](const void* p, size_t n) {
            if (out && pos + n <= cap) std::memcpy(out + pos, p, n);
            pos += n;
        };
        put(g->dims, 12);
        uint64_t i = 0;
        while (i < g->V) {
            uint16_t v = h[i];
            uint64_t j = i;
            while (j < g->V && h[j] == v && j - i < 0xFFFFFFFFull) ++j;
            uint32_t n = (uint32_t)(j - i);
            put(&v, 2);
            put(&n, 4);
            i = j;
        }
        if (size) *size = pos;
    });
}

int vf_grid_import_rle(vf_grid* g, const uint8_t* in, size_t n)
{
    return guard([&] {
        checked(g);
        if (!in || n < 12 || (n - 12) % 6) fail(VF_ERR_INVALID_ARG, "RLE stream is truncated");
        uint32_t d[3];
        std::memcpy(d, in, 12);
        if (d[0] != g->dims[0] || d[1] != g->dims[1] || d[2] != g->dims[2])
            fail(VF_ERR_INVALID_ARG, "RLE dims %ux%ux%u differ from the grid's", d[0], d[1], d[2]);
        std::vector<uint16_t> h(g->V);
        uint64_t pos = 0;
        for (size_t o = 12; o < n; o += 6) {
            uint16_t v;
            uint32_t c;
            std::memcpy(&v, in + o, 2);
            std::memcpy(&c, in + o + 2, 4);
            if (pos + c > g->V) fail(VF_ERR_INVALID_ARG, "RLE runs exceed the grid size");
            std::fill(h.begin() + (ptrdiff_t)pos, h.begin() + (ptrdiff_t)(pos + c), v);
            pos += c;
        }
        if (pos != g->V) fail(VF_ERR_INVALID_ARG, "RLE runs cover %llu of %llu cells", (unsigned long long)pos,
                              (unsigned long long)g->V);
        int rc = vf_grid_write(g, h.data());
        if (rc != VF_OK) throw Exception{ rc };
    });
}

}  // extern "C"

// ---- device memory helpers (mesh residency for jobs with mesh_on_device) ----
extern "C" {

int vf_device_alloc(size_t bytes, void** ptr)
{
    return guard([&] {
        if (!ptr) fail(VF_ERR_INVALID_ARG, "ptr is NULL");
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *ptr = nullptr;
            fail(VF_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed", bytes);
        }
    });
}

int vf_device_free(void* ptr)
{
    return guard([&] {
        if (ptr) VF_HIP(hipFree(ptr));
    });
}

int vf_memcpy_htod(void* dst, const void* src, size_t bytes)
{
    return guard([&] {
        if (bytes && (!dst || !src)) fail(VF_ERR_INVALID_ARG, "NULL pointer");
        if (bytes) VF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

}  // extern "C"
