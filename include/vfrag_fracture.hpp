// vfrag_fracture.hpp — C++ mirror of VoxelFragmentML's fracture API over the
// libvfrag C ABI (vfrag.h), so code written against the reference's
// RegularGrid / FractureParameters / fracturer::* / Seeder calls the MI355X
// path with the same names, enums, argument meaning and seed layout.
//
// Reference surfaces mirrored (paths relative to MeshFragments/Source):
//   RegularGrid                      DataStructures/RegularGrid.h:15-367
//   FractureParameters               Graphics/Core/FractureParameters.h:14-140
//   fracturer::DistanceFunction      Fracturer/Fracturer.h:10-14
//   fracturer::Fracturer             Fracturer/Fracturer.h:28-57
//   fracturer::NaiveFracturer        Fracturer/NaiveFracturer.h:13-70 (Singleton)
//   fracturer::FloodFracturer        Fracturer/FloodFracturer.h:13-75 (Singleton)
//   fracturer::Seeder::uniform       Fracturer/Seeder.h:24, 69-76
//   fracturer::Seeder::nearSeeds     Fracturer/Seeder.h:62, Seeder.cpp:49-113
//   fracturer::Seeder::mergeSeeds    Fracturer/Seeder.h:67, Seeder.cpp:115-152
//
// Differences a caller sees: the grid lives in device memory and data() /
// at() / set() work on a host mirror that is synchronised explicitly, exactly
// as the reference's updateGrid() / updateSSBO() pair does for its SSBO
// (RegularGrid.cpp:505-514) — the mirror is refreshed lazily after device
// work; fill() takes the mesh as arrays instead of a Model3D*; errors are
// thrown as vfrag::Error (a std::runtime_error) carrying the library message.
// Everything is header-only and links against libvfrag.so.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "vfrag.h"

namespace vfrag {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
    if (rc != VF_OK) throw Error(rc, vf_last_error());
}

struct ivec3 {
    int x, y, z;
    ivec3(int v = 0) : x(v), y(v), z(v) {}
    ivec3(int a, int b, int c) : x(a), y(b), z(c) {}
};
struct uvec3 {
    uint32_t x, y, z;
};
struct vec3 {
    float x, y, z;
};
// glm::uvec4 seed {x, y, z, label} (Seeder.h:69-76)
struct uvec4 {
    uint32_t x, y, z, w;
};
static_assert(sizeof(uvec4) == sizeof(vf_seed), "seed layout");

// AABB (Geometry/3D/AABB.h): min / max corners
struct AABB {
    vec3 _min, _max;
    vec3 min() const { return _min; }
    vec3 max() const { return _max; }
};

// The mesh a RegularGrid is filled from: FP32 vertex positions and triangles
// (the positions and faces of Model3D's VertexGPUData / FaceGPUData,
// Model3D.h:37-59).
struct TriangleMesh {
    std::vector<float> xyz;     // 3 per vertex
    std::vector<uint32_t> tri;  // 3 per face
};

// FractureParameters (FractureParameters.h:14-140): the enums and the fields
// the voxelize -> fracture path reads, with the reference's defaults.
struct FractureParameters {
    enum FractureAlgorithm : uint8_t { NAIVE, FLOOD, VORONOI, BASE_ALGORITHMS };
    enum DistanceFunction : uint8_t { EUCLIDEAN, MANHATTAN, CHEBYSHEV, DISTANCE_FUNCTIONS };
    enum NeighbourhoodType { VON_NEUMANN, MOORE, NUM_NEIGHBOURHOODS };
    enum ErosionType { SQUARE, ELLIPSE, CROSS, NUM_EROSION_CONVOLUTIONS };

    int _fractureAlgorithm = FLOOD;
    int _distanceFunction = CHEBYSHEV;
    int _neighbourhoodType = VON_NEUMANN;  // the reference's flood ignores it (the stencil follows the distance)
    int _numSeeds = 8;
    int _numExtraSeeds = 16;               // BOTH seeds merged into the fragments as sub-seeds (Seeder::mergeSeeds)
    int _mergeSeedsDistanceFunction = EUCLIDEAN;
    int _numImpacts = 0;                   // Seeder::nearSeeds impacts; 0 = uniform fragment seeds only
    int _biasSeeds = 32;                   // nearSeeds: seeds over all impacts
    int _biasFocus = 5;                    // nearSeeds: uniform draws summed per axis (spreading)
    // erosion instead of detectBoundaries at the end of fractureModel (CADScene.cpp:680-689)
    bool _erode = false;
    int _erosionConvolution = ELLIPSE;
    float _erosionProbability = .5f;
    int _erosionIterations = 3;
    int _erosionSize = 3;
    float _erosionThreshold = 0.5f;
    bool _removeIsolatedRegions = true;    // the reference's GPU pass is an identity; see NaiveFracturer below
    bool _launchGPU = true;
    int _seed = 80;
    int _marchingCubesSubdivisions = 1;
    ivec3 _voxelizationSize = ivec3(128);
    // marching-cubes smoothing: iterations = (unsigned)(max _voxelizationSize edge * factor)
    float _boundaryMCWeight = 0.2f, _boundaryMCIterations = 0.048f;
    float _nonBoundaryMCWeight = 0.9f, _nonBoundaryMCIterations = 0.048f;

    // Flood boundary noise: node weights 1 + hash(noise seed, cell) in
    // [0, amplitude].  Not in the reference (amplitude 0 = plain BFS order).
    uint64_t _noiseSeed = 0;
    uint32_t _noiseAmplitude = 0;
};

// RegularGrid (RegularGrid.h:15-367).  uint16 cells, index x*ny*nz + y*nz + z.
class RegularGrid {
public:
    static constexpr uint16_t VOXEL_EMPTY = 0;  // RegularGrid.h:12
    static constexpr uint16_t VOXEL_FREE = 1;   // RegularGrid.h:13
    static constexpr unsigned MASK_POSITION = 15;
    using CellGrid = uint16_t;

    RegularGrid(const AABB& aabb, const ivec3& subdivisions)
    {
        const uint32_t dims[3] = { (uint32_t)subdivisions.x, (uint32_t)subdivisions.y, (uint32_t)subdivisions.z };
        const float mn[3] = { aabb._min.x, aabb._min.y, aabb._min.z }, mx[3] = { aabb._max.x, aabb._max.y, aabb._max.z };
        check(vf_grid_create(dims, mn, mx, &g_));
        aabb_ = aabb;
        dims_ = { dims[0], dims[1], dims[2] };
    }
    RegularGrid(const RegularGrid&) = delete;
    RegularGrid& operator=(const RegularGrid&) = delete;
    virtual ~RegularGrid() { vf_grid_destroy(g_); }

    // RegularGrid::setAABB (RegularGrid.cpp:426-441): new box and dims, grid cleared
    void setAABB(const AABB& aabb, const ivec3& gridDims)
    {
        const uint32_t dims[3] = { (uint32_t)gridDims.x, (uint32_t)gridDims.y, (uint32_t)gridDims.z };
        const float mn[3] = { aabb._min.x, aabb._min.y, aabb._min.z }, mx[3] = { aabb._max.x, aabb._max.y, aabb._max.z };
        check(vf_grid_set_aabb(g_, mn, mx, dims));
        aabb_ = aabb;
        dims_ = { dims[0], dims[1], dims[2] };
        host_.clear();
        host_valid_ = false;
    }
    AABB getAABB() const { return aabb_; }
    uvec3 getNumSubdivisions() const { return dims_; }
    size_t length() const { return (size_t)dims_.x * dims_.y * dims_.z; }

    // RegularGrid::fill(Model3D*) (RegularGrid.cpp:173-212): occupied cells become FREE
    void fill(const TriangleMesh& mesh)
    {
        check(vf_voxelize(g_, mesh.xyz.data(), (uint32_t)(mesh.xyz.size() / 3), mesh.tri.data(),
                          (uint32_t)(mesh.tri.size() / 3)));
        host_valid_ = false;
    }

    // RegularGrid::homogenize (RegularGrid.cpp:533-541)
    void homogenize() { check(vf_homogenize(g_)); host_valid_ = false; }
    // RegularGrid::detectBoundaries (RegularGrid.cpp:64-80)
    void detectBoundaries(int boundarySize) { check(vf_detect_boundaries(g_, boundarySize)); host_valid_ = false; }
    // RegularGrid::undoMask (RegularGrid.cpp:488-503)
    void undoMask() { check(vf_undo_mask(g_)); host_valid_ = false; }
    // RegularGrid::erode (RegularGrid.cpp:82-159); the noise buffer is a
    // counter-based draw keyed by noiseSeed
    void erode(int erosionType, uint32_t convolutionSize, uint16_t numIterations, float erosionProbability,
               float erosionThreshold, uint64_t noiseSeed = 0)
    {
        check(vf_erode(g_, erosionType, convolutionSize, numIterations, erosionProbability, erosionThreshold, noiseSeed));
        host_valid_ = false;
    }
    // RegularGrid::removeIsolatedRegions (RegularGrid.cpp:1006-1015)
    void removeIsolatedRegions() { check(vf_remove_isolated_cells(g_)); host_valid_ = false; }

    // RegularGrid::countValues (RegularGrid.cpp:601-625): count per label > FREE
    // (boundary bit removed); returns the number of counted cells
    size_t countValues(std::unordered_map<uint16_t, unsigned>& values)
    {
        std::vector<uint32_t> counts(65536);
        check(vf_histogram(g_, counts.data(), 65536));
        size_t total = 0;
        for (uint32_t v = 2; v < 65536; ++v)
            if (counts[v]) {
                values[(uint16_t)v] += counts[v];
                total += counts[v];
            }
        return total;
    }

    // The FragmentationProcedure::FragmentMetadata entries (FragmentationProcedure.h:22-37)
    // RegularGrid::toTriangleMesh fills (RegularGrid.cpp:443-471): one per label in
    // ascending order, percentage = voxels / occupied voxels
    std::vector<vf_fragment_meta> fragmentMetadata()
    {
        std::vector<uint32_t> counts(65536);
        check(vf_histogram(g_, counts.data(), 65536));
        uint32_t d[3], n = 0;
        check(vf_grid_dims(g_, d));
        check(vf_fragment_metadata(counts.data(), d, nullptr, 0, &n));
        std::vector<vf_fragment_meta> out(n);
        check(vf_fragment_metadata(counts.data(), d, out.data(), n, &n));
        return out;
    }

    // RegularGrid::numOccupiedVoxels
    unsigned numOccupiedVoxels()
    {
        uint64_t n = 0;
        check(vf_grid_count_occupied(g_, &n));
        return (unsigned)n;
    }

    // Host mirror (RegularGrid::data, at, set, isOccupied, isEmpty)
    CellGrid* data()
    {
        pull();
        return host_.data();
    }
    uint16_t at(int x, int y, int z) { pull(); return host_[index(x, y, z)]; }
    void set(int x, int y, int z, uint16_t i) { pull(); host_[index(x, y, z)] = i; host_dirty_ = true; }
    bool isOccupied(int x, int y, int z) { return at(x, y, z) != VOXEL_EMPTY; }
    bool isEmpty(int x, int y, int z) { return at(x, y, z) == VOXEL_EMPTY; }

    // RegularGrid::updateSSBO (RegularGrid.cpp:511-514): host mirror -> device
    void updateSSBO()
    {
        if (!host_valid_) return;
        check(vf_grid_write(g_, host_.data()));
        host_dirty_ = false;
    }
    // RegularGrid::updateGrid (RegularGrid.cpp:505-509): device -> host mirror
    void updateGrid()
    {
        host_.resize(length());
        check(vf_grid_read(g_, host_.data()));
        host_valid_ = true;
        host_dirty_ = false;
    }

    // RegularGrid::exportRLE (RegularGrid.cpp:672-714), to memory
    std::vector<uint8_t> exportRLE()
    {
        sync_device();
        size_t n = 0;
        check(vf_grid_export_rle(g_, nullptr, 0, &n));
        std::vector<uint8_t> out(n);
        check(vf_grid_export_rle(g_, out.data(), n, &n));
        return out;
    }

    // Marching cubes over every label (RegularGrid::toTriangleMesh,
    // RegularGrid.cpp:443-486): one mesh per label, FP32 world coordinates.
    // With parameters, the meshes are smoothed as MarchingCubes.cpp:368-401
    // does: a non-boundary pass, then a boundary pass, each of
    // (unsigned)(max _voxelizationSize edge * iteration factor) steps.
    struct LabelMesh {
        uint16_t label;
        std::vector<float> xyz;
        std::vector<uint32_t> tri;
        std::vector<uint8_t> boundary;  // per triangle: the boundary bit of its cell
    };
    std::vector<LabelMesh> toTriangleMesh() { return triangulate(nullptr); }
    std::vector<LabelMesh> toTriangleMesh(const FractureParameters& params) { return triangulate(&params); }

    // One mesh of the last toTriangleMesh call (index into its list) in the
    // layout CADModel::saveBinary writes (CADModel.cpp:839-855), as the
    // reference's dataset export does: positions set, other vertex fields zero.
    std::vector<uint8_t> exportBinaryMesh(uint32_t meshIndex)
    {
        size_t n = 0;
        check(vf_mesh_export_binm(g_, meshIndex, nullptr, 0, &n));
        std::vector<uint8_t> out(n);
        check(vf_mesh_export_binm(g_, meshIndex, out.data(), n, &n));
        return out;
    }

    vf_grid* handle()
    {
        sync_device();
        return g_;
    }
    // the device owns the grid again (after device work through handle())
    void invalidateHost() { host_valid_ = false; }

private:
    std::vector<LabelMesh> triangulate(const FractureParameters* params)
    {
        sync_device();
        vf_mesh_info info{};
        check(vf_marching_cubes(g_, VF_ALL_LABELS, &info));
        if (params) {
            const ivec3& v = params->_voxelizationSize;
            const unsigned maxVoxels = (unsigned)std::max(v.x, std::max(v.y, v.z));
            check(vf_mesh_smooth(g_, (unsigned)(maxVoxels * params->_nonBoundaryMCIterations),
                                 params->_nonBoundaryMCWeight,
                                 (unsigned)(maxVoxels * params->_boundaryMCIterations), params->_boundaryMCWeight));
        }
        std::vector<float> xyz(3ull * info.num_vertices);
        std::vector<uint32_t> tri(3ull * info.num_triangles);
        std::vector<uint8_t> flag(info.num_triangles);
        std::vector<uint16_t> labels(info.num_labels);
        std::vector<uint32_t> voff(info.num_labels + 1), toff(info.num_labels + 1);
        check(vf_mesh_read(g_, xyz.data(), tri.data(), flag.data(), labels.data(), voff.data(), toff.data()));
        std::vector<LabelMesh> out(info.num_labels);
        for (uint32_t l = 0; l < info.num_labels; ++l) {
            out[l].label = labels[l];
            out[l].xyz.assign(xyz.begin() + 3ull * voff[l], xyz.begin() + 3ull * voff[l + 1]);
            out[l].tri.assign(tri.begin() + 3ull * toff[l], tri.begin() + 3ull * toff[l + 1]);
            out[l].boundary.assign(flag.begin() + toff[l], flag.begin() + toff[l + 1]);
        }
        return out;
    }

    size_t index(int x, int y, int z) const { return ((size_t)x * dims_.y + (size_t)y) * dims_.z + (size_t)z; }
    void pull()
    {
        if (!host_valid_) updateGrid();
    }
    // host edits not yet uploaded reach the device before device work
    void sync_device()
    {
        if (host_valid_ && host_dirty_) updateSSBO();
    }

    vf_grid* g_ = nullptr;
    AABB aabb_{};
    uvec3 dims_{ 0, 0, 0 };
    std::vector<uint16_t> host_;
    bool host_valid_ = false, host_dirty_ = false;

    friend class FractureAccess;
};

class FractureAccess {
public:
    static vf_grid* device(RegularGrid& g) { return g.handle(); }
    static void touched(RegularGrid& g) { g.invalidateHost(); }
};

namespace fracturer {

// Fracturer.h:10-14
enum DistanceFunction : uint32_t {
    EUCLIDEAN_DISTANCE = 0,
    MANHATTAN_DISTANCE = 1,
    CHEBYSHEV_DISTANCE = 2
};

// Fracturer.h:28-57
class Fracturer {
public:
    virtual ~Fracturer() = default;
    virtual void build(RegularGrid& grid, const std::vector<uvec4>& seeds, FractureParameters* fractParameters) = 0;
    virtual void destroy() = 0;
    virtual void init(FractureParameters* fractParameters) = 0;
    virtual void prepareSSBOs(FractureParameters* fractParameters) = 0;
    virtual bool setDistanceFunction(DistanceFunction dfunc) = 0;
};

// Utilities/Singleton.h: one instance per process
template <class T>
class Singleton {
public:
    static T* getInstance()
    {
        static T instance;
        return &instance;
    }
};

namespace detail {
inline bool valid_distance(uint32_t d) { return d <= CHEBYSHEV_DISTANCE; }
}  // namespace detail

// NaiveFracturer (NaiveFracturer.h:13-70; build :215-225, buildCPU :26-68).
// Every non-EMPTY cell takes the label of the nearest seed, ties to the first
// seed in list order.  _removeIsolatedRegions follows removeIsolatedRegionsCPU
// (:111-150) only when removeIsolatedRegionsCPU(true) is set: the reference's
// default GPU pass is an identity (its output buffer starts as a copy).
class NaiveFracturer : public Singleton<NaiveFracturer>, public Fracturer {
public:
    void build(RegularGrid& grid, const std::vector<uvec4>& seeds, FractureParameters* p) override
    {
        const uint32_t d = p ? (uint32_t)p->_distanceFunction : dfunc_;
        if (!detail::valid_distance(d)) throw Error(VF_ERR_INVALID_ARG, "Invalid distance function");
        vf_grid* g = FractureAccess::device(grid);
        check(vf_fracture_naive(g, reinterpret_cast<const vf_seed*>(seeds.data()), (uint32_t)seeds.size(), (int)d));
        if (cpu_isolated_ && p && p->_removeIsolatedRegions)
            check(vf_remove_isolated_regions(g, reinterpret_cast<const vf_seed*>(seeds.data()), (uint32_t)seeds.size()));
        FractureAccess::touched(grid);
    }
    void destroy() override {}
    void init(FractureParameters*) override {}
    void prepareSSBOs(FractureParameters*) override {}
    bool setDistanceFunction(DistanceFunction dfunc) override
    {
        if (!detail::valid_distance(dfunc)) return false;
        dfunc_ = dfunc;
        return true;
    }
    void removeIsolatedRegionsCPU(bool on) { cpu_isolated_ = on; }

private:
    friend class Singleton<NaiveFracturer>;
    NaiveFracturer() = default;
    uint32_t dfunc_ = EUCLIDEAN_DISTANCE;
    bool cpu_isolated_ = false;
};

// FloodFracturer (FloodFracturer.h:13-75; build FloodFracturer.cpp:98-191).
// Region growing from the seeds; the stencil follows the distance function
// (MANHATTAN -> 6 neighbours, else 26, FloodFracturer.cpp:114); claims are
// settled by (distance, seed index); fragment ids are the low 8 bits of the
// seed label (labels 2..255).
class FloodFracturer : public Singleton<FloodFracturer>, public Fracturer {
public:
    void build(RegularGrid& grid, const std::vector<uvec4>& seeds, FractureParameters* p) override
    {
        const uint32_t d = p ? (uint32_t)p->_distanceFunction : dfunc_;
        if (!detail::valid_distance(d)) throw Error(VF_ERR_INVALID_ARG, "Invalid distance function");
        vf_grid* g = FractureAccess::device(grid);
        check(vf_fracture_flood(g, reinterpret_cast<const vf_seed*>(seeds.data()), (uint32_t)seeds.size(), (int)d,
                                VF_NEIGH_FROM_DISTANCE, p ? p->_noiseSeed : 0ull, p ? p->_noiseAmplitude : 0u));
        FractureAccess::touched(grid);
    }
    void destroy() override {}
    void init(FractureParameters*) override {}
    void prepareSSBOs(FractureParameters*) override {}
    bool setDistanceFunction(DistanceFunction dfunc) override
    {
        if (!detail::valid_distance(dfunc)) return false;
        dfunc_ = dfunc;
        return true;
    }

private:
    friend class Singleton<FloodFracturer>;
    FloodFracturer() = default;
    uint32_t dfunc_ = EUCLIDEAN_DISTANCE;
};

// Seeder (Seeder.h:24-76).  uniform() draws nseeds distinct candidate cells
// with a counter-based generator keyed by randomSeed (the reference's
// std/boost generators are not reproducible across platforms) and returns
// them in (x, y, z) order labelled 2..nseeds+1.
class Seeder {
public:
    static constexpr uint32_t VOXEL_ID_POSITION = 8;  // Seeder.h:12: sub-seed ids live above the 8-bit fragment id
    enum Location { INNER, OUTER, BOTH };

    // nearSeeds(grid, frags, numImpacts, numSeeds, spreading): frags followed
    // by impact seeds drawn around randomly chosen fragment seeds (occupied
    // surface cells within half the smallest grid edge), in (x, y, z) order
    // with labels after the last fragment's.  Counter-based, keyed by
    // randomSeed; the reference's endless retry loop becomes an error after
    // 10^6 draws per impact.
    static std::vector<uvec4> nearSeeds(RegularGrid& grid, const std::vector<uvec4>& frags, unsigned numImpacts,
                                        unsigned numSeeds, unsigned spreading, uint64_t randomSeed = 0)
    {
        std::vector<uvec4> out(frags.size() + numSeeds);
        uint32_t n = 0;
        check(vf_seed_near(FractureAccess::device(grid), reinterpret_cast<const vf_seed*>(frags.data()),
                           (uint32_t)frags.size(), numImpacts, numSeeds, spreading, randomSeed,
                           reinterpret_cast<vf_seed*>(out.data()), &n));
        out.resize(n);
        return out;
    }

    // mergeSeeds(frags, seeds, dfunc): each seed takes its nearest fragment
    // seed's label plus that fragment's running sub-seed id << 8.
    static void mergeSeeds(const std::vector<uvec4>& frags, std::vector<uvec4>& seeds, DistanceFunction dfunc)
    {
        check(vf_seed_merge(reinterpret_cast<const vf_seed*>(frags.data()), (uint32_t)frags.size(),
                            reinterpret_cast<vf_seed*>(seeds.data()), (uint32_t)seeds.size(), (int)dfunc));
    }

    static std::vector<uvec4> uniform(RegularGrid& grid, unsigned int nseeds, uint64_t randomSeed,
                                      Location location = OUTER)
    {
        static_assert(INNER == VF_SEED_INNER && OUTER == VF_SEED_OUTER && BOTH == VF_SEED_BOTH, "Seeder::Location");
        std::vector<uvec4> out(nseeds);
        check(vf_seed_uniform(FractureAccess::device(grid), nseeds, randomSeed, (int)location,
                              reinterpret_cast<vf_seed*>(out.data())));
        return out;
    }
};

}  // namespace fracturer
}  // namespace vfrag

#ifdef VFRAG_REFERENCE_NAMES
// Expose the reference's global names for code written against it.
using vfrag::FractureParameters;
using vfrag::RegularGrid;
namespace fracturer {
using namespace vfrag::fracturer;
}
#endif
