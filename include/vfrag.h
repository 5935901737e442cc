/*
 * vfrag.h — C ABI of libvfrag, the MI355X (gfx950) implementation of
 * VoxelFragmentML's voxelize -> fracture path.
 *
 * Every function is extern "C", takes plain pointers and sizes, and returns an
 * int status (VF_OK or a negative VF_ERR_*); vf_last_error() gives the
 * thread-local message of the last failure.  No exceptions cross this ABI.
 * Host arrays are borrowed for the duration of a call; grids, batches and
 * servers are opaque handles owned by the library.  Grid values follow the
 * reference: uint16 per cell, linear index x*ny*nz + y*nz + z, 0 = EMPTY,
 * 1 = FREE, >= 2 = fragment label, bit 15 = boundary mask.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference root, MF = MeshFragments).
 */
#ifndef VFRAG_H
#define VFRAG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define VF_OK 0
#define VF_ERR_INVALID_ARG (-1)   /* bad dims, pointer, seed or enum value */
#define VF_ERR_OUT_OF_MEMORY (-2) /* device or host allocation failed */
#define VF_ERR_DEVICE (-3)        /* a HIP call failed */
#define VF_ERR_TOO_MANY_SEEDS (-4)/* flood labels must fit the 8-bit id field (2..255) */
#define VF_ERR_BUDGET (-5)        /* job does not fit the server's HBM budget */
#define VF_ERR_CANCELLED (-6)     /* job cancelled before completion */
#define VF_ERR_NOT_READY (-7)     /* poll: job still running */
#define VF_ERR_OVERFLOW (-8)      /* a packed key or counter would overflow */
#define VF_ERR_SEEDING (-9)       /* fewer candidate cells than seeds (SeederSearchError) */
#define VF_ERR_POISON (-10)       /* job flagged as poison by the caller (failure-path tests) */

/* FractureParameters::DistanceFunction / fracturer::DistanceFunction
 * (MF/Source/Graphics/Core/FractureParameters.h:21, MF/Source/Fracturer/Fracturer.h:10-14) */
#define VF_EUCLIDEAN 0
#define VF_MANHATTAN 1
#define VF_CHEBYSHEV 2

/* FractureParameters::FractureAlgorithm (FractureParameters.h:18) */
#define VF_NAIVE 0
#define VF_FLOOD 1

/* flood neighbourhood: VF_NEIGH_FROM_DISTANCE follows FloodFracturer::build
 * (FloodFracturer.cpp:114: MANHATTAN -> 6 neighbours, anything else -> 26) */
#define VF_NEIGH_FROM_DISTANCE 0
#define VF_NEIGH_VON_NEUMANN 6
#define VF_NEIGH_MOORE 26

/* Seeder::Location { INNER, OUTER, BOTH } (MF/Source/Fracturer/Seeder.h:24) */
#define VF_SEED_INNER 0
#define VF_SEED_OUTER 1
#define VF_SEED_BOTH 2
/* FractureParameters::ErosionType (FractureParameters.h:27) */
#define VF_EROSION_SQUARE 0
#define VF_EROSION_ELLIPSE 1
#define VF_EROSION_CROSS 2

#define VF_ALL_LABELS 0xFFFFFFFFu

typedef struct vf_grid vf_grid;

/* glm::uvec4 {x, y, z, label} seed layout (Seeder.h:69-76, Fracturer.h:35) */
typedef struct {
    uint32_t x, y, z, w;
} vf_seed;

/* ---- library ------------------------------------------------------------- */
const char* vf_last_error(void);
int vf_version(void);
/* Number of visible HIP devices (0 when none; never fails). */
int vf_device_count(void);

/* ---- grid handles ---------------------------------------------------------- */
/* RegularGrid(const AABB&, const ivec3&) (MF/Source/DataStructures/RegularGrid.h:152,
 * RegularGrid.cpp:21-27): device-resident uint16 grid, cleared to EMPTY. */
int vf_grid_create(const uint32_t dims[3], const float aabb_min[3], const float aabb_max[3], vf_grid** out);
int vf_grid_destroy(vf_grid* g);
/* RegularGrid::setAABB (RegularGrid.cpp:426-441): new box and dims, grid cleared. */
int vf_grid_set_aabb(vf_grid* g, const float aabb_min[3], const float aabb_max[3], const uint32_t dims[3]);
/* RegularGrid::getNumSubdivisions (RegularGrid.h:314) */
int vf_grid_dims(const vf_grid* g, uint32_t dims[3]);
/* RegularGrid::updateGrid / data() (RegularGrid.cpp:505-509, :518-521): device -> host copy */
int vf_grid_read(vf_grid* g, uint16_t* out);
/* RegularGrid::updateSSBO (RegularGrid.cpp:511-514): host -> device copy */
int vf_grid_write(vf_grid* g, const uint16_t* in);
/* uint8 occupancy (1 = cell not EMPTY) */
int vf_occupancy_read(vf_grid* g, uint8_t* out);
/* number of non-EMPTY cells (RegularGrid::numOccupiedVoxels, RegularGrid.h:233) */
int vf_grid_count_occupied(vf_grid* g, uint64_t* out);
/* device pointer to the uint16 cell array (RegularGrid::ssbo(), RegularGrid.h:263) */
int vf_grid_device_ptr(vf_grid* g, void** ptr);
/* block until every operation queued on the grid has finished */
int vf_grid_synchronize(vf_grid* g);

/* ---- voxelization ---------------------------------------------------------- */
/* RegularGrid::fill(Model3D*) (RegularGrid.cpp:173-212): conservative FP32
 * triangle/box SAT shell plus z-parity interior; cells become FREE.
 * xyz: nv*3 floats, tri: nt*3 vertex indices (host memory). */
int vf_voxelize(vf_grid* g, const float* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt);
/* same, with the mesh already in device memory */
int vf_voxelize_device(vf_grid* g, const float* d_xyz, uint32_t nv, const uint32_t* d_tri, uint32_t nt);

/* ---- seeding --------------------------------------------------------------- */
/* Seeder::uniform (MF/Source/Fracturer/Seeder.cpp:154-208), counter-based: S
 * distinct candidate cells drawn uniformly, in (x,y,z) order, labels 2..S+1. */
int vf_seed_uniform(vf_grid* g, uint32_t S, uint64_t seed, int location, vf_seed* out);
/* Seeder::nearSeeds (Seeder.cpp:49-113), counter-based: impacts times, pick a
 * fragment seed and a count of 1..pending-1 new seeds (of num_seeds in all),
 * then draw cells around it by the biased sum of `spreading` uniform offsets
 * per axis (RandomUtilities::getBiasedRandomInt), wrapped around the grid, and
 * keep those within min(dims)/2 of it that are occupied, on the surface
 * (RegularGrid::isBoundary) and new.  out (F + num_seeds entries) receives
 * the F fragment seeds, then the new seeds in (x,y,z) order labelled after the
 * last fragment's label; *count the number written.  spreading 1..64 and at
 * most the smallest grid edge; F >= 1.  An impact that finds too few cells in
 * 10^6 draws fails with VF_ERR_INVALID_ARG (the reference loops forever). */
int vf_seed_near(vf_grid* g, const vf_seed* frags, uint32_t F, uint32_t impacts, uint32_t num_seeds,
                 uint32_t spreading, uint64_t seed, vf_seed* out, uint32_t* count);
/* Seeder::mergeSeeds (Seeder.cpp:115-152), host only: each of seeds[0..N)
 * takes the label of its nearest fragment seed (distance VF_*, first on ties)
 * with that fragment's running sub-seed id (1, 2, ...) in bits 8..15.
 * Fragment labels must fit 8 bits; more than 255 sub-seeds on one fragment is
 * VF_ERR_TOO_MANY_SEEDS (the reference's uint16 cell would wrap).  The flood
 * then keeps, per fragment, only the connected part holding its smallest id
 * and regrows the rest (vf_fracture_flood). */
int vf_seed_merge(const vf_seed* frags, uint32_t F, vf_seed* seeds, uint32_t N, int distance);

/* ---- fracture -------------------------------------------------------------- */
/* NaiveFracturer::build (MF/Source/Fracturer/NaiveFracturer.cpp:215-225,
 * buildCPU :26-68, naiveFracturer-comp.glsl:19-42): every non-EMPTY cell takes
 * the label of the nearest seed, ties to the lowest seed index. */
int vf_fracture_naive(vf_grid* g, const vf_seed* seeds, uint32_t S, int distance);
/* NaiveFracturer::removeIsolatedRegionsCPU (NaiveFracturer.cpp:111-150) */
int vf_remove_isolated_regions(vf_grid* g, const vf_seed* seeds, uint32_t S);
/* FloodFracturer::build (MF/Source/Fracturer/FloodFracturer.cpp:98-191): region
 * growing from the seeds, deterministic (distance, seed index) claims, node
 * weights 1 + hash noise in [0, noise_amp].  Labels are fragment | sub-seed
 * id << 8 (fragment 2..255): when a fragment's seeds carry different ids,
 * each connected part of the fragment keeps the smallest id of its seeds, the
 * parts without the fragment's smallest id are freed and regrown from all
 * other labelled cells at once (FloodFracturer.cpp:134-177, ties to the lower
 * owning seed index).  Cells end with the fragment id (unmaskRightMost). */
int vf_fracture_flood(vf_grid* g, const vf_seed* seeds, uint32_t S, int distance, int neighbourhood,
                      uint64_t noise_seed, uint32_t noise_amp);

/* ---- post-fracture --------------------------------------------------------- */
/* RegularGrid::detectBoundaries (RegularGrid.cpp:64-80, detectBoundaries-comp.glsl:18-43) */
int vf_detect_boundaries(vf_grid* g, int size);
/* RegularGrid::undoMask (RegularGrid.cpp:488-503) */
int vf_undo_mask(vf_grid* g);
/* RegularGrid::erode (RegularGrid.cpp:82-159): iterations times
 * detectBoundaries(1) then one erosion step (erodeGrid-comp.glsl:26-59); then
 * removeIsolatedRegions.  A cell above FREE whose noise draw
 * (hash32(noise_seed, cell) >> 8) / 2^24 is below probability becomes EMPTY
 * when the mask-weighted count of window cells with its raw value (boundary
 * bit included), over the in-bounds window cells, is below activations *
 * threshold (FP32).  The window is size^3 (even sizes grow by one, at most
 * 15): VF_EROSION_SQUARE, _CROSS or _ELLIPSE, as the reference builds them.
 * Leaves the boundary bits of the last detection set. */
int vf_erode(vf_grid* g, int convolution, uint32_t size, uint32_t iterations, float probability, float threshold,
             uint64_t noise_seed);
/* RegularGrid::removeIsolatedRegions (RegularGrid.cpp:1006-1015,
 * removeIsolatedRegionsGrid-comp.glsl:16-38): a cell with fewer than 6 other
 * cells of its raw value in its clamped 3x3x3 window becomes EMPTY (read from
 * the grid before the pass: the shader's in-place update is thread-order
 * dependent). */
int vf_remove_isolated_cells(vf_grid* g);
/* RegularGrid::homogenize (RegularGrid.cpp:533-541) */
int vf_homogenize(vf_grid* g);
/* RegularGrid::countValues (RegularGrid.cpp:601-625): counts[v & 0x7FFF] for v > FREE.
 * counts has ncounts entries (65536 covers every label); values >= ncounts are
 * not reported. */
int vf_histogram(vf_grid* g, uint32_t* counts, uint32_t ncounts);

/* The per-fragment metadata RegularGrid::toTriangleMesh fills from countValues
 * (RegularGrid.cpp:443-471; FragmentationProcedure::FragmentMetadata,
 * FragmentationProcedure.h:22-37; written by CADScene::exportMetadata,
 * CADScene.cpp:580-622): one entry per value above FREE (counts[0] and counts[1]
 * are ignored) with a nonzero count, in ascending
 * value order (the reference assigns entry value - 2, i.e. the same order when the
 * labels are 2, 3, ...), with id = the entry's index, voxels = its count,
 * occupied_voxels = the sum of all counts and percentage = (float)voxels /
 * (float)occupied_voxels.  Host-only (no device work): counts are the 65536 counts
 * vf_histogram returns.  *n receives the number of entries; at most cap are
 * written. */
typedef struct {
    uint32_t id;
    uint32_t value;             /* the fragment's label */
    uint32_t voxels;
    uint32_t occupied_voxels;
    float percentage;
    uint32_t voxelization_size[3];
} vf_fragment_meta;
int vf_fragment_metadata(const uint32_t* counts, const uint32_t dims[3], vf_fragment_meta* out, uint32_t cap,
                         uint32_t* n);

/* ---- marching cubes -------------------------------------------------------- */
/* RegularGrid::toTriangleMesh -> MarchingCubes::triangulateFieldGPU
 * (RegularGrid.cpp:443-486, MF/Source/Graphics/Core/MarchingCubes.cpp:364-433,
 * marchingCubes-comp.glsl:96-156), all labels in one pass with an exact
 * lattice-edge weld.  Two-call protocol: vf_marching_cubes computes the mesh
 * on the device and reports sizes; vf_mesh_read copies it out. */
typedef struct {
    uint32_t num_vertices;
    uint32_t num_triangles;
    uint32_t num_labels;
} vf_mesh_info;
int vf_marching_cubes(vf_grid* g, uint32_t label /* VF_ALL_LABELS or one label */, vf_mesh_info* info);
/* xyz: 3*nv FP32; tri: 3*nt vertex indices local to their label's vertex range;
 * tri_flag: nt boundary flags (may be NULL); labels: nl; vert_offsets and
 * tri_offsets: nl+1 each (may be NULL). */
int vf_mesh_read(vf_grid* g, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels,
                 uint32_t* vert_offsets, uint32_t* tri_offsets);

/* Laplacian smoothing of the mesh computed by the last vf_marching_cubes call,
 * in place (MarchingCubes::smoothSurface, MarchingCubes.cpp:395-401, 491-514;
 * laplacianSmoothing-comp.glsl, finishLaplacianSmoothing-comp.glsl): first
 * nonboundary_iterations steps moving the non-boundary vertices (faces whose
 * three vertices are non-boundary) with nonboundary_weight, then
 * boundary_iterations steps moving the boundary vertices (all faces) with
 * boundary_weight.  A step moves a vertex to mix(p, average of its face
 * neighbours, weight), the average taken over positions scaled by 10^4 and
 * truncated to integers.  A vertex is a boundary vertex when any of its
 * triangles is flagged.  The reference's iteration counts are
 * (uint32_t)(max grid edge * 0.048f) with weights 0.9 and 0.2
 * (FractureParameters.h:93-110). */
int vf_mesh_smooth(vf_grid* g, uint32_t nonboundary_iterations, float nonboundary_weight, uint32_t boundary_iterations,
                   float boundary_weight);

/* One fragment mesh of the last vf_marching_cubes call (index into its label
 * list) in the reference's binary mesh format, CADModel::saveBinary
 * (MF/Source/Graphics/Core/CADModel.cpp:839-855): uint32 vertex count, then per
 * vertex the 64-byte Model3D::VertexGPUData {position, pad, normal, pad,
 * texcoord, pad2, tangent, pad} (Model3D.h:37-50), uint32 face count, then per
 * face the 16-byte FaceGPUData {3 vertex indices, component id}.  As in the
 * reference's dataset export (endInsertionBatch without computeMeshData,
 * CADModel.cpp:66-82) only positions are set; the other fields are zero.
 * Writes at most cap bytes; *size gets the full size. */
int vf_mesh_export_binm(vf_grid* g, uint32_t label_index, uint8_t* out, size_t cap, size_t* size);

/* ---- no device allocation --------------------------------------------------- */
/* Largest grid, and the device scratch held per GPU, of vf_fracture_naive_noalloc. */
#define VF_NOALLOC_MAX_CELLS (1ull << 24)
#define VF_NOALLOC_SCRATCH_BYTES (66560ull + VF_NOALLOC_MAX_CELLS)
/* vf_fracture_naive on a caller-owned device grid, with no device allocation
 * and no allocating library call.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (serving loops, harnesses that
 * forbid hipMalloc, code that must not allocate while it runs).
 * d_cells: device pointer (16-byte aligned) to dims[0]*dims[1]*dims[2] uint16
 * cells in the reference's layout, relabelled in place as vf_fracture_naive
 * relabels a grid; seeds, S and distance as for vf_fracture_naive (host
 * array); stream: a hipStream_t, or NULL for the default stream; the call
 * returns when the work on that stream has finished.  Everything beyond the
 * cells (the grid's occupancy bytes, the launch descriptor, the seeds) lives in
 * one static device array of VF_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends: grids of more than VF_NOALLOC_MAX_CELLS cells fail with
 * VF_ERR_INVALID_ARG.  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU race. */
int vf_fracture_naive_noalloc(uint16_t* d_cells, const uint32_t dims[3], const vf_seed* seeds, uint32_t S,
                              int distance, void* stream);

/* ---- grid file formats ----------------------------------------------------- */
/* RegularGrid::exportRLE (RegularGrid.cpp:672-714): uvec3 dims + {uint16 value,
 * uint32 count} runs.  Writes at most cap bytes; *size gets the full size. */
int vf_grid_export_rle(vf_grid* g, uint8_t* out, size_t cap, size_t* size);
/* inverse of the above (docs/decompress/decompress_grid.py:16-32); dims must match */
int vf_grid_import_rle(vf_grid* g, const uint8_t* in, size_t n);

/* ---- synthetic input ------------------------------------------------------- */
/* Closed lathe-revolved vessel (outer wall, inner wall, bottom disks, rim),
 * spline profile from seed, normalised like CADModel::load
 * (MF/Source/Graphics/Core/CADModel.cpp:148-156): largest half extent 1.
 * Pass NULL buffers to query nv/nt. */
int vf_vessel_mesh(uint64_t seed, uint32_t radial, uint32_t height_segments, float* xyz, uint32_t* nv,
                   uint32_t* tri, uint32_t* nt);

/* ---- device memory ----------------------------------------------------------- */
/* Plain device buffers for callers that keep meshes resident in HBM
 * (vf_job.mesh_on_device, vf_voxelize_device).  No reference counterpart: the
 * reference keeps meshes in GL buffers owned by Model3D. */
int vf_device_alloc(size_t bytes, void** ptr);
int vf_device_free(void* ptr);
int vf_memcpy_htod(void* dst, const void* src, size_t bytes);

/* ---- timing ---------------------------------------------------------------- */
/* Per-kernel HIP-event timing on the stream each kernel is launched on.
 * When enabled, every launch of a named kernel is bracketed by events;
 * vf_timing_read returns total milliseconds and launch count per name. */
int vf_timing_enable(int enable);
int vf_timing_reset(void);
int vf_timing_read(const char* kernel, double* total_ms, uint64_t* launches);

/* ---- batched form ------------------------------------------------------------ */
/* Host buffers for one job's marching-cubes mesh, in vf_mesh_read's layout:
 * positions, triangles with indices local to their label's vertex range,
 * boundary flags, labels ascending and per-label vertex / triangle offsets.
 * The run writes num_*; the mesh is copied (copied = 1) only when it fits the
 * capacities. */
typedef struct {
    float* xyz;              /* 3 * cap_vertices */
    uint32_t* tri;           /* 3 * cap_triangles */
    uint8_t* tri_flag;       /* cap_triangles, may be NULL */
    uint16_t* labels;        /* cap_labels, may be NULL */
    uint32_t* vert_offsets;  /* cap_labels + 1, may be NULL */
    uint32_t* tri_offsets;   /* cap_labels + 1, may be NULL */
    uint32_t cap_vertices, cap_triangles, cap_labels;
    uint32_t num_vertices, num_triangles, num_labels;
    int copied;
} vf_mesh_out;

/* One fracture job: voxelize a mesh, seed, fracture, then boundaries, histogram
 * and (optionally) marching cubes, with the grid resident on the device.
 * The same sequence CADScene::fractureModel + prepareScene drives
 * (MF/Source/Graphics/Application/CADScene.cpp:624-691, :791-833). */
typedef struct {
    const float* xyz;      /* host mesh (or device when mesh_on_device) */
    const uint32_t* tri;
    uint32_t nv, nt;
    uint32_t dims[3];
    float aabb_min[3], aabb_max[3];
    uint32_t num_seeds;
    uint64_t seed;         /* seeding RNG key */
    int seed_location;     /* VF_SEED_* */
    int algorithm;         /* VF_NAIVE / VF_FLOOD */
    int distance;          /* VF_EUCLIDEAN / VF_MANHATTAN / VF_CHEBYSHEV */
    int neighbourhood;     /* flood only, VF_NEIGH_* */
    uint64_t noise_seed;   /* flood only */
    uint32_t noise_amp;    /* flood only */
    int boundary_size;     /* detectBoundaries size; 0 skips the stage */
    int marching_cubes;    /* 1: extract every label's surface (RegularGrid::toTriangleMesh) */
    /* Laplacian smoothing of the extracted meshes (vf_mesh_smooth): non-boundary
     * steps and weight, then boundary steps and weight; zero steps skip it */
    uint32_t smooth_iterations[2];
    float smooth_weights[2];
    vf_mesh_out* mesh_out; /* optional host copy of the mesh (marching_cubes = 1) */
    /* Seeder::mergeSeeds as CADScene::fractureModel runs it (CADScene.cpp:647-655):
     * extra_seeds more seeds drawn anywhere (VF_SEED_BOTH, key extra_seed) follow
     * copies of the num_seeds fragment seeds; each of those takes its nearest
     * fragment's label (merge_distance, first on ties) with a running sub-seed id
     * in bits 8..15, and the job runs on fragments + merged list.  The flood then
     * runs its sub-seed pass.  0 skips it. */
    uint32_t extra_seeds;
    uint64_t extra_seed;
    int merge_distance;
    /* RegularGrid::erode in place of detectBoundaries (CADScene.cpp:675-687):
     * erosion_iterations steps of boundary detection and erosion with the
     * VF_EROSION_* window of side erosion_size, then the isolated-cell pass
     * (vf_erode); boundary_size is not used then. */
    int erode;
    int erosion_convolution;
    uint32_t erosion_size, erosion_iterations;
    float erosion_probability, erosion_threshold;
    uint64_t erosion_seed;
    /* The caller's fragment seeds (num_seeds entries, host memory, labels in .w)
     * in place of the uniform draw: CADScene::fractureModel's impact seeds
     * (_impactSeeds, CADScene.cpp:643-646).  NULL draws them (seed,
     * seed_location).  Flood labels keep the fragment id in the low 8 bits. */
    const vf_seed* seeds;
    /* Seeder::nearSeeds on the fragment seeds (CADScene.cpp:636-646, vf_seed_near):
     * impacts times, up to impact_seeds new surface seeds in all around a
     * fragment seed, spreading impact_spreading (1..64), key impact_seed.  The
     * new seeds join the fragments, labelled after the last one, before
     * mergeSeeds.  0 impacts skips it. */
    uint32_t impacts, impact_seeds, impact_spreading;
    uint64_t impact_seed;
    int mesh_on_device;
    int poison;            /* failure-path testing: the job fails with VF_ERR_POISON */
    uint16_t* labels_out;  /* optional host copy of the final grid (V entries) */
    uint32_t* histogram_out; /* optional host copy of counts[0..65535] */
} vf_job;

typedef struct {
    int status;
    uint64_t voxels;        /* V = nx*ny*nz */
    uint64_t occupied;      /* non-EMPTY cells after voxelization */
    uint64_t checksum;      /* order-independent grid checksum: sum (mod 2^64) over non-EMPTY cells c (linear
                               index) holding v of (a << 32) | fmix32(a ^ c ^ 0x7F4A7C15), a = fmix32(c ^ v * 0x9E3779B1),
                               fmix32 the MurmurHash3 finalizer */
    uint32_t num_labels;    /* distinct labels with a nonzero count */
    uint32_t mc_vertices, mc_triangles;
    uint32_t flood_levels;  /* flood relaxation rounds run (brick worklist rounds until drained) */
    double latency_ms;      /* submit -> completion (server) or batch wall time */
    double queue_ms;        /* submit -> admission (server) */
    uint32_t flood_bricks;  /* flood work: brick visits over all rounds */
    uint32_t flood_sweeps;  /* flood work: in-brick relaxation sweeps over all visits */
} vf_result;

typedef struct vf_batch vf_batch;
/* Tuning: tile/block shapes come from the autotuner's table when tune_path is
 * non-NULL and exists; otherwise defaults. */
int vf_batch_create(vf_batch** out);
int vf_batch_destroy(vf_batch* b);
/* Runs n jobs together: every stage is one launch over all jobs' tiles (the
 * flood runs all jobs' BFS levels in shared launches), captured in a HIP graph
 * per launch shape.  results: n entries.  Device state and graphs are kept
 * while a call passes the same job descriptors as the last; host meshes are
 * validated and uploaded again on every call, so buffers refilled in place
 * are seen. */
int vf_batch_run(vf_batch* b, const vf_job* jobs, uint32_t n, vf_result* results);
/* Plain form: the same jobs one after another, one job per launch. */
int vf_plain_run(const vf_job* jobs, uint32_t n, vf_result* results);
int vf_batch_set_graphs(vf_batch* b, int use_graphs);

/* ---- served form ------------------------------------------------------------- */
/* Continuous batching: submitted jobs are admitted against an HBM budget as
 * slots and memory free up, join the running batch at the next BFS level, and
 * leave when done.  Thread-safe submit/poll. */
typedef struct vf_server vf_server;
int vf_server_create(size_t hbm_budget_bytes, uint32_t max_slots, vf_server** out);
int vf_server_destroy(vf_server* s);
/* The job (and its mesh) is copied; ticket identifies it. */
int vf_server_submit(vf_server* s, const vf_job* job, uint64_t* ticket);
int vf_server_cancel(vf_server* s, uint64_t ticket);
/* VF_ERR_NOT_READY while running; otherwise fills result (result->status holds
 * the job's own status) and forgets the ticket. */
int vf_server_poll(vf_server* s, uint64_t ticket, vf_result* result);
/* blocks until the ticket completes */
int vf_server_wait(vf_server* s, uint64_t ticket, vf_result* result);
/* bytes currently reserved by admitted jobs and number of jobs in flight */
int vf_server_stats(vf_server* s, size_t* reserved_bytes, uint32_t* in_flight, uint64_t* completed);
/* Device memory audit of the scheduler: bytes reserved by admitted jobs (each
 * job reserves its own grid block, rounded to its size class, plus its share of
 * the scratch its launches use), the device memory the scheduler held at its
 * last allocation step (blocks in use plus blocks cached for reuse), the peak of
 * that figure (cached blocks are freed before it would pass the budget), and the
 * largest excess of in-use bytes over reservations seen so far (0 when the
 * budget accounting holds).  Any pointer may be NULL.  No reference
 * counterpart: the reference has no server. */
int vf_server_memory(vf_server* s, size_t* reserved_bytes, size_t* device_bytes, size_t* peak_device_bytes,
                     size_t* max_unreserved_bytes);

/* ---- autotuner ---------------------------------------------------------------- */
/* Launch shapes per grid-edge bucket (largest edge <= 64, <= 128, <= 256,
 * larger): workgroup counts of the cell, voxelizer-pair, boundary-brick and
 * flood stages, walker run lengths (naive tiles, boundary bricks per step)
 * and the batched flood's round chunking.  A batch uses the bucket of its
 * largest edge.  No reference counterpart: the reference sizes its GL
 * dispatches from fixed work-group sizes (ComputeShader.cpp:152-155).
 *
 * vf_tune_run times the batched form on synthetic vessel jobs of each bucket
 * (coordinate descent; quick = 1: fewer jobs and repeats, buckets above 64),
 * installs the winners and writes the table to path (text).  vf_tune_jobs
 * tunes the bucket of the caller's own batch on that batch, reporting its
 * median time before and after (either may be NULL) and, when path is not
 * NULL, writing the table.  vf_tune_load installs a table for later launches.
 * Shapes change timing only, never results. */
int vf_tune_run(const char* path, int quick);
int vf_tune_jobs(const vf_job* jobs, uint32_t n, const char* path, int quick, double* before_ms, double* after_ms);
int vf_tune_load(const char* path);

#ifdef __cplusplus
}
#endif
#endif
