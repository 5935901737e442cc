// Device-side job descriptor and the stage launchers.
//
// Every stage kernel is written once in multi-job form: it receives an array
// of JobDev descriptors plus a prefix table mapping global tiles (or
// triangles) to jobs, and walks the tiles with a persistent grid-stride loop.
// The single-grid API launches it with one job; the batched form launches it
// once for all jobs of a batch, so uneven grids share one launch.
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"

namespace vf {

constexpr uint32_t kMaxNoiseAmp = 2;  // flood cell weights are 1 + [0, amp]
constexpr uint32_t kRelaxBrick = 8;   // flood relaxation brick edge (cells)
constexpr uint32_t kTileCells = 2048; // cells per tile for cell-parallel stages (256 threads x 8 cells)

enum StatSlot : int {
    kStatOccupied = 0,   // non-EMPTY cells after voxelization
    kStatChecksum = 1,   // order-independent grid checksum
    kStatShellPairs = 2, // (triangle, cell) SAT candidates
    kStatColPairs = 3,   // (triangle, column) parity candidates
    kStatError = 4,      // first error flag raised on the device
    kStatFloodWork = 5,  // flood work: (sweeps << 32) + brick visits
    kStatFloodDone = 6,  // flood relaxation rounds run (0 while running)
    kStatLabels = 7,     // labels >= 2 with a nonzero count
    kStatCount = 8
};

// Device error flags (kStatError)
enum DevError : uint32_t {
    kDevErrNone = 0,
    kDevErrSeeding = 1,     // fewer candidates than seeds
    kDevErrKeyOverflow = 2, // flood distance no longer fits the packed key
    kDevErrRelax = 4,       // a flood brick did not settle within its sweep limit
    kDevErrMesh = 8,        // a triangle references a vertex index >= nv (the triangle is skipped)
    kDevErrMerge = 16,      // mergeSeeds: a fragment would get more than 255 sub-seeds
};

struct TriSetup {
    float P[3][3];             // FP32 vertices (SAT input)
    int32_t X[3], Y[3], Z[3];  // snapped fixed-point vertices (CCW in xy)
    int64_t area;              // twice the projected area (> 0) or 0 = no parity work
    int32_t lo[3], hi[3];      // SAT candidate cell range (inclusive)
    int32_t i0, i1, j0, j1;    // parity column range (inclusive)
    uint32_t job;
};

struct JobDev {
    uint32_t nx, ny, nz, wp;  // dims; wp = 32-bit words per z column in the bitmasks (power of two)
    uint64_t V;
    float mn[3], mx[3], cs[3];
    uint16_t* labels;
    uint8_t* occ;
    // voxelizer input / scratch
    const float* xyz;
    const uint32_t* tri;
    uint32_t nt, nv;
    uint32_t* shell_bits;   // nx*ny*wp words
    uint32_t* toggle_bits;  // nx*ny*wp words
    // seeds {x,y,z,label}
    uint4* seeds;
    uint32_t S;             // seeds the fracture uses
    uint32_t S_draw;        // seeds a uniform draw picks (k_seed_pick), written from seeds[0]
    uint32_t S0, extra;     // mergeSeeds: fragment seeds and extra seeds (extra = 0: no merge)
    uint32_t F_in;          // fragment seeds before nearSeeds (drawn or given); nearSeeds sets S0 and S
    uint32_t xoff;          // mergeSeeds: the extra draw sits at seeds[xoff .. xoff + extra)
    uint32_t impacts, near_seeds, spreading;  // nearSeeds (impacts = 0: none)
    uint64_t near_key;
    int merge_dist;         // mergeSeeds distance (VF_*)
    int dist;               // VF_* distance, or kDistEuclidF32 for naive Euclidean ranked by float roots
    uint32_t key_bits;      // bits for the seed index in packed keys
    int neighbourhood;      // 6 or 26
    uint64_t noise_seed;
    uint32_t noise_amp;
    int seed_location;
    uint64_t seed_key;
    uint32_t* keys;         // flood keys, V: kFloodEmptyKey = EMPTY, else (K << 2) | weight (1..3), K = (distance << key_bits) | seed; kFloodUnreachedKey | weight unreached
    uint32_t* sub;          // sub-seed pass (null when the seeds carry one id per fragment): S union-find parents, S keep flags
    unsigned long long* bflags;  // brick flag bits for even / odd rounds (relax_words_per_parity words each), then the brick liveness words
    uint32_t nbricks;       // relaxation bricks (kRelaxBrick^3 cells each)
    uint32_t* fctl;         // flood control: [0] round, [1] rounds run once drained (0 = running), [2..3] flags non-empty, [4..5] round started by the compaction of an even / odd round of a chunk
    unsigned long long* stats;  // kStatCount slots
    uint32_t* hist;         // 65536 counts
    int active;             // 0: the job is skipped by multi-job launches
    int algo;               // VF_NAIVE / VF_FLOOD: selects which fracture launches process the job
    int bsize;              // detectBoundaries window half-size used when a launch passes size < 0
    int naive_key64;        // naive packed keys need 64 bits
};

// Flood keys (see flood.hip): all ones = EMPTY, unreached occupied cells kFloodUnreachedKey | weight.
constexpr uint32_t kFloodEmptyKey = 0xFFFFFFFFu;
constexpr uint32_t kFloodUnreachedKey = 0xFFFFFFF0u;
__host__ __device__ inline uint32_t flood_weight(uint64_t noise_seed, uint32_t cell, uint32_t amp)
{
    return 1u + (amp ? noise_delta(noise_seed, cell, amp) : 0u);
}

uint32_t key_bits_for(uint32_t S);  // bits of the seed index in packed keys

// Naive Euclidean distances: the reference ranks seeds by the float sqrt of the exact integer
// squared distance (NaiveFracturer.cpp:12-15).  Below kSqrtInjectiveMax distinct integers have
// distinct float roots, so the squared distance orders identically; grids whose largest squared
// distance exceeds it rank by the float root's bit pattern (distance code kDistEuclidF32, 64-bit
// keys), where distinct squared distances can tie as they do in the reference.
constexpr uint64_t kSqrtInjectiveMax = 4197200;  // sqrtf(4197200) == sqrtf(4197201)
constexpr int kDistEuclidF32 = 3;
struct NaiveKeyForm {
    int dist;   // distance code the kernel evaluates (VF_* or kDistEuclidF32)
    int key64;  // packed keys need 64 bits
};
NaiveKeyForm naive_key_form(const uint32_t dims[3], int distance, uint32_t S);

__device__ inline uint32_t find_job(const uint32_t* prefix, uint32_t njobs, uint32_t t)
{
    // largest j with prefix[j] <= t  (prefix has njobs+1 entries)
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ inline uint64_t find_job64(const unsigned long long* prefix, uint32_t n, unsigned long long t)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// --------------------------------------------------------------- launchers
// All launchers enqueue on stream s; none synchronises.  `jobs` and the
// prefix tables are device pointers.
struct Launch {
    const JobDev* jobs;          // device array
    uint32_t njobs;
    const uint32_t* tile_prefix; // device, njobs+1 (kTileCells tiles)
    uint32_t total_tiles;
    hipStream_t s;
};

// voxelize.hip
struct VoxScratch {
    TriSetup* setup;                   // total triangles, as 8 planes of 16-byte pieces (store_setup / load_setup)
    unsigned long long* shell_off;     // 2 * total triangles + 1: exclusive prefix of counts (shell work counts,
                                       // then parity column counts; the parity offsets start at shell_off + total)
    uint32_t* counts;                  // 2*total triangles: shell work counts (column segments of one z word
                                       // per candidate (x, y)), then parity column counts
    const uint32_t* tri_prefix;        // njobs+1 prefix over triangles (device)
    uint32_t total_tris;
    void* scan_tmp;                    // scan scratch
    size_t scan_tmp_bytes;
};
// tile_cnt (or null): also count the occupied cells per tile (L.tile_prefix) of jobs that seed anywhere
// (VF_SEED_BOTH), as launch_seed_count does; launch_seed_count(..., true) then skips those jobs
void launch_voxelize(const Launch& L, const VoxScratch& X, const uint32_t* col_prefix, uint32_t total_col_groups,
                     uint32_t* tile_cnt);
void launch_clear_bits(const Launch& L, const uint32_t* col_prefix, uint32_t total_col_groups, uint32_t* tile_cnt);

// grid_ops.hip
void launch_occ_to_labels(const Launch& L);
void launch_labels_to_occ(const Launch& L);
void launch_homogenize(const Launch& L);
void launch_undo_mask(const Launch& L, uint16_t keep_mask);
void launch_clear_hist(const Launch& L);
void launch_histogram(const Launch& L);
void launch_boundaries(const Launch& L, int size /* < 0: per-job bsize */, int with_hist);
// detectBoundaries(size <= 1 per job) fused with the histogram, 8x8x32 LDS
// bricks; bt_prefix: njobs+1 prefix of bricks; jobs with size > 1 go through
// the generic kernel when any_wide.
constexpr uint32_t kBrickX = 8, kBrickY = 8, kBrickZ = 32;
inline uint32_t brick_count(uint32_t nx, uint32_t ny, uint32_t nz)
{
    return ((nx + kBrickX - 1) / kBrickX) * ((ny + kBrickY - 1) / kBrickY) * ((nz + kBrickZ - 1) / kBrickZ);
}
// with_checksum: also sum every job's grid checksum into kStatChecksum
void launch_boundaries_tiled(const Launch& L, const uint32_t* bt_prefix, uint32_t total_btiles, bool any_wide,
                             bool with_checksum);
void launch_job_init(const Launch& L);     // zero stats, histogram and flood control of every job
void launch_gather(const Launch& L, uint32_t* out);  // 16 words per job: stats (low 32 bits) + flood state
void launch_checksum(const Launch& L);
void launch_count_occupied(const Launch& L);
void launch_hist_summary(const Launch& L);  // kStatLabels of every job (added: zeroed by job_init)

// naive.hip
// runs every job with algo == VF_NAIVE and matching key width; distance and
// the z-run fast path are chosen per job inside the kernel
void launch_naive(const Launch& L, bool key64);
void launch_remove_isolated(hipStream_t s, uint16_t* labels, uint16_t* out, uint32_t* frontier_a, uint32_t* frontier_b,
                            uint32_t* counters, uint32_t* visited, const uint32_t dims[3], const uint4* seeds, uint32_t S,
                            uint32_t* host_counter);

// flood.hip
constexpr uint32_t kFloodCtlWords = 6;
constexpr uint32_t kCtlSeedCounts = 16;  // fctl words 16..17: S and S0 as nearSeeds left them
constexpr uint32_t kFloodMaxJobs = 1024;  // jobs per multi-job flood launch
inline uint32_t relax_bricks(uint32_t nx, uint32_t ny, uint32_t nz)
{
    return ((nx + kRelaxBrick - 1) / kRelaxBrick) * ((ny + kRelaxBrick - 1) / kRelaxBrick) *
           ((nz + kRelaxBrick - 1) / kRelaxBrick);
}
// 64-bit brick flag words per round parity, and compaction tasks (64 words) of a job
__host__ __device__ inline uint32_t relax_words_per_parity(uint32_t nbricks) { return (nbricks + 63u) / 64u; }
inline uint32_t relax_tasks(uint32_t nx, uint32_t ny, uint32_t nz)
{
    return (relax_words_per_parity(relax_bricks(nx, ny, nz)) + 63u) / 64u;
}
// Brick liveness (bricks holding an occupied cell, whose keys the init writes), after the two flag
// parities: a bitmap with one 64-bit word per 64 bricks of a brick column (bx, by), bit bz % 64 of word
// (bx * nby + by) * wpc + bz / 64, then one 32-bit word per brick with the liveness of the brick and its
// 26 neighbours, bit (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1).
__host__ __device__ inline uint32_t live_words_per_column(uint32_t nbz) { return (nbz + 63u) / 64u; }
__host__ __device__ inline uint32_t live_words(uint32_t nbx, uint32_t nby, uint32_t nbz)
{
    return nbx * nby * live_words_per_column(nbz);
}
inline uint64_t relax_flag_bytes(uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint32_t nbx = (nx + kRelaxBrick - 1) / kRelaxBrick, nby = (ny + kRelaxBrick - 1) / kRelaxBrick,
                   nbz = (nz + kRelaxBrick - 1) / kRelaxBrick, nb = relax_bricks(nx, ny, nz);
    return 16ull * relax_words_per_parity(nb) + 8ull * live_words(nbx, nby, nbz) + 4ull * nb;
}
// Device state of the flood rounds of one launch table.
struct FloodRoundScratch {
    const uint32_t* task_prefix;   // njobs + 1: prefix of relax_tasks over the jobs (0 for non-flood jobs)
    uint32_t total_tasks;
    unsigned long long* lists;     // 2 x kFloodXcds brick worklists (even and odd rounds of a chunk), list_cap entries each
    uint32_t list_cap;             // >= flood_list_cap(total_tasks)
    uint32_t* counts;              // 2 x kFloodXcds list lengths; zero before the first round (a chunk leaves them zero)
};
constexpr uint32_t kFloodXcds = 8;
uint32_t flood_list_cap(uint32_t total_tasks);  // entries per XCD list
// Key init and brick liveness, then seed stamping and the bricks' neighbour liveness (W's task prefix
// deals the work: a task is 4096 bricks of a job).  Both brick flag parities of every flood job must be
// zero before launch_flood_seed (launch_job_init clears them).
void launch_flood_init(const Launch& L, const FloodRoundScratch& W);
void launch_flood_seed(const Launch& L, const FloodRoundScratch& W);
// `rounds` (rounded up to an even count) relaxation rounds for every running flood job of L.  Each job
// continues at its own round, so jobs may join a table between calls; rounds
// after a job drained do nothing for it, so the launches can be replayed from
// a graph.
// blocks: workgroups of one wave for the compaction and relaxation launches
void launch_flood_rounds(const Launch& L, const FloodRoundScratch& W, uint32_t rounds, uint32_t blocks);
void launch_flood_finish(const Launch& L);
// erode.hip: RegularGrid::erode.  erosion_mask returns the odd window side (or -1) and fills
// side^3 mask bytes and the active fraction; the stencils read L.jobs[0]'s labels.
constexpr uint32_t kMaxErosionSide = 15;
int erosion_mask(int type, uint32_t size, std::vector<uint8_t>& mask, float& activations);
void launch_erode_step(const Launch& L, uint16_t* out, const uint8_t* mask, uint32_t side, float limit,
                       float probability, uint64_t noise_seed, uint64_t V);
void launch_isolated_cells(const Launch& L, uint16_t* out, uint64_t V);

// Sub-seed pass after the first drain (jobs with a non-null sub): merge touching
// same-fragment seed regions, keep each fragment's smallest-id component, and,
// where a region was freed, reset the keys and restart the rounds (fctl[1] = 0).
void launch_flood_sub(const Launch& L);

// seed.hip
constexpr uint32_t kMaxDeviceSeeds = 4096;
void launch_seed_count(const Launch& L, uint32_t* tile_cnt, bool skip_both);
// tile_off: exclusive prefix over the global tiles (total_tiles + 1 entries)
void launch_seed_pick(const Launch& L, const unsigned long long* tile_off);
// Seeder::mergeSeeds for every job with extra seeds (one block each): seeds[S0..2 S0) get copies of the
// fragment seeds [0, S0), seeds[2 S0..2 S0 + extra) hold the extra draw; each of these S0 + extra seeds
// takes its nearest fragment's label with a running sub-seed id in bits 8..15 (more than 255 on one
// fragment flags kDevErrMerge).
void launch_seed_merge(const Launch& L);
// Seeder::nearSeeds on job 0 of L; out: F + num_seeds entries; result[2]: count, draw-cap flag
void launch_seed_near(const Launch& L, const uint4* frags, uint32_t F, uint32_t impacts, uint32_t num_seeds,
                      uint32_t spreading, uint64_t key, uint4* out, uint32_t* result);
// nearSeeds for every job of L with impacts, in place on its fragment seeds (one block per job); sets
// the job's S0 and S in the table.  A draw cap raises kDevErrSeeding.
void launch_seed_near_jobs(const Launch& L);
// copies S and S0 that nearSeeds left in each job's control words into the table L (jobs with impacts whose
// front stage ran before: tables made after it)
void launch_seed_counts(const Launch& L);

// mc.hip: marching cubes of every label of one or many device grids (one
// launch set for all of them).  Scratch and results live in W's buffers, taken
// from `cache` (its owner's block cache, or W's own, which release_all trims).
struct McScratch {
    static constexpr int kBuffers = 32;
    CacheBuf buf[kBuffers];
    BlockCache own;
    BlockCache* cache = &own;
    uint32_t nv = 0, nt = 0, nl = 0;  // totals of the last run; nl counts (grid, label) segments
    bool valid = false;
    std::vector<uint32_t> label_base;  // first segment of each grid of the last run, and the total
    void ensure(int slot, size_t n);
    void release_all();               // every buffer back to the cache (the result is dropped)
    McScratch() = default;
    McScratch(const McScratch&) = delete;
    McScratch& operator=(const McScratch&) = delete;
    ~McScratch();
};
struct McJobSpec {
    const uint16_t* grid;
    const uint32_t* hist;  // the grid's countValues histogram (65536 counts, device)
    uint32_t dims[3];
    float mn[3], cs[3];
    uint32_t want;         // VF_ALL_LABELS or one label
};
struct McJobResult {
    uint32_t nv, nt, nl;   // the grid's vertices, triangles and labels
    uint64_t v0, t0;       // its first vertex and triangle in W
    uint32_t seg0;         // its first segment in W
    int status;            // VF_ERR_INVALID_ARG: more labels than marching cubes supports
};
// reserve (optional): asked for the bytes of each of the two phases' buffers before they are taken; a
// refusal ends the run with every grid's status VF_ERR_BUDGET
void mc_run_jobs(McScratch& W, hipStream_t s, const McJobSpec* jobs, uint32_t n, McJobResult* out,
                 const std::function<bool(size_t)>& reserve);
// copies each grid's mesh of the last run to its vf_mesh_out (NULL entries skipped) when it fits
void mc_read_jobs(McScratch& W, hipStream_t s, const McJobResult* res, vf_mesh_out* const* outs, uint32_t n);
void mc_run(McScratch& W, hipStream_t s, const uint16_t* grid, const uint32_t dims[3], const float mn[3], const float cs[3],
            const uint32_t* hist_dev, uint32_t want, vf_mesh_info* info);
void mc_read(McScratch& W, hipStream_t s, float* xyz, uint32_t* tri, uint8_t* tri_flag, uint16_t* labels, uint32_t* voff,
             uint32_t* toff);
// segments [seg0, seg0 + nseg) of the last run: offsets relative to the first of them
void mc_read_range(McScratch& W, hipStream_t s, uint32_t seg0, uint32_t nseg, float* xyz, uint32_t* tri, uint8_t* tri_flag,
                   uint16_t* labels, uint32_t* voff, uint32_t* toff);
// Laplacian smoothing of the last mesh in W, in place (non-boundary pass, then boundary pass); with
// several grids, each with its own steps and weights
struct McSmooth {
    uint32_t nb_iters;
    float nb_weight;
    uint32_t b_iters;
    float b_weight;
};
void mc_smooth(McScratch& W, hipStream_t s, uint32_t nb_iters, float nb_weight, uint32_t b_iters, float b_weight);
void mc_smooth_jobs(McScratch& W, hipStream_t s, const McSmooth* per_job, uint32_t njobs);

// scan (scan.hip): exclusive prefix sums, graph-capturable, no host sync
size_t scan_tmp_bytes(uint64_t n);
void scan_exclusive_u32_to_u64(const uint32_t* in, unsigned long long* out, uint64_t n, void* tmp, hipStream_t s);

}  // namespace vf
