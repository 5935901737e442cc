// Multi-job execution engine shared by the batched, plain and served forms.
//
// A fracture job runs voxelize -> seed -> (naive | flood rounds) -> flood
// finish -> boundaries + histogram -> checksum [-> marching cubes], the
// sequence CADScene::fractureModel + prepareScene drives
// (MF/Source/Graphics/Application/CADScene.cpp:624-691, :791-833), with the
// grid resident on the device throughout.  Every stage is one launch over a
// Table: a compact array of job descriptors plus prefix tables, so any subset
// of jobs (one for the plain form, a whole batch, or whichever jobs a server
// tick has in that phase) shares each launch.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "jobs.hpp"
#include "tune.hpp"

namespace vf {

struct EJob {
    vf_job spec{};
    std::vector<float> xyz_copy;     // server: the job owns a copy of its mesh
    std::vector<uint32_t> tri_copy;
    std::vector<vf_seed> seeds_copy;  // server: and of the caller's fragment seeds
    JobDev d{};
    void* mem = nullptr;             // one device block per job
    size_t bytes = 0;
    uint64_t words = 0;
    uint32_t tiles = 0, groups = 0;
    vf_result res{};
    uint64_t ticket = 0;
    double t_submit = 0, t_admit = 0;
    int phase = 0;
    int vstatus = 0;                 // job_validate's status when the job was prepared
    bool sub_done = false;           // the flood's sub-seed pass has run (served form)
    bool fronted = false;            // the front stage is queued: nearSeeds' counts are in fctl
    uint16_t* erode_out = nullptr;   // erosion output grid and window (jobs with erode set)
    uint8_t* erode_mask = nullptr;
    uint32_t erode_side = 0;
    float erode_limit = 0;
    bool cancelled = false;
    bool admitted = false;           // server: holds a slot and budget until completed
};

// Device bytes a job needs while resident: its own block (grid, occupancy,
// bitmasks, flood keys and brick flags, seeds, histogram, mesh copy) plus its
// share of the engine scratch its launches use (voxelizer setup records and
// scans, seed counts, launch table entries, flood worklists).
size_t job_footprint(const vf_job& j);
// VF_OK or the status the job fails with before touching the device.
int job_validate(const vf_job& j, std::string& why);

struct Table {
    std::vector<EJob*> jobs;
    void* dev = nullptr;
    Launch L{};
    const uint32_t* tri_prefix = nullptr;
    const uint32_t* col_prefix = nullptr;
    uint32_t total_tris = 0, total_groups = 0;
    const uint32_t* bt_prefix = nullptr;  // boundary bricks per job
    uint32_t total_btiles = 0;
    bool any_wide = false;                // some job uses a boundary window wider than 1
    uint32_t* gather = nullptr;       // 16 words per job
    FloodRoundScratch flood{};      // chunk prefix + worklists of the flood rounds
    int bucket = 0;                 // launch-shape bucket (largest grid edge of the table)
    void* flood_lists = nullptr;    // separate allocation of flood.lists
    size_t dev_bytes = 0, list_bytes = 0;  // class sizes of the two blocks
    bool any_flood = false, any_naive32 = false, any_naive64 = false;
    bool any_extra = false, any_sub = false, any_erode = false, any_near = false;
    Launch xL{};                          // the jobs as their extra-seed draw sees them
    const JobDev* ejobs = nullptr;        // the jobs as their erosion passes see them (boundary size 1)
    const uint32_t* local_prefix = nullptr;  // per job {0, tiles, 0, bricks}: one-job launches
};

class Engine {
public:
    Engine();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    hipStream_t s = nullptr;
    hipStream_t side = nullptr;        // the batched form's naive fracture, beside the flood rounds
    hipEvent_t side_go = nullptr, side_done = nullptr;
    int device = 0;

    void alloc(EJob& j);    // stream-ordered allocation + descriptor + mesh upload
    void upload_inputs(EJob& j);  // host mesh and given seeds -> the job's device copies
    void release(EJob& j);  // stream-ordered free
    Table* make_table(const std::vector<EJob*>& js);
    void drop_table(Table* t);
    void reserve_front(const Table& t);  // size the front scratch (not capturable)
    void release_front();                // return the front scratch (the next reserve_front takes it again)
    // device bytes held: job blocks, launch tables, worklists and scratch buffers, in use or cached
    size_t device_bytes() const;
    size_t cached_bytes() const { return cache_.cached(); }
    // free cached blocks until `need` more bytes fit in `budget` (or nothing is left to free)
    void make_room(size_t need, size_t budget);
    void trim_cache(size_t keep = 0) { cache_.trim(keep); }
    // launches only; split: without the fracture (the batched form's naive_side and flood_start)
    void front(Table& t, bool split = false);
    void flood_start(Table& t);  // the flood's key init and seeding
    // the naive fracture of t's naive jobs on the side stream, after the work queued so far on s; s waits
    // for it at join_side()
    void naive_side(Table& t);
    void join_side();
    void rounds(Table& t, uint32_t n);  // flood relaxation rounds
    void post(Table& t);
    void erode(Table& t);      // per-job erosion (post calls it)
    void flood_sub(Table& t);  // the flood's sub-seed pass for the jobs with merged seeds
    void gather(Table& t, uint32_t* host16);  // async copy of the gathered stats
    // fill j.res from 16 gathered words and make the optional grid and histogram copies
    void finish(EJob& j, const uint32_t* w16);
    // marching cubes (and smoothing) of every listed job with marching_cubes set and no error, in one
    // launch set: counts into the results, meshes into the jobs' mesh_out.  reserve: see mc_run_jobs.
    void extract(const std::vector<EJob*>& js, const std::function<bool(size_t)>& reserve = nullptr);
    void release_extract();  // the meshes and their scratch back to the block cache

    uint32_t* pinned = nullptr;     // gathered stats staging (16 words x kMaxTableJobs)
    static constexpr uint32_t kMaxTableJobs = 1024;

private:
    bool side_pending_ = false;  // naive work queued on the side stream that s has not joined yet
    BlockCache cache_;       // job blocks, tables, worklists and front scratch
    CacheBuf vox_setup_, vox_counts_, vox_shell_off_, scan_tmp_, seed_cnt_, seed_off_;
    McScratch mc_;
};

double now_ms();

}  // namespace vf
