// Autotuner: picks launch shapes per grid-edge bucket (workgroup counts, walker
// run lengths, flood round chunking) by timing the batched form on synthetic
// vessel jobs of that bucket, or on the caller's own batch (coordinate descent
// over the parameters, median of repeated runs), and stores / loads the table
// as text.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "tune.hpp"

namespace vf {

namespace {

std::mutex g_mu;
// Defaults measured on MI355X with the mixed 64-job batch.  Cell kernels: 8192 workgroups of 4 waves
// (8 waves per SIMD for the streaming kernels; 2048 left them at 2 and latency-bound).  Boundary
// bricks: 16384.  Flood relaxation: 4096 one-wave workgroups (16 per CU), all resident at once, taking
// short runs of their XCD's brick list round-robin (batch step on one box: 10.91-10.93 ms at 4096,
// 10.95-10.98 at 4608, 11.02 at 3584, 11.30 at 5120).
// Walker runs: 8 naive tiles, 4 boundary bricks.  Flood: chunks of 8 rounds, the first drain check after 3.
LaunchShape g_shapes[kTuneBuckets] = {
    { 8192, 4096, 8192, 16384, 8, 4, 8, 3, 2048 },
    { 8192, 4096, 8192, 16384, 8, 4, 8, 3, 2048 },
    { 8192, 4096, 8192, 16384, 8, 4, 8, 3, 2048 },
    { 8192, 4096, 8192, 16384, 8, 4, 8, 3, 2048 },
};
uint64_t g_generation = 1;
thread_local int t_bucket = kTuneBuckets - 1;

constexpr uint32_t kEdgeOfBucket[kTuneBuckets] = { 64, 128, 256, 384 };

}  // namespace

bool shape_ok(const LaunchShape& s)
{
    auto ok = [](uint32_t v) { return v >= 64 && v <= (1u << 20); };
    return ok(s.cell_blocks) && ok(s.flood_blocks) && ok(s.pair_blocks) && ok(s.brick_blocks) &&
           (s.naive_run == 4 || s.naive_run == 8 || s.naive_run == 16) &&
           (s.brick_run == 2 || s.brick_run == 4 || s.brick_run == 8) && s.round_chunk >= 2 && s.round_chunk <= 32 &&
           s.lead_chunks >= 1 && s.lead_chunks <= 8 && ok(s.mc_blocks) && s.mc_blocks % 8 == 0;
}

int tune_bucket(uint32_t max_edge)
{
    if (max_edge <= 64) return 0;
    if (max_edge <= 128) return 1;
    if (max_edge <= 256) return 2;
    return 3;
}

LaunchShape active_shape()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_shapes[t_bucket];
}

void set_active_bucket(int bucket) { t_bucket = std::min(std::max(bucket, 0), kTuneBuckets - 1); }

LaunchShape get_shape(int bucket)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_shapes[std::min(std::max(bucket, 0), kTuneBuckets - 1)];
}

void set_shape(int bucket, const LaunchShape& s)
{
    if (bucket < 0 || bucket >= kTuneBuckets || !shape_ok(s)) fail(VF_ERR_INVALID_ARG, "bad launch shape");
    std::lock_guard<std::mutex> lk(g_mu);
    g_shapes[bucket] = s;
    ++g_generation;
}

uint64_t tune_generation()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_generation;
}

}  // namespace vf

using namespace vf;

namespace {

// A synthetic batch of the bucket: 16 vessel jobs (8 in quick mode) with edges
// in (edge/2, edge], seeds 2-64, naive and flood alternating; meshes resident
// in device memory, so a timed run is device work only.
struct TuneJobs {
    std::vector<std::vector<float>> xyz;
    std::vector<std::vector<uint32_t>> tri;
    std::vector<void*> dev;  // device copies of the meshes
    std::vector<vf_job> jobs;
    ~TuneJobs()
    {
        for (void* p : dev) vf_device_free(p);
    }
};

TuneJobs make_jobs(int bucket, bool quick)
{
    TuneJobs T;
    const uint32_t emax = kEdgeOfBucket[bucket], n = quick ? 8 : 16;
    uint64_t st = 0x9E3779B97F4A7C15ull * (uint64_t)(bucket + 1);
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t tris = 5000 + (uint32_t)(next() % 100000);
        const uint32_t r = std::max(3u, (uint32_t)std::lround(std::sqrt(tris / 10.0)));
        const uint32_t h = std::max(1u, (uint32_t)std::lround(2.5 * r));
        uint32_t nv = 0, nt = 0;
        if (vf_vessel_mesh(next(), r, h, nullptr, &nv, nullptr, &nt) != VF_OK) fail(VF_ERR_INVALID_ARG, "vessel");
        T.xyz.emplace_back(3ull * nv);
        T.tri.emplace_back(3ull * nt);
        if (vf_vessel_mesh(st, r, h, T.xyz.back().data(), &nv, T.tri.back().data(), &nt) != VF_OK)
            fail(VF_ERR_INVALID_ARG, "vessel");
    }
    for (uint32_t i = 0; i < n; ++i) {
        vf_job j{};
        const size_t bx = T.xyz[i].size() * sizeof(float), bt = T.tri[i].size() * sizeof(uint32_t);
        void *dx = nullptr, *dt = nullptr;
        if (vf_device_alloc(bx, &dx) != VF_OK || vf_device_alloc(bt, &dt) != VF_OK)
            fail(VF_ERR_OUT_OF_MEMORY, "tuning meshes");
        T.dev.push_back(dx);
        T.dev.push_back(dt);
        if (vf_memcpy_htod(dx, T.xyz[i].data(), bx) != VF_OK || vf_memcpy_htod(dt, T.tri[i].data(), bt) != VF_OK)
            fail(VF_ERR_DEVICE, "tuning meshes");
        j.xyz = static_cast<const float*>(dx);
        j.tri = static_cast<const uint32_t*>(dt);
        j.mesh_on_device = 1;
        j.nv = (uint32_t)(T.xyz[i].size() / 3);
        j.nt = (uint32_t)(T.tri[i].size() / 3);
        const uint32_t e = std::max<uint32_t>(16, (emax / 2 + (uint32_t)(next() % (emax / 2 + 1))) / 4 * 4);
        j.dims[0] = j.dims[1] = j.dims[2] = e;
        for (int q = 0; q < 3; ++q) {
            j.aabb_min[q] = -1.03f;
            j.aabb_max[q] = 1.03f;
        }
        j.num_seeds = 2 + (uint32_t)(next() % 63);
        j.seed = next() | 1;
        j.seed_location = VF_SEED_BOTH;
        j.algorithm = (i & 1) ? VF_FLOOD : VF_NAIVE;
        j.distance = (int)(next() % 3);
        j.neighbourhood = VF_NEIGH_FROM_DISTANCE;
        j.noise_seed = next();
        j.noise_amp = j.algorithm == VF_FLOOD ? 2 : 0;
        j.boundary_size = 1;
        T.jobs.push_back(j);
    }
    return T;
}

void run_or_fail(vf_batch* b, const std::vector<vf_job>& jobs, std::vector<vf_result>& res)
{
    const int rc = vf_batch_run(b, jobs.data(), (uint32_t)jobs.size(), res.data());
    if (rc != VF_OK) {
        const std::string msg = vf_last_error();
        fail(rc, "tuning batch failed: %s", msg.c_str());
    }
}

double time_batch(vf_batch* b, const std::vector<vf_job>& jobs, int reps)
{
    std::vector<vf_result> res(jobs.size());
    run_or_fail(b, jobs, res);
    std::vector<double> t;
    for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        run_or_fail(b, jobs, res);
        t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

void write_table(const char* path)
{
    FILE* f = std::fopen(path, "w");
    if (!f) fail(VF_ERR_INVALID_ARG, "cannot write %s", path);
    std::fprintf(f, "# vfrag launch shapes: bucket max_edge cell_blocks flood_blocks pair_blocks brick_blocks "
                    "naive_run brick_run round_chunk lead_chunks mc_blocks\n");
    for (int b = 0; b < kTuneBuckets; ++b) {
        const LaunchShape s = get_shape(b);
        std::fprintf(f, "%d %u %u %u %u %u %u %u %u %u %u\n", b, kEdgeOfBucket[b], s.cell_blocks, s.flood_blocks,
                     s.pair_blocks, s.brick_blocks, s.naive_run, s.brick_run, s.round_chunk, s.lead_chunks,
                     s.mc_blocks);
    }
    std::fclose(f);
}

// Coordinate descent over the bucket's shape on one batch: each parameter in turn takes each of its
// candidates, a change is kept only when the median batch time drops by more than 2 %; the walk over
// all parameters runs twice (once in quick mode).  Returns the starting and the final median.
// flood_blocks is swept only on a caller's batch: its best value follows the length of the brick worklist,
// which grows with the number of jobs in a batch, not with the bucket's edge, so a bucket's synthetic
// jobs timed alone do not predict it.
void tune_bucket_on(vf_batch* b, const std::vector<vf_job>& jobs, int bucket, bool quick, bool flood_blocks,
                    double* before, double* after)
{
    const int reps = quick ? 5 : 7;
    LaunchShape best = get_shape(bucket);
    double best_ms = time_batch(b, jobs, reps);
    if (before) *before = best_ms;
    auto sweep = [&](uint32_t LaunchShape::*field, const std::vector<uint32_t>& cands) {
        for (uint32_t v : cands) {
            if (best.*field == v) continue;
            LaunchShape s = best;
            s.*field = v;
            set_shape(bucket, s);
            const double ms = time_batch(b, jobs, reps);
            if (ms < best_ms * 0.98) {
                best_ms = ms;
                best = s;
            }
        }
        set_shape(bucket, best);
    };
    for (int pass = 0; pass < (quick ? 1 : 2); ++pass) {
        sweep(&LaunchShape::cell_blocks, { 4096, 8192, 16384 });
        sweep(&LaunchShape::naive_run, { 4, 8, 16 });
        sweep(&LaunchShape::pair_blocks, { 8192, 16384, 32768 });
        sweep(&LaunchShape::brick_blocks, { 4096, 8192, 16384 });
        sweep(&LaunchShape::brick_run, { 2, 4, 8 });
        if (flood_blocks) sweep(&LaunchShape::flood_blocks, { 3072, 4096, 4608, 6144 });
        sweep(&LaunchShape::round_chunk, { 4, 8, 16 });
        sweep(&LaunchShape::lead_chunks, { 1, 2, 3, 4, 6 });
        sweep(&LaunchShape::mc_blocks, { 1024, 1280, 2048, 2560, 4096 });
    }
    // the final shape timed again, so the figure reported is not the luckiest of the sweep's runs
    if (after) *after = time_batch(b, jobs, reps);
}

struct BatchHandle {
    vf_batch* b = nullptr;
    BatchHandle()
    {
        if (vf_batch_create(&b) != VF_OK) fail(VF_ERR_DEVICE, "cannot create a tuning batch");
    }
    ~BatchHandle() { vf_batch_destroy(b); }
};

}  // namespace

extern "C" {

int vf_tune_run(const char* path, int quick)
{
    return guard([&] {
        if (!path) fail(VF_ERR_INVALID_ARG, "path is NULL");
        if (vf_device_count() < 1) fail(VF_ERR_DEVICE, "no HIP device available");
        for (int bucket = quick ? 1 : 0; bucket < kTuneBuckets; ++bucket) {
            TuneJobs T = make_jobs(bucket, quick != 0);
            BatchHandle h;
            tune_bucket_on(h.b, T.jobs, bucket, quick != 0, false, nullptr, nullptr);
        }
        write_table(path);
    });
}

int vf_tune_jobs(const vf_job* jobs, uint32_t n, const char* path, int quick, double* before_ms, double* after_ms)
{
    return guard([&] {
        if (!jobs || n == 0) fail(VF_ERR_INVALID_ARG, "no jobs to tune on");
        if (vf_device_count() < 1) fail(VF_ERR_DEVICE, "no HIP device available");
        uint32_t edge = 1;
        for (uint32_t i = 0; i < n; ++i)
            edge = std::max(edge, std::max(jobs[i].dims[0], std::max(jobs[i].dims[1], jobs[i].dims[2])));
        const std::vector<vf_job> js(jobs, jobs + n);
        BatchHandle h;
        tune_bucket_on(h.b, js, tune_bucket(edge), quick != 0, true, before_ms, after_ms);
        if (path) write_table(path);
    });
}

int vf_tune_load(const char* path)
{
    return guard([&] {
        if (!path) fail(VF_ERR_INVALID_ARG, "path is NULL");
        FILE* f = std::fopen(path, "r");
        if (!f) fail(VF_ERR_INVALID_ARG, "cannot read %s", path);
        char line[256];
        int loaded = 0;
        while (std::fgets(line, sizeof line, f)) {
            if (line[0] == '#') continue;
            int b = -1;
            unsigned edge = 0;
            LaunchShape s = get_shape(kTuneBuckets - 1);
            const int got = std::sscanf(line, "%d %u %u %u %u %u %u %u %u %u %u", &b, &edge, &s.cell_blocks,
                                        &s.flood_blocks, &s.pair_blocks, &s.brick_blocks, &s.naive_run, &s.brick_run,
                                        &s.round_chunk, &s.lead_chunks, &s.mc_blocks);
            if (got != 11 || b < 0 || b >= kTuneBuckets || !shape_ok(s)) {
                std::fclose(f);
                fail(VF_ERR_INVALID_ARG, "malformed tuning line: %s", line);
            }
            set_shape(b, s);
            ++loaded;
        }
        std::fclose(f);
        if (!loaded) fail(VF_ERR_INVALID_ARG, "no tuning entries in %s", path);
    });
}

}  // extern "C"
