// Engine implementation: per-job device blocks, table upload, and the stage
// sequences.  All work is enqueued on the engine's stream; only gather
// readbacks synchronise.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace vf {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Layout {
    size_t labels, occ, bits, seeds, hist, stats, fctl, keys, bricks, sub, erode, mask, mesh_xyz, mesh_tri, total;
    uint64_t V, words;
};

// fragment seeds at most: the drawn or given ones, then those nearSeeds adds
uint32_t frag_max(const vf_job& j) { return j.num_seeds + (j.impacts ? j.impact_seeds : 0); }
// seed slots: the fragments, then (mergeSeeds) their copies and the extra draw
uint32_t seed_total(const vf_job& j) { return frag_max(j) + (j.extra_seeds ? frag_max(j) + j.extra_seeds : 0); }
// the flood's sub-seed pass: merged sub-seeds, or given seeds that carry sub-seed ids (bits 8..15)
bool needs_sub(const vf_job& j)
{
    if (j.algorithm != VF_FLOOD) return false;
    if (j.extra_seeds) return true;
    if (j.seeds)
        for (uint32_t i = 0; i < j.num_seeds; ++i)
            if (j.seeds[i].w >> 8) return true;
    return false;
}
constexpr uint32_t kMaxMergeSeeds = 2048;  // fragments + extra seeds of one job (k_seed_merge)
constexpr uint32_t kMaxErosionIterations = 64;

int resolved_neigh(const vf_job& j)
{
    if (j.neighbourhood == VF_NEIGH_FROM_DISTANCE) return j.distance == VF_MANHATTAN ? 6 : 26;
    return j.neighbourhood;
}

Layout layout_of(const vf_job& j)
{
    Layout L{};
    L.V = (uint64_t)j.dims[0] * j.dims[1] * j.dims[2];
    const uint32_t wp = next_pow2(ceil_div(j.dims[2], 32));
    L.words = (uint64_t)j.dims[0] * j.dims[1] * wp;
    const bool flood = j.algorithm == VF_FLOOD;
    size_t o = 0;
    L.labels = o; o = align_up(o + L.V * 2);
    L.occ = o; o = align_up(o + L.V);
    L.bits = o; o = align_up(o + L.words * 8);
    L.seeds = o; o = align_up(o + (size_t)std::max<uint32_t>(seed_total(j), 1) * 16);
    L.hist = o; o = align_up(o + 65536 * 4);
    L.stats = o; o = align_up(o + kStatCount * 8);
    L.fctl = o; o = align_up(o + 64 * 4);
    L.keys = o; if (flood) o = align_up(o + L.V * 4);
    L.bricks = o; if (flood) o = align_up(o + relax_flag_bytes(j.dims[0], j.dims[1], j.dims[2]));
    L.sub = o; if (needs_sub(j)) o = align_up(o + 8ull * seed_total(j));  // union-find parents + keep flags
    L.erode = o; if (j.erode) o = align_up(o + L.V * 2);                          // erosion output grid
    L.mask = o; if (j.erode) o = align_up(o + (size_t)kMaxErosionSide * kMaxErosionSide * kMaxErosionSide);
    L.mesh_xyz = o; if (!j.mesh_on_device) o = align_up(o + (size_t)j.nv * 12);
    L.mesh_tri = o; if (!j.mesh_on_device) o = align_up(o + (size_t)j.nt * 12);
    L.total = o;
    return L;
}

}  // namespace

size_t job_footprint(const vf_job& j)
{
    const Layout Lo = layout_of(j);
    const uint64_t tiles = ceil_div(Lo.V, kTileCells);
    // the job's share of the engine scratch its launches use: the voxelizer's per-triangle setup records,
    // work counts and their 64-bit scans, the seed-candidate counts per tile and their scan (each buffer
    // at least 256 bytes plus 64 bytes of slack), the launch table's descriptor and prefix entries (with
    // the alignment of its nine arrays), and the flood worklists (2 sets x 8 XCD lists of
    // ceil(tasks / 8) x 4096 eight-byte entries: 64 KiB per task, plus one rounding step)
    const size_t front = (size_t)j.nt * (sizeof(TriSetup) + 8 + 16) + scan_tmp_bytes(2ull * j.nt) +
                         (size_t)tiles * 12 + scan_tmp_bytes(tiles) + 6 * (256 + 64);
    const size_t table = sizeof(JobDev) + 16 * 4 + 6 * 4 + 9 * kAlign;
    const size_t lists = j.algorithm == VF_FLOOD
                             ? 65536ull * relax_tasks(j.dims[0], j.dims[1], j.dims[2]) + 2ull * kFloodXcds * 4096 * 8
                             : 0;
    // blocks come from the engine's size classes: at most an eighth more than asked, and 256 bytes at least
    const size_t shared = front + table + lists;
    return BlockCache::class_size(Lo.total) + shared + shared / 8 + 8 * 256;
}

int job_validate(const vf_job& j, std::string& why)
{
    if (j.poison) { why = "job flagged as poison"; return VF_ERR_POISON; }
    for (int q = 0; q < 3; ++q)
        if (j.dims[q] == 0 || j.dims[q] > 2048) { why = "grid edge out of range"; return VF_ERR_INVALID_ARG; }
    if ((uint64_t)j.dims[0] * j.dims[1] * j.dims[2] >= (1ull << 31)) { why = "grid too large"; return VF_ERR_INVALID_ARG; }
    for (int q = 0; q < 3; ++q)
        if (!(j.aabb_max[q] > j.aabb_min[q])) { why = "empty aabb"; return VF_ERR_INVALID_ARG; }
    if (j.nt && (!j.xyz || !j.tri)) { why = "mesh pointer is NULL"; return VF_ERR_INVALID_ARG; }
    if (j.num_seeds == 0 || j.num_seeds > kMaxDeviceSeeds) { why = "seed count out of range"; return VF_ERR_INVALID_ARG; }
    if (j.seeds) {  // the caller's fragment seeds
        for (uint32_t i = 0; i < j.num_seeds; ++i) {
            const vf_seed& q = j.seeds[i];
            if (q.x >= j.dims[0] || q.y >= j.dims[1] || q.z >= j.dims[2]) { why = "seed outside the grid"; return VF_ERR_INVALID_ARG; }
            if (q.w > 0xFFFF) { why = "seed label above 16 bits"; return VF_ERR_INVALID_ARG; }
            if (j.algorithm == VF_FLOOD && (q.w & 0xFFu) < 2) { why = "flood fragment id (low 8 bits) below 2"; return VF_ERR_TOO_MANY_SEEDS; }
        }
    }
    if (j.impacts) {
        const uint32_t e = std::min(j.dims[0], std::min(j.dims[1], j.dims[2]));
        if (j.impact_spreading == 0 || j.impact_spreading > 64 || j.impact_spreading > e) {
            why = "impact spreading outside 1..min(64, smallest edge)";
            return VF_ERR_INVALID_ARG;
        }
        if (frag_max(j) > kMaxDeviceSeeds) { why = "too many impact seeds"; return VF_ERR_INVALID_ARG; }
        const uint32_t last = j.seeds ? j.seeds[j.num_seeds - 1].w : j.num_seeds + 1;
        if ((uint64_t)last + j.impact_seeds > 0xFFFF) { why = "impact seed labels above 16 bits"; return VF_ERR_INVALID_ARG; }
    }
    if (j.extra_seeds) {
        // mergeSeeds keeps the fragment label in 8 bits
        const uint32_t last = j.seeds ? j.seeds[j.num_seeds - 1].w : j.num_seeds + 1;
        uint32_t top = last + (j.impacts ? j.impact_seeds : 0);
        for (uint32_t i = 0; j.seeds && i < j.num_seeds; ++i) top = std::max(top, j.seeds[i].w);
        if (top > 255) { why = "fragment labels above 8 bits cannot take sub-seeds"; return VF_ERR_TOO_MANY_SEEDS; }
        if (seed_total(j) > kMaxDeviceSeeds || frag_max(j) + j.extra_seeds > kMaxMergeSeeds) {
            why = "too many seeds to merge";
            return VF_ERR_INVALID_ARG;
        }
        if (j.merge_distance < VF_EUCLIDEAN || j.merge_distance > VF_CHEBYSHEV) { why = "unknown merge distance"; return VF_ERR_INVALID_ARG; }
    }
    if (j.erode) {
        std::vector<uint8_t> mask;
        float act = 0;
        if (j.erosion_size == 0 || j.erosion_size > kMaxErosionSide || erosion_mask(j.erosion_convolution, j.erosion_size, mask, act) < 0) {
            why = "bad erosion window";
            return VF_ERR_INVALID_ARG;
        }
        if (j.erosion_iterations > kMaxErosionIterations) { why = "too many erosion iterations"; return VF_ERR_INVALID_ARG; }
    }
    if (j.algorithm != VF_NAIVE && j.algorithm != VF_FLOOD) { why = "unknown algorithm"; return VF_ERR_INVALID_ARG; }
    if (j.distance < VF_EUCLIDEAN || j.distance > VF_CHEBYSHEV) { why = "unknown distance"; return VF_ERR_INVALID_ARG; }
    if (j.seed_location < VF_SEED_INNER || j.seed_location > VF_SEED_BOTH) { why = "unknown seed location"; return VF_ERR_INVALID_ARG; }
    if (j.boundary_size < 0 || j.boundary_size > 64) { why = "boundary size out of range"; return VF_ERR_INVALID_ARG; }
    if (j.algorithm == VF_FLOOD) {
        if (!j.seeds && frag_max(j) > 254) { why = "flood labels must fit 8 bits"; return VF_ERR_TOO_MANY_SEEDS; }
        if (j.seeds && j.impacts && (j.seeds[j.num_seeds - 1].w & 0xFFu) + j.impact_seeds > 255) {
            why = "flood impact seed labels must fit 8 bits";
            return VF_ERR_TOO_MANY_SEEDS;
        }
        const int nb = resolved_neigh(j);
        if (nb != 6 && nb != 26) { why = "neighbourhood must be 0, 6 or 26"; return VF_ERR_INVALID_ARG; }
        if (j.noise_amp > kMaxNoiseAmp) { why = "noise amplitude too large"; return VF_ERR_INVALID_ARG; }
    }
    if (!j.mesh_on_device)
        for (uint64_t i = 0; i < 3ull * j.nt; ++i)
            if (j.tri[i] >= j.nv) { why = "triangle index out of range"; return VF_ERR_INVALID_ARG; }
    return VF_OK;
}

Engine::Engine()
{
    VF_HIP(hipGetDevice(&device));
    VF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    VF_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    VF_HIP(hipEventCreateWithFlags(&side_go, hipEventDisableTiming));
    VF_HIP(hipEventCreateWithFlags(&side_done, hipEventDisableTiming));
    VF_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), 16 * 4 * kMaxTableJobs));
    mc_.cache = &cache_;  // marching-cubes scratch and meshes come from the same block cache
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t thr = ~0ull;  // keep freed blocks cached in the pool
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipGetLastError();
}

Engine::~Engine()
{
    if (s) {
        for (CacheBuf* b : { &vox_setup_, &vox_counts_, &vox_shell_off_, &scan_tmp_, &seed_cnt_, &seed_off_ }) b->release(cache_);
        (void)hipStreamSynchronize(s);
    }
    if (side) (void)hipStreamSynchronize(side);
    if (pinned) (void)hipHostFree(pinned);
    if (side_go) (void)hipEventDestroy(side_go);
    if (side_done) (void)hipEventDestroy(side_done);
    if (side) (void)hipStreamDestroy(side);
    if (s) (void)hipStreamDestroy(s);
}

void Engine::alloc(EJob& e)
{
    const vf_job& j = e.spec;
    const Layout Lo = layout_of(j);
    e.bytes = Lo.total;
    e.mem = cache_.get(Lo.total, &e.bytes);
    auto* base = static_cast<uint8_t*>(e.mem);
    JobDev& d = e.d;
    d = JobDev{};
    d.nx = j.dims[0];
    d.ny = j.dims[1];
    d.nz = j.dims[2];
    d.wp = next_pow2(ceil_div(j.dims[2], 32));
    d.V = Lo.V;
    for (int q = 0; q < 3; ++q) {
        d.mn[q] = j.aabb_min[q];
        d.mx[q] = j.aabb_max[q];
        d.cs[q] = (j.aabb_max[q] - j.aabb_min[q]) / (float)j.dims[q];  // RegularGrid::setAABB
    }
    d.labels = reinterpret_cast<uint16_t*>(base + Lo.labels);
    d.occ = base + Lo.occ;
    d.shell_bits = reinterpret_cast<uint32_t*>(base + Lo.bits);
    d.toggle_bits = d.shell_bits + Lo.words;
    d.seeds = reinterpret_cast<uint4*>(base + Lo.seeds);
    d.S = j.num_seeds + (j.extra_seeds ? j.num_seeds + j.extra_seeds : 0);  // nearSeeds resets it on the device
    d.S_draw = j.seeds ? 0 : j.num_seeds;  // the caller's seeds are copied, not drawn
    d.S0 = j.num_seeds;
    d.F_in = j.num_seeds;
    d.extra = j.extra_seeds;
    d.xoff = 2 * frag_max(j);
    d.impacts = j.impacts;
    d.near_seeds = j.impact_seeds;
    d.spreading = j.impact_spreading;
    d.near_key = j.impact_seed;
    d.merge_dist = j.merge_distance;
    d.dist = j.distance;
    d.key_bits = key_bits_for(seed_total(j));  // any count nearSeeds leaves fits
    d.neighbourhood = resolved_neigh(j);
    d.noise_seed = j.noise_seed;
    d.noise_amp = j.noise_amp;
    d.seed_location = j.seed_location;
    d.seed_key = j.seed;
    d.stats = reinterpret_cast<unsigned long long*>(base + Lo.stats);
    d.hist = reinterpret_cast<uint32_t*>(base + Lo.hist);
    d.fctl = reinterpret_cast<uint32_t*>(base + Lo.fctl);
    d.active = 1;
    d.algo = j.algorithm;
    d.bsize = j.erode ? 0 : j.boundary_size;  // erosion runs its own boundary detection
    if (j.algorithm == VF_FLOOD) {
        d.keys = reinterpret_cast<uint32_t*>(base + Lo.keys);
        d.bflags = reinterpret_cast<unsigned long long*>(base + Lo.bricks);
        d.nbricks = relax_bricks(d.nx, d.ny, d.nz);
        d.sub = needs_sub(j) ? reinterpret_cast<uint32_t*>(base + Lo.sub) : nullptr;
    } else {
        const NaiveKeyForm kf = naive_key_form(j.dims, j.distance, seed_total(j));
        d.dist = kf.dist;
        d.naive_key64 = kf.key64;
    }
    if (j.mesh_on_device) {
        d.xyz = j.xyz;
        d.tri = j.tri;
    } else {
        d.xyz = reinterpret_cast<const float*>(base + Lo.mesh_xyz);
        d.tri = reinterpret_cast<const uint32_t*>(base + Lo.mesh_tri);
    }
    upload_inputs(e);
    d.nt = j.nt;
    d.nv = j.nv;
    e.erode_out = nullptr;
    if (j.erode) {  // RegularGrid::erode's window and output grid
        std::vector<uint8_t> mask;
        float act = 0;
        e.erode_side = (uint32_t)erosion_mask(j.erosion_convolution, j.erosion_size, mask, act);
        e.erode_limit = act * j.erosion_threshold;
        e.erode_out = reinterpret_cast<uint16_t*>(base + Lo.erode);
        e.erode_mask = base + Lo.mask;
        VF_HIP(hipMemcpyAsync(e.erode_mask, mask.data(), mask.size(), hipMemcpyHostToDevice, s));
        VF_HIP(hipStreamSynchronize(s));  // the pageable mask goes out of scope
    }
    e.words = Lo.words;
    e.tiles = ceil_div(Lo.V, kTileCells);
    e.groups = ceil_div(Lo.words, 256);
}

void Engine::upload_inputs(EJob& e)
{
    const vf_job& j = e.spec;
    if (j.seeds) VF_HIP(hipMemcpyAsync(e.d.seeds, j.seeds, (size_t)j.num_seeds * sizeof(vf_seed), hipMemcpyHostToDevice, s));
    if (j.mesh_on_device) return;
    if (j.nv) VF_HIP(hipMemcpyAsync(const_cast<float*>(e.d.xyz), j.xyz, (size_t)j.nv * 12, hipMemcpyHostToDevice, s));
    if (j.nt) VF_HIP(hipMemcpyAsync(const_cast<uint32_t*>(e.d.tri), j.tri, (size_t)j.nt * 12, hipMemcpyHostToDevice, s));
}

void Engine::release(EJob& e)
{
    cache_.put(e.mem, e.bytes);  // reused only by work queued after the job's own
    e.mem = nullptr;
}

Table* Engine::make_table(const std::vector<EJob*>& js)
{
    const uint32_t n = (uint32_t)js.size();
    if (n == 0 || n > kMaxTableJobs) fail(VF_ERR_INVALID_ARG, "table of %u jobs (1..%u)", n, kMaxTableJobs);
    auto* t = new Table();
    t->jobs = js;
    const size_t o_jobs = 0;
    const size_t o_tile = align_up(o_jobs + n * sizeof(JobDev));
    const size_t o_tri = align_up(o_tile + (n + 1) * 4);
    const size_t o_col = align_up(o_tri + (n + 1) * 4);
    const size_t o_bt = align_up(o_col + (n + 1) * 4);
    const size_t o_gather = align_up(o_bt + (n + 1) * 4);
    const size_t o_done = align_up(o_gather + 16 * 4 * (size_t)n);
    const size_t o_fcnt = align_up(o_done + 4 * ((size_t)n + 1));
    const size_t o_xjobs = align_up(o_fcnt + 4 * 2 * kFloodXcds);  // extra-seed draws (mergeSeeds)
    const size_t o_ejobs = align_up(o_xjobs + n * sizeof(JobDev)); // erosion passes (boundary size 1)
    const size_t o_local = align_up(o_ejobs + n * sizeof(JobDev)); // per job: {0, tiles}, {0, bricks}
    const size_t total = align_up(o_local + 16 * (size_t)n);
    std::vector<uint8_t> img(total, 0);
    auto* xj = reinterpret_cast<JobDev*>(img.data() + o_xjobs);
    auto* ej = reinterpret_cast<JobDev*>(img.data() + o_ejobs);
    auto* lp = reinterpret_cast<uint32_t*>(img.data() + o_local);
    auto* jd = reinterpret_cast<JobDev*>(img.data() + o_jobs);
    auto* tp = reinterpret_cast<uint32_t*>(img.data() + o_tile);
    auto* rp = reinterpret_cast<uint32_t*>(img.data() + o_tri);
    auto* cp = reinterpret_cast<uint32_t*>(img.data() + o_col);
    auto* bp = reinterpret_cast<uint32_t*>(img.data() + o_bt);
    auto* fp = reinterpret_cast<uint32_t*>(img.data() + o_done);
    uint64_t tiles = 0, tris = 0, groups = 0, btiles = 0, fchunks = 0;
    uint32_t max_edge = 1;
    for (uint32_t i = 0; i < n; ++i) {
        EJob& e = *js[i];
        jd[i] = e.d;
        tp[i] = (uint32_t)tiles;
        rp[i] = (uint32_t)tris;
        cp[i] = (uint32_t)groups;
        bp[i] = (uint32_t)btiles;
        fp[i] = (uint32_t)fchunks;
        btiles += brick_count(e.d.nx, e.d.ny, e.d.nz);
        if (e.d.algo == VF_FLOOD) fchunks += relax_tasks(e.d.nx, e.d.ny, e.d.nz);
        max_edge = std::max(max_edge, std::max(e.d.nx, std::max(e.d.ny, e.d.nz)));
        t->any_wide |= e.d.bsize > 1;
        tiles += e.tiles;
        tris += e.d.nt;
        groups += e.groups;
        t->any_flood |= e.d.algo == VF_FLOOD;
        t->any_naive32 |= e.d.algo == VF_NAIVE && !e.d.naive_key64;
        t->any_naive64 |= e.d.algo == VF_NAIVE && e.d.naive_key64;
        // the extra draw: S_draw = extra seeds anywhere, with their own key, written after the fragments'
        // copies; jobs without extra seeds sit it out
        xj[i] = e.d;
        xj[i].active = e.d.extra ? 1 : 0;
        xj[i].S_draw = e.d.extra;
        xj[i].seed_key = e.spec.extra_seed;
        xj[i].seed_location = VF_SEED_BOTH;
        xj[i].seeds = e.d.seeds + e.d.xoff;
        t->any_extra |= e.d.extra != 0;
        t->any_near |= e.d.impacts != 0;
        t->any_sub |= e.d.sub != nullptr;
        ej[i] = e.d;
        ej[i].bsize = 1;
        t->any_erode |= e.spec.erode != 0;
        lp[4 * i + 0] = 0;
        lp[4 * i + 1] = e.tiles;
        lp[4 * i + 2] = 0;
        lp[4 * i + 3] = brick_count(e.d.nx, e.d.ny, e.d.nz);
    }
    if (tiles > 0xFFFFFFFFull || tris > 0xFFFFFFFFull || groups > 0xFFFFFFFFull || fchunks > 0xFFFFFFFFull)
        fail(VF_ERR_INVALID_ARG, "table too large for 32-bit prefixes");
    tp[n] = (uint32_t)tiles;
    rp[n] = (uint32_t)tris;
    cp[n] = (uint32_t)groups;
    bp[n] = (uint32_t)btiles;
    fp[n] = (uint32_t)fchunks;
    t->dev = cache_.get(total, &t->dev_bytes);
    VF_HIP(hipMemcpyAsync(t->dev, img.data(), total, hipMemcpyHostToDevice, s));
    VF_HIP(hipStreamSynchronize(s));  // the pageable source must outlive the copy
    auto* base = static_cast<uint8_t*>(t->dev);
    t->L.jobs = reinterpret_cast<const JobDev*>(base + o_jobs);
    t->L.njobs = n;
    t->L.tile_prefix = reinterpret_cast<const uint32_t*>(base + o_tile);
    t->L.total_tiles = (uint32_t)tiles;
    t->L.s = s;
    t->tri_prefix = reinterpret_cast<const uint32_t*>(base + o_tri);
    t->col_prefix = reinterpret_cast<const uint32_t*>(base + o_col);
    t->bt_prefix = reinterpret_cast<const uint32_t*>(base + o_bt);
    t->total_btiles = (uint32_t)btiles;
    t->total_tris = (uint32_t)tris;
    t->total_groups = (uint32_t)groups;
    t->gather = reinterpret_cast<uint32_t*>(base + o_gather);
    t->bucket = tune_bucket(max_edge);
    t->xL = t->L;
    t->xL.jobs = reinterpret_cast<const JobDev*>(base + o_xjobs);
    t->ejobs = reinterpret_cast<const JobDev*>(base + o_ejobs);
    t->local_prefix = reinterpret_cast<const uint32_t*>(base + o_local);
    t->flood.task_prefix = reinterpret_cast<const uint32_t*>(base + o_done);
    t->flood.total_tasks = (uint32_t)fchunks;
    t->flood.counts = reinterpret_cast<uint32_t*>(base + o_fcnt);
    if (fchunks) {
        t->flood.list_cap = flood_list_cap((uint32_t)fchunks);
        t->flood_lists = cache_.get(2ull * kFloodXcds * t->flood.list_cap * 8, &t->list_bytes);
        t->flood.lists = static_cast<unsigned long long*>(t->flood_lists);
    }
    bool counts = false;  // jobs whose nearSeeds already ran: take its seed counts
    for (const EJob* e : js) counts |= e->d.impacts && e->fronted;
    if (counts) launch_seed_counts(t->L);
    return t;
}

void Engine::drop_table(Table* t)
{
    if (!t) return;
    cache_.put(t->dev, t->dev_bytes);
    cache_.put(t->flood_lists, t->list_bytes);
    delete t;
}

void Engine::reserve_front(const Table& t)
{
    const uint32_t nt = t.total_tris;
    vox_setup_.ensure((size_t)std::max<uint32_t>(nt, 1) * sizeof(TriSetup), cache_);
    vox_counts_.ensure((size_t)std::max<uint32_t>(nt, 1) * 8, cache_);
    vox_shell_off_.ensure((2 * (size_t)nt + 1) * 8, cache_);
    scan_tmp_.ensure(std::max(scan_tmp_bytes(2 * (uint64_t)nt), scan_tmp_bytes(t.L.total_tiles)) + 64, cache_);
    seed_cnt_.ensure((size_t)t.L.total_tiles * 4 + 64, cache_);
    seed_off_.ensure(((size_t)t.L.total_tiles + 1) * 8 + 64, cache_);
}

void Engine::release_front()
{
    for (CacheBuf* b : { &vox_setup_, &vox_counts_, &vox_shell_off_, &scan_tmp_, &seed_cnt_, &seed_off_ }) b->release(cache_);
}

size_t Engine::device_bytes() const { return cache_.held(); }

void Engine::make_room(size_t need, size_t budget)
{
    const size_t held = device_bytes();
    if (held + need <= budget) return;
    const size_t excess = held + need - budget;
    cache_.trim(cache_.cached() > excess ? cache_.cached() - excess : 0);
}

// Launch-only (safe inside stream capture): reserve_front must have sized the scratch.
void Engine::front(Table& t, bool split)
{
    set_active_bucket(t.bucket);
    for (EJob* e : t.jobs) e->fronted = true;
    const Launch& L = t.L;
    launch_job_init(L);
    VoxScratch X;
    const uint32_t nt = t.total_tris;
    X.setup = vox_setup_.as<TriSetup>();
    X.counts = vox_counts_.as<uint32_t>();
    X.shell_off = vox_shell_off_.as<unsigned long long>();
    X.tri_prefix = t.tri_prefix;
    X.total_tris = nt;
    X.scan_tmp = scan_tmp_.p;
    X.scan_tmp_bytes = scan_tmp_.bytes;
    launch_voxelize(L, X, t.col_prefix, t.total_groups, seed_cnt_.as<uint32_t>());  // counts tiles of VF_SEED_BOTH jobs
    launch_seed_count(L, seed_cnt_.as<uint32_t>(), true);
    scan_exclusive_u32_to_u64(seed_cnt_.as<uint32_t>(), seed_off_.as<unsigned long long>(), L.total_tiles, scan_tmp_.p, s);
    launch_seed_pick(L, seed_off_.as<unsigned long long>());
    if (t.any_near) launch_seed_near_jobs(L);  // nearSeeds after the fragment seeds
    if (t.any_extra) {  // mergeSeeds: the extra draw over all occupied cells, then the merge
        launch_seed_count(t.xL, seed_cnt_.as<uint32_t>(), false);
        scan_exclusive_u32_to_u64(seed_cnt_.as<uint32_t>(), seed_off_.as<unsigned long long>(), L.total_tiles, scan_tmp_.p, s);
        launch_seed_pick(t.xL, seed_off_.as<unsigned long long>());
        launch_seed_merge(L);
    }
    if (split) return;
    if (t.any_naive32) launch_naive(L, false);
    if (t.any_naive64) launch_naive(L, true);
    flood_start(t);
}

void Engine::flood_start(Table& t)
{
    if (!t.any_flood) return;
    set_active_bucket(t.bucket);
    launch_flood_init(t.L, t.flood);
    launch_flood_seed(t.L, t.flood);
}

// The naive jobs and the flood jobs of a table share no memory, so the naive fracture (VALU-bound) can run
// on a second stream while the flood's key init (HBM-bound) and its rounds, whose later ones leave most of
// the device idle, run on s.
void Engine::naive_side(Table& t)
{
    if (!t.any_naive32 && !t.any_naive64) return;
    set_active_bucket(t.bucket);
    VF_HIP(hipEventRecord(side_go, s));
    VF_HIP(hipStreamWaitEvent(side, side_go, 0));
    Launch L = t.L;
    L.s = side;
    if (t.any_naive32) launch_naive(L, false);
    if (t.any_naive64) launch_naive(L, true);
    VF_HIP(hipEventRecord(side_done, side));
    side_pending_ = true;
}

void Engine::join_side()
{
    if (!side_pending_) return;
    VF_HIP(hipStreamWaitEvent(s, side_done, 0));
    side_pending_ = false;
}

void Engine::rounds(Table& t, uint32_t n)
{
    set_active_bucket(t.bucket);
    if (t.flood.total_tasks) launch_flood_rounds(t.L, t.flood, n, active_shape().flood_blocks);
}

void Engine::post(Table& t)
{
    set_active_bucket(t.bucket);
    const Launch& L = t.L;
    if (t.any_flood) launch_flood_finish(L);
    if (t.any_erode) erode(t);
    launch_boundaries_tiled(L, t.bt_prefix, t.total_btiles, t.any_wide, true);  // + grid checksum
    launch_hist_summary(L);
    launch_gather(L, t.gather);
}

// RegularGrid::erode (RegularGrid.cpp:82-159) for each job that asks for it, in place of boundary detection:
// iterations x (detectBoundaries(1), one erosion step), then the isolated-cell pass; its histogram is
// cleared afterwards (the boundary passes count into it) and recounted by the post stage.  Launch-only.
void Engine::erode(Table& t)
{
    for (uint32_t i = 0; i < t.L.njobs; ++i) {
        const EJob& e = *t.jobs[i];
        if (!e.spec.erode) continue;
        Launch one = t.L;
        one.jobs = t.ejobs + i;
        one.njobs = 1;
        one.tile_prefix = t.local_prefix + 4 * i;
        one.total_tiles = e.tiles;
        const uint32_t bricks = brick_count(e.d.nx, e.d.ny, e.d.nz);
        for (uint32_t it = 0; it < e.spec.erosion_iterations; ++it) {
            launch_boundaries_tiled(one, t.local_prefix + 4 * i + 2, bricks, false, false);
            launch_erode_step(one, e.erode_out, e.erode_mask, e.erode_side, e.erode_limit, e.spec.erosion_probability,
                              e.spec.erosion_seed, e.d.V);
            VF_HIP(hipMemcpyAsync(e.d.labels, e.erode_out, e.d.V * 2, hipMemcpyDeviceToDevice, s));
        }
        launch_isolated_cells(one, e.erode_out, e.d.V);
        VF_HIP(hipMemcpyAsync(e.d.labels, e.erode_out, e.d.V * 2, hipMemcpyDeviceToDevice, s));
        launch_clear_hist(one);
    }
}

void Engine::flood_sub(Table& t)
{
    set_active_bucket(t.bucket);
    launch_flood_sub(t.L);
}

void Engine::gather(Table& t, uint32_t* host16)
{
    launch_gather(t.L, t.gather);
    VF_HIP(hipMemcpyAsync(host16, t.gather, 16 * 4 * (size_t)t.L.njobs, hipMemcpyDeviceToHost, s));
}

void Engine::finish(EJob& e, const uint32_t* w)
{
    auto stat = [&](int k) { return (uint64_t)w[2 * k] | ((uint64_t)w[2 * k + 1] << 32); };
    vf_result& r = e.res;
    r.voxels = e.d.V;
    r.occupied = stat(kStatOccupied);
    r.checksum = stat(kStatChecksum);
    r.num_labels = (uint32_t)stat(kStatLabels);
    r.flood_levels = (uint32_t)stat(kStatFloodDone);
    r.flood_bricks = (uint32_t)stat(kStatFloodWork);
    r.flood_sweeps = (uint32_t)(stat(kStatFloodWork) >> 32);
    const uint64_t err = stat(kStatError);
    r.status = VF_OK;
    if (err & kDevErrMesh) r.status = VF_ERR_INVALID_ARG;
    else if (err & kDevErrSeeding) r.status = VF_ERR_SEEDING;
    else if (err & kDevErrMerge) r.status = VF_ERR_TOO_MANY_SEEDS;
    else if (err) r.status = VF_ERR_OVERFLOW;
    r.mc_vertices = r.mc_triangles = 0;
    if (e.spec.labels_out)
        VF_HIP(hipMemcpyAsync(e.spec.labels_out, e.d.labels, e.d.V * 2, hipMemcpyDeviceToHost, s));
    if (e.spec.histogram_out)
        VF_HIP(hipMemcpyAsync(e.spec.histogram_out, e.d.hist, 65536 * 4, hipMemcpyDeviceToHost, s));
    if (e.spec.labels_out || e.spec.histogram_out) VF_HIP(hipStreamSynchronize(s));
}

void Engine::extract(const std::vector<EJob*>& js, const std::function<bool(size_t)>& reserve)
{
    std::vector<EJob*> mj;
    for (EJob* e : js)
        if (e->spec.marching_cubes && e->res.status == VF_OK) mj.push_back(e);
    if (mj.empty()) return;
    const uint32_t n = (uint32_t)mj.size();
    std::vector<McJobSpec> spec(n);
    std::vector<McJobResult> res(n);
    std::vector<McSmooth> sm(n);
    std::vector<vf_mesh_out*> outs(n);
    bool smooth = false;
    for (uint32_t i = 0; i < n; ++i) {
        const EJob& e = *mj[i];
        McJobSpec& q = spec[i];
        q.grid = e.d.labels;
        q.hist = e.d.hist;
        q.dims[0] = e.d.nx;
        q.dims[1] = e.d.ny;
        q.dims[2] = e.d.nz;
        for (int a = 0; a < 3; ++a) {
            q.mn[a] = e.d.mn[a];
            q.cs[a] = e.d.cs[a];
        }
        q.want = VF_ALL_LABELS;
        sm[i] = McSmooth{ e.spec.smooth_iterations[0], e.spec.smooth_weights[0], e.spec.smooth_iterations[1],
                          e.spec.smooth_weights[1] };
        smooth |= sm[i].nb_iters || sm[i].b_iters;
        outs[i] = e.spec.mesh_out;
    }
    mc_run_jobs(mc_, s, spec.data(), n, res.data(), reserve);
    if (mc_.valid && smooth) mc_smooth_jobs(mc_, s, sm.data(), n);
    mc_read_jobs(mc_, s, res.data(), outs.data(), n);
    for (uint32_t i = 0; i < n; ++i) {
        vf_result& r = mj[i]->res;
        r.mc_vertices = res[i].nv;
        r.mc_triangles = res[i].nt;
        if (res[i].status != VF_OK) r.status = res[i].status;
    }
}

void Engine::release_extract() { mc_.release_all(); }

}  // namespace vf
