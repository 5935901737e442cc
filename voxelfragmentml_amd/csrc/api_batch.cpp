// C ABI for the batched, plain and served forms.
//
//   vf_batch_run   all jobs share every stage launch; the front, the flood
//                  round chunk and the post stages are captured once into HIP
//                  graphs and replayed while the job set is unchanged.
//   vf_plain_run   the same stages one job at a time, one job per launch, with
//                  a host check after every few flood rounds (the reference's
//                  per-dispatch round trip, FloodFracturer.cpp:143-158, batched
//                  to chunks).
//   vf_server_*    continuous batching: a scheduler thread admits queued jobs
//                  against an HBM budget and a slot limit, runs their front
//                  stages together, lets new flood jobs join the shared round
//                  launches at their own round, and retires drained jobs, so
//                  work of jobs that arrived at different times overlaps.
#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "engine.hpp"

using namespace vf;

namespace {

constexpr uint32_t kServerRounds = 4; // flood rounds per scheduler tick
constexpr size_t kPlainCacheKeep = 16ull << 30;  // device bytes the plain form keeps cached between calls

uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

uint64_t job_signature(const vf_job* jobs, uint32_t n)
{
    uint64_t h = n;
    for (uint32_t i = 0; i < n; ++i) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(jobs + i);
        for (size_t k = 0; k + 8 <= sizeof(vf_job); k += 8) {
            uint64_t v;
            std::memcpy(&v, p + k, 8);
            h = mix64(h, v);
        }
    }
    return h;
}

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    void reset()
    {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        x = nullptr;
        g = nullptr;
    }
    ~Graph() { reset(); }
};

template <class F>
void run_captured(hipStream_t s, Graph& G, bool use, F&& body)
{
    if (!use) {
        body();
        return;
    }
    if (!G.x) {
        VF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(s, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        VF_HIP(hipStreamEndCapture(s, &G.g));
        VF_HIP(hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0));
    }
    VF_HIP(hipGraphLaunch(G.x, s));
}

bool all_drained(const Table& t, const uint32_t* w)
{
    for (size_t i = 0; i < t.jobs.size(); ++i)
        if (t.jobs[i]->d.algo == VF_FLOOD && w[16 * i + 2 * kStatFloodDone] == 0 &&
            w[16 * i + 2 * kStatFloodDone + 1] == 0 && w[16 * i + 2 * kStatError] == 0)
            return false;
    return true;
}

}  // namespace

// ---------------------------------------------------------------- batch ---
struct vf_batch {
    Engine eng;
    bool use_graphs = true;
    uint64_t sig = 0;
    uint64_t tune_gen = 0;           // launch shapes the graphs were captured with
    std::vector<EJob> jobs;          // every submitted job (invalid ones keep their status)
    std::vector<EJob*> live;         // jobs on the device
    Table* table = nullptr;
    Graph g_front, g_fstart, g_rounds, g_post;
    uint32_t drain_chunks[2] = { 0, 0 };  // chunks the last run needed before its drain and after the sub-seed pass

    void clear()
    {
        g_front.reset();
        g_fstart.reset();
        g_rounds.reset();
        g_post.reset();
        if (table) eng.drop_table(table);
        table = nullptr;
        for (auto* e : live) eng.release(*e);
        live.clear();
        jobs.clear();
        (void)hipStreamSynchronize(eng.s);
        sig = 0;
        drain_chunks[0] = drain_chunks[1] = 0;
    }
    ~vf_batch() { clear(); }
};

static void batch_prepare(vf_batch* b, const vf_job* jobs, uint32_t n)
{
    const uint64_t gen = tune_generation();
    if (gen != b->tune_gen) {  // launch shapes changed: recapture
        b->g_front.reset();
        b->g_fstart.reset();
        b->g_rounds.reset();
        b->g_post.reset();
        b->tune_gen = gen;
    }
    const uint64_t sig = job_signature(jobs, n);
    if (sig == b->sig && b->jobs.size() == n) {
        // the same descriptors: device state is reused, but host meshes and seeds are read again (a
        // caller may have refilled the buffers in place), after validating them again
        bool same = true;
        for (uint32_t i = 0; i < n && same; ++i) {
            std::string why;
            if (!jobs[i].mesh_on_device || jobs[i].seeds) same = job_validate(jobs[i], why) == b->jobs[i].vstatus;
        }
        if (same) {
            for (auto* e : b->live) b->eng.upload_inputs(*e);
            return;
        }
    }
    b->clear();
    b->jobs.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        EJob& e = b->jobs[i];
        e.spec = jobs[i];
        std::string why;
        e.res = vf_result{};
        e.res.status = e.vstatus = job_validate(e.spec, why);
        e.res.voxels = (uint64_t)e.spec.dims[0] * e.spec.dims[1] * e.spec.dims[2];
        if (e.res.status == VF_OK) b->live.push_back(&e);
    }
    if (b->live.size() > Engine::kMaxTableJobs) fail(VF_ERR_INVALID_ARG, "at most %u jobs per batch", Engine::kMaxTableJobs);
    for (auto* e : b->live) b->eng.alloc(*e);
    if (!b->live.empty()) {
        b->table = b->eng.make_table(b->live);
        b->eng.reserve_front(*b->table);
    }
    b->sig = sig;
}

extern "C" {

int vf_batch_create(vf_batch** out)
{
    return guard([&] {
        if (!out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        if (vf_device_count() < 1) fail(VF_ERR_DEVICE, "no HIP device available");
        *out = new vf_batch();
    });
}

int vf_batch_destroy(vf_batch* b)
{
    return guard([&] { delete b; });
}

int vf_batch_set_graphs(vf_batch* b, int use_graphs)
{
    return guard([&] {
        if (!b) fail(VF_ERR_INVALID_ARG, "batch is NULL");
        b->clear();
        b->use_graphs = use_graphs != 0;
    });
}

int vf_batch_run(vf_batch* b, const vf_job* jobs, uint32_t n, vf_result* results)
{
    return guard([&] {
        if (!b || (n && (!jobs || !results))) fail(VF_ERR_INVALID_ARG, "NULL argument");
        const double t0 = now_ms();
        batch_prepare(b, jobs, n);
        Engine& E = b->eng;
        if (b->table) {
            Table& T = *b->table;
            run_captured(E.s, b->g_front, b->use_graphs, [&] { E.front(T, T.any_flood); });
            if (T.any_flood) {
                E.naive_side(T);  // beside the flood's init and rounds
                run_captured(E.s, b->g_fstart, b->use_graphs, [&] { E.flood_start(T); });
                // the first chunks run without a check; then one readback per chunk.  Jobs with merged
                // seeds run the sub-seed pass after the first drain and grow again where it freed cells.
                // A batch run again runs as many chunks before its first check as its last run needed
                // (the rounds a flood needs hardly change between runs of the same jobs; rounds after a
                // drain find nothing to do), so a replay waits on the host about once per phase.
                const LaunchShape sh = get_shape(T.bucket);
                uint32_t chunks = 0, lead = std::max(sh.lead_chunks, b->drain_chunks[0]);
                bool sub_done = !T.any_sub;
                for (;;) {
                    run_captured(E.s, b->g_rounds, b->use_graphs, [&] { E.rounds(T, sh.round_chunk); });
                    if (++chunks < lead) continue;
                    E.gather(T, E.pinned);
                    VF_HIP(hipStreamSynchronize(E.s));
                    if (all_drained(T, E.pinned)) {
                        b->drain_chunks[sub_done && T.any_sub ? 1 : 0] = chunks;
                        if (sub_done) break;
                        E.flood_sub(T);
                        sub_done = true;
                        chunks = 0;
                        lead = std::max(1u, b->drain_chunks[1]);  // check after the next chunk, or as the last run
                    }
                    if ((uint64_t)chunks * sh.round_chunk > (1ull << 24)) fail(VF_ERR_OVERFLOW, "flood did not drain");
                }
            }
            E.join_side();
            run_captured(E.s, b->g_post, b->use_graphs, [&] { E.post(T); });
            VF_HIP(hipMemcpyAsync(E.pinned, T.gather, 16 * 4 * (size_t)T.L.njobs, hipMemcpyDeviceToHost, E.s));
            VF_HIP(hipStreamSynchronize(E.s));
            std::vector<uint32_t> w(E.pinned, E.pinned + 16 * T.L.njobs);
            for (size_t i = 0; i < T.jobs.size(); ++i) E.finish(*T.jobs[i], w.data() + 16 * i);
            E.extract(T.jobs);  // marching cubes of the jobs that ask for it, all in one launch set
        }
        const double t1 = now_ms();
        for (uint32_t i = 0; i < n; ++i) {
            results[i] = b->jobs[i].res;
            results[i].latency_ms = t1 - t0;
            results[i].queue_ms = 0;
        }
    });
}

int vf_plain_run(const vf_job* jobs, uint32_t n, vf_result* results)
{
    return guard([&] {
        if (n && (!jobs || !results)) fail(VF_ERR_INVALID_ARG, "NULL argument");
        static std::mutex mu;
        std::lock_guard<std::mutex> lk(mu);
        static Engine* E = nullptr;
        if (!E) E = new Engine();
        for (uint32_t i = 0; i < n; ++i) {
            const double t0 = now_ms();
            EJob e;
            e.spec = jobs[i];
            std::string why;
            e.res = vf_result{};
            e.res.voxels = (uint64_t)e.spec.dims[0] * e.spec.dims[1] * e.spec.dims[2];
            e.res.status = job_validate(e.spec, why);
            if (e.res.status == VF_OK) {
                E->alloc(e);
                std::vector<EJob*> one{ &e };
                Table* T = E->make_table(one);
                E->reserve_front(*T);
                E->front(*T);
                if (T->any_flood) {
                    uint32_t chunk = 4, total = 0;
                    bool sub_done = !T->any_sub;
                    for (;;) {
                        E->rounds(*T, chunk);
                        total += chunk;
                        E->gather(*T, E->pinned);
                        VF_HIP(hipStreamSynchronize(E->s));
                        if (all_drained(*T, E->pinned)) {
                            if (sub_done) break;
                            E->flood_sub(*T);  // merged seeds: sub-seed pass, then grow again
                            sub_done = true;
                            E->gather(*T, E->pinned);
                            VF_HIP(hipStreamSynchronize(E->s));
                            if (all_drained(*T, E->pinned)) break;
                            chunk = 4;
                        }
                        chunk = std::min<uint32_t>(chunk * 2, 32);
                        if (total > (1u << 24)) fail(VF_ERR_OVERFLOW, "flood did not drain");
                    }
                }
                E->post(*T);
                VF_HIP(hipMemcpyAsync(E->pinned, T->gather, 64, hipMemcpyDeviceToHost, E->s));
                VF_HIP(hipStreamSynchronize(E->s));
                std::vector<uint32_t> w(E->pinned, E->pinned + 16);
                E->finish(e, w.data());
                E->extract({ &e });
                E->release_extract();
                E->drop_table(T);
                E->release(e);
                VF_HIP(hipStreamSynchronize(E->s));
            }
            results[i] = e.res;
            results[i].latency_ms = now_ms() - t0;
            results[i].queue_ms = 0;
        }
        E->trim_cache(kPlainCacheKeep);  // blocks are reused across jobs and calls, up to a bound
    });
}

}  // extern "C"

// --------------------------------------------------------------- server ---
struct vf_server {
    size_t budget = 0;
    uint32_t max_slots = 0;
    std::mutex mu;
    std::condition_variable cv;       // wakes the scheduler
    std::condition_variable done_cv;  // wakes waiters
    std::deque<EJob*> queue;
    std::unordered_map<uint64_t, EJob*> jobs;  // every ticket not yet collected
    size_t reserved = 0;
    size_t device_bytes = 0;          // engine device memory at the last audit (in use + cached)
    size_t peak_device_bytes = 0;     // largest such figure seen
    size_t max_unreserved = 0;        // largest excess of the engine's in-use bytes over reserved bytes seen
    uint32_t in_flight = 0;
    uint64_t completed = 0;
    uint64_t next_ticket = 1;
    bool stop = false;
    int device = 0;
    std::thread th;
    std::string fatal;  // scheduler failure, reported to every caller

    void loop();
};

// Wall time of the scheduler's phases, printed at shutdown when VF_SERVER_TRACE is set.
struct ServerTrace {
    bool on = std::getenv("VF_SERVER_TRACE") != nullptr;
    double t[12] = {};
    uint64_t n[12] = {};
    uint64_t ticks = 0, round_jobs = 0;
    static constexpr const char* names[12] = { "admit: job blocks", "front", "naive post+sync", "flood table",
                                                "rounds+gather+sync", "drained post+sync", "idle wait", "other",
                                                "admit: table", "admit: scratch", "drained: table", "drained: release" };
    double mark = now_ms();
    void step(int k)
    {
        if (!on) return;
        const double t1 = now_ms();
        t[k] += t1 - mark;
        ++n[k];
        mark = t1;
    }
    void print() const
    {
        if (!on) return;
        std::fprintf(stderr, "vf_server trace: %llu ticks, %llu job-rounds chunks\n", (unsigned long long)ticks,
                     (unsigned long long)round_jobs);
        for (int k = 0; k < 12; ++k)
            if (n[k]) std::fprintf(stderr, "  %-22s %10.1f ms  %8llu times  %8.3f ms each\n", names[k], t[k],
                                   (unsigned long long)n[k], t[k] / n[k]);
    }
};

void vf_server::loop()
{
    (void)hipSetDevice(device);
    ServerTrace tr;
    try {
        Engine E;
        // engine allocations never exceed the bytes reserved by admitted jobs: checked after every step
        // that allocates (job blocks, tables, worklists, front scratch)
        auto audit = [&] {
            const size_t dev = E.device_bytes(), in_use = dev - E.cached_bytes();
            std::lock_guard<std::mutex> lk(mu);
            device_bytes = dev;
            peak_device_bytes = std::max(peak_device_bytes, dev);
            if (in_use > reserved) max_unreserved = std::max(max_unreserved, in_use - reserved);
        };
        // marching cubes of a group of finished jobs: its buffers are sized only once the surfaces are
        // counted, so they are reserved then (cached blocks give way first); a group that cannot fit the
        // budget fails with VF_ERR_BUDGET.  The meshes leave with the results, the buffers go back.
        size_t mc_reserved = 0;
        auto extract = [&](std::vector<EJob*>& group) {
            auto reserve = [&](size_t bytes) {
                E.make_room(bytes, budget);
                std::lock_guard<std::mutex> lk(mu);
                if (reserved + bytes > budget) return false;
                reserved += bytes;
                mc_reserved += bytes;
                return true;
            };
            E.extract(group, reserve);
            audit();
            E.release_extract();
            std::lock_guard<std::mutex> lk(mu);
            reserved -= mc_reserved;
            mc_reserved = 0;
        };
        std::vector<EJob*> leveling;  // flood jobs between front and drain
        Table* lev = nullptr;
        bool lev_dirty = false;
        auto complete = [&](EJob* e, int status) {
            std::lock_guard<std::mutex> lk(mu);
            if (e->admitted) {  // return the slot and the budget (the block itself is freed stream-ordered)
                reserved -= job_footprint(e->spec);
                --in_flight;
                e->admitted = false;
            }
            e->res.status = e->cancelled ? VF_ERR_CANCELLED : status;
            e->res.latency_ms = now_ms() - e->t_submit;
            e->res.queue_ms = (e->t_admit > 0 ? e->t_admit : now_ms()) - e->t_submit;
            e->phase = 2;
            ++completed;
            done_cv.notify_all();
        };
        for (;;) {
            std::vector<EJob*> admitted;
            ++tr.ticks;
            tr.step(7);
            {
                std::unique_lock<std::mutex> lk(mu);
                if (queue.empty() && leveling.empty())
                    cv.wait_for(lk, std::chrono::milliseconds(2), [&] { return stop || !queue.empty(); });
                if (stop && queue.empty() && leveling.empty()) break;
                // queued work that cannot be admitted yet and nothing in flight to wait on:
                // sleep briefly instead of spinning (only a cancel or a new job changes this)
                if (!queue.empty() && leveling.empty() && in_flight >= max_slots)
                    cv.wait_for(lk, std::chrono::milliseconds(1));
                while (!queue.empty() && in_flight + admitted.size() < max_slots) {
                    EJob* e = queue.front();
                    if (e->cancelled) {
                        queue.pop_front();
                        e->res.status = VF_ERR_CANCELLED;
                        e->res.latency_ms = now_ms() - e->t_submit;
                        e->phase = 2;
                        ++completed;
                        done_cv.notify_all();
                        continue;
                    }
                    const size_t fp = job_footprint(e->spec);
                    if (fp > budget) {  // can never fit
                        queue.pop_front();
                        e->res.status = VF_ERR_BUDGET;
                        e->res.latency_ms = now_ms() - e->t_submit;
                        e->phase = 2;
                        ++completed;
                        done_cv.notify_all();
                        continue;
                    }
                    if (reserved + fp > budget) break;  // wait for memory to free up
                    queue.pop_front();
                    reserved += fp;
                    e->admitted = true;
                    admitted.push_back(e);
                }
                in_flight += (uint32_t)admitted.size();
            }
            tr.step(6);
            // front stages of the newly admitted jobs, in one set of launches
            if (!admitted.empty()) {
                size_t need = 0;
                for (auto* e : admitted) need += job_footprint(e->spec);
                E.make_room(need, budget);  // cached blocks give way to the budget
                for (auto* e : admitted) {
                    e->t_admit = now_ms();
                    E.alloc(*e);
                }
                tr.step(0);
                Table* T = E.make_table(admitted);
                tr.step(8);
                E.reserve_front(*T);
                audit();
                tr.step(9);
                E.front(*T);
                E.drop_table(T);
                E.release_front();  // admitted jobs reserve their share of it only while they are in flight
                tr.step(1);
                std::vector<EJob*> naive_done;
                for (auto* e : admitted) {
                    if (e->d.algo == VF_FLOOD) {
                        leveling.push_back(e);
                        lev_dirty = true;
                    } else {
                        naive_done.push_back(e);
                    }
                }
                if (!naive_done.empty()) {
                    Table* P = E.make_table(naive_done);
                    audit();
                    E.post(*P);
                    VF_HIP(hipMemcpyAsync(E.pinned, P->gather, 16 * 4 * naive_done.size(), hipMemcpyDeviceToHost, E.s));
                    VF_HIP(hipStreamSynchronize(E.s));
                    std::vector<uint32_t> w(E.pinned, E.pinned + 16 * naive_done.size());
                    for (size_t i = 0; i < naive_done.size(); ++i) E.finish(*naive_done[i], w.data() + 16 * i);
                    extract(naive_done);
                    for (auto* e : naive_done) {
                        E.release(*e);
                        complete(e, e->res.status);
                    }
                    E.drop_table(P);
                    tr.step(2);
                }
            }
            // one chunk of relaxation rounds for every flood job in flight
            if (!leveling.empty()) {
                if (lev_dirty || !lev) {
                    if (lev) E.drop_table(lev);
                    lev = E.make_table(leveling);
                    lev_dirty = false;
                    audit();
                    tr.step(3);
                }
                E.rounds(*lev, kServerRounds);
                E.gather(*lev, E.pinned);
                VF_HIP(hipStreamSynchronize(E.s));
                tr.step(4);
                tr.round_jobs += leveling.size();
                std::vector<uint32_t> w(E.pinned, E.pinned + 16 * leveling.size());
                std::vector<EJob*> drained, still, subs;
                for (size_t i = 0; i < leveling.size(); ++i) {
                    const uint32_t* wi = w.data() + 16 * i;
                    const bool done = wi[2 * kStatFloodDone] || wi[2 * kStatFloodDone + 1] || wi[2 * kStatError];
                    const bool sub = done && !wi[2 * kStatError] && leveling[i]->d.sub && !leveling[i]->sub_done;
                    (sub ? subs : done ? drained : still).push_back(leveling[i]);
                }
                if (!subs.empty()) {  // merged seeds: the sub-seed pass, then they grow again or are done
                    Table* U = E.make_table(subs);
                    audit();
                    E.flood_sub(*U);
                    E.gather(*U, E.pinned);
                    VF_HIP(hipStreamSynchronize(E.s));
                    for (size_t i = 0; i < subs.size(); ++i) {
                        subs[i]->sub_done = true;
                        const uint32_t* wi = E.pinned + 16 * i;
                        const bool done = wi[2 * kStatFloodDone] || wi[2 * kStatFloodDone + 1] || wi[2 * kStatError];
                        (done ? drained : still).push_back(subs[i]);
                    }
                    E.drop_table(U);
                    if (drained.empty()) {  // the restarted jobs rejoin the rounds with a fresh table
                        leveling.swap(still);
                        E.drop_table(lev);
                        lev = nullptr;
                        lev_dirty = true;
                    }
                }
                if (!drained.empty()) {
                    Table* P = E.make_table(drained);
                    audit();
                    tr.step(10);
                    E.post(*P);
                    VF_HIP(hipMemcpyAsync(E.pinned, P->gather, 16 * 4 * drained.size(), hipMemcpyDeviceToHost, E.s));
                    VF_HIP(hipStreamSynchronize(E.s));
                    tr.step(5);
                    std::vector<uint32_t> wd(E.pinned, E.pinned + 16 * drained.size());
                    for (size_t i = 0; i < drained.size(); ++i) E.finish(*drained[i], wd.data() + 16 * i);
                    extract(drained);
                    for (auto* e : drained) {
                        E.release(*e);
                        complete(e, e->res.status);
                    }
                    tr.step(11);
                    E.drop_table(P);
                    leveling.swap(still);
                    // the table (and its worklists) of the jobs just retired goes now: their reservations
                    // were returned above
                    E.drop_table(lev);
                    lev = nullptr;
                    lev_dirty = true;
                    tr.step(11);
                }
            }
        }
        if (lev) E.drop_table(lev);
        VF_HIP(hipStreamSynchronize(E.s));
        tr.print();
    } catch (const Exception& ex) {
        std::lock_guard<std::mutex> lk(mu);
        fatal = vf_last_error();
        for (auto& kv : jobs)
            if (kv.second->phase != 2) {
                kv.second->res.status = ex.code;
                kv.second->phase = 2;
            }
        done_cv.notify_all();
    }
}

extern "C" {

int vf_server_create(size_t hbm_budget_bytes, uint32_t max_slots, vf_server** out)
{
    return guard([&] {
        if (!out) fail(VF_ERR_INVALID_ARG, "out is NULL");
        if (max_slots == 0 || max_slots > Engine::kMaxTableJobs) fail(VF_ERR_INVALID_ARG, "max_slots out of range");
        if (vf_device_count() < 1) fail(VF_ERR_DEVICE, "no HIP device available");
        auto* s = new vf_server();
        s->budget = hbm_budget_bytes;
        s->max_slots = max_slots;
        VF_HIP(hipGetDevice(&s->device));
        s->th = std::thread([s] { s->loop(); });
        *out = s;
    });
}

int vf_server_destroy(vf_server* s)
{
    return guard([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        if (s->th.joinable()) s->th.join();
        for (auto& kv : s->jobs) delete kv.second;
        delete s;
    });
}

int vf_server_submit(vf_server* s, const vf_job* job, uint64_t* ticket)
{
    return guard([&] {
        if (!s || !job || !ticket) fail(VF_ERR_INVALID_ARG, "NULL argument");
        auto* e = new EJob();
        e->spec = *job;
        std::string why;
        const int st = job_validate(e->spec, why);
        if (st == VF_OK && !e->spec.mesh_on_device && e->spec.nt) {
            e->xyz_copy.assign(job->xyz, job->xyz + 3 * (size_t)job->nv);
            e->tri_copy.assign(job->tri, job->tri + 3 * (size_t)job->nt);
            e->spec.xyz = e->xyz_copy.data();
            e->spec.tri = e->tri_copy.data();
        }
        if (st == VF_OK && job->seeds) {  // the caller's fragment seeds, kept with the job
            e->seeds_copy.assign(job->seeds, job->seeds + job->num_seeds);
            e->spec.seeds = e->seeds_copy.data();
        }
        e->spec.labels_out = job->labels_out;
        e->spec.histogram_out = job->histogram_out;
        e->t_submit = now_ms();
        e->res.voxels = (uint64_t)job->dims[0] * job->dims[1] * job->dims[2];
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->fatal.empty()) fail(VF_ERR_DEVICE, "server stopped: %s", s->fatal.c_str());
        e->ticket = s->next_ticket++;
        *ticket = e->ticket;
        s->jobs[e->ticket] = e;
        if (st != VF_OK) {
            e->res.status = st;
            e->phase = 2;
            ++s->completed;
            s->done_cv.notify_all();
            return;
        }
        s->queue.push_back(e);
        s->cv.notify_all();
    });
}

int vf_server_cancel(vf_server* s, uint64_t ticket)
{
    return guard([&] {
        if (!s) fail(VF_ERR_INVALID_ARG, "server is NULL");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(ticket);
        if (it == s->jobs.end()) fail(VF_ERR_INVALID_ARG, "unknown ticket");
        it->second->cancelled = true;
        s->cv.notify_all();
    });
}

static int collect(vf_server* s, uint64_t ticket, vf_result* r, bool block)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) {
        set_error("unknown ticket %llu", (unsigned long long)ticket);
        return VF_ERR_INVALID_ARG;
    }
    EJob* e = it->second;
    if (block) s->done_cv.wait(lk, [&] { return e->phase == 2; });
    if (e->phase != 2) return VF_ERR_NOT_READY;
    if (r) *r = e->res;
    s->jobs.erase(it);
    delete e;
    return VF_OK;
}

int vf_server_poll(vf_server* s, uint64_t ticket, vf_result* result)
{
    if (!s) {
        set_error("server is NULL");
        return VF_ERR_INVALID_ARG;
    }
    return collect(s, ticket, result, false);
}

int vf_server_wait(vf_server* s, uint64_t ticket, vf_result* result)
{
    if (!s) {
        set_error("server is NULL");
        return VF_ERR_INVALID_ARG;
    }
    return collect(s, ticket, result, true);
}

int vf_server_memory(vf_server* s, size_t* reserved_bytes, size_t* device_bytes, size_t* peak_device_bytes,
                     size_t* max_unreserved_bytes)
{
    return guard([&] {
        if (!s) fail(VF_ERR_INVALID_ARG, "server is NULL");
        std::lock_guard<std::mutex> lk(s->mu);
        if (reserved_bytes) *reserved_bytes = s->reserved;
        if (device_bytes) *device_bytes = s->device_bytes;
        if (peak_device_bytes) *peak_device_bytes = s->peak_device_bytes;
        if (max_unreserved_bytes) *max_unreserved_bytes = s->max_unreserved;
    });
}

int vf_server_stats(vf_server* s, size_t* reserved_bytes, uint32_t* in_flight, uint64_t* completed)
{
    return guard([&] {
        if (!s) fail(VF_ERR_INVALID_ARG, "server is NULL");
        std::lock_guard<std::mutex> lk(s->mu);
        if (reserved_bytes) *reserved_bytes = s->reserved;
        if (in_flight) *in_flight = s->in_flight;
        if (completed) *completed = s->completed;
    });
}

}  // extern "C"
