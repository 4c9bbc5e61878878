// ============================================================================
// serve.hip -- request server: continuous batching of force-evaluation and
// integration jobs of different N and expansion order p on one GPU.
//
//  * Jobs are admitted in arrival order while their estimated device
//    footprint fits the budget (afmm_job_bytes); later arrivals wait, and are
//    admitted as running jobs finish and release memory.
//  * One worker per expansion order p owns a solver context (stream, pools,
//    order-p tables) and a device arena for its jobs' resident arrays. Each
//    iteration it takes the jobs admitted since the last one and runs ONE
//    batched step over all of its active jobs: drift, one forest evaluation
//    (one launch per FMM phase for all systems), kick. A new job's first
//    evaluation is its VelocityVerlet3D::initialize (Integrator3D.hpp:19-21);
//    an integration job then takes `steps` steps (:23-40), a force job
//    returns the accelerations of that first evaluation.
//  * Workers of different orders run concurrently on their own streams, so
//    their batches overlap on the device.
//  * A job's result does not depend on its batch mates: every tree, list and
//    sum of the forest is the one of that system alone.
// The reference has no server; its drivers call the solver once per step
// (gpu/galaxy_sim.cu:230-237).
// ============================================================================
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "afmm_ctx.h"
#include "afmm_math.h"

using namespace afmm;

namespace {

using Clock = std::chrono::steady_clock;

// Grow-only device arena with first-fit reuse: job arrays come and go without
// hipFree (which would synchronise the device under the other workers).
class Arena {
public:
    ~Arena() {
        for (auto& c : chunks_) (void)hipFree(c.base);
    }
    char* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto& c : chunks_)
            for (auto it = c.free.begin(); it != c.free.end(); ++it)
                if (it->second >= bytes) {
                    const size_t off = it->first, rem = it->second - bytes;
                    c.free.erase(it);
                    if (rem) c.free[off + bytes] = rem;
                    return c.base + off;
                }
        const size_t sz = std::max(bytes, kChunk);
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) throw HipError(e, "hipMalloc(job arena)");
        chunks_.push_back(Chunk{static_cast<char*>(p), sz, {}});
        if (sz > bytes) chunks_.back().free[bytes] = sz - bytes;
        return static_cast<char*>(p);
    }
    void release(char* p, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto& c : chunks_) {
            if (p < c.base || p >= c.base + c.size) continue;
            size_t off = (size_t)(p - c.base);
            auto nx = c.free.lower_bound(off);
            if (nx != c.free.end() && nx->first == off + bytes) {  // merge with the next block
                bytes += nx->second;
                nx = c.free.erase(nx);
            }
            if (nx != c.free.begin()) {  // merge with the previous block
                auto pv = std::prev(nx);
                if (pv->first + pv->second == off) {
                    off = pv->first;
                    bytes += pv->second;
                    c.free.erase(pv);
                }
            }
            c.free[off] = bytes;
            return;
        }
    }

private:
    static constexpr size_t kChunk = size_t(1) << 30;
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> free;  // offset -> bytes
    };
    std::vector<Chunk> chunks_;
};

struct Job {
    int64_t id = 0;
    afmm_job_desc d{};
    std::vector<double> pos, vel, mass, acc;  // inputs, then results
    int64_t bytes = 0;
    int state = AFMM_JOB_QUEUED;
    int error = 0;
    std::string msg;
    bool init = false;
    std::atomic<int> steps_done{0}, batches{0};
    double t_submit = 0, t_admit = 0, t_done = 0;
    // device (arena block)
    char* block = nullptr;
    size_t block_bytes = 0;
    double *dpos = nullptr, *dvel = nullptr, *dacc = nullptr, *dmass = nullptr;
};

struct Worker {
    int p = 0;
    afmm_ctx* ctx = nullptr;
    std::thread th;
    std::deque<std::shared_ptr<Job>> inbox;  // admitted, not yet started (guarded by the server mutex)
    std::vector<std::shared_ptr<Job>> active;
    Arena arena;
};

}  // namespace

struct afmm_server {
    afmm_server_params prm{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<std::shared_ptr<Job>> pending;  // arrival order
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    std::map<int, std::unique_ptr<Worker>> workers;  // by order p
    int64_t next_id = 1, hbm_used = 0, unfinished = 0;
    bool stop = false;
    std::string err;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void worker_loop(afmm_server* sv, Worker* w);

// Footprint estimate at leaf capacity cap: resident arrays 80 B/body; batch
// copies, keys, sort and sorted bodies ~420 B/body; per box the two
// expansions (2 S doubles) and the lists (~1.3 KB). Boxes number about
// 4 n / cap + 64 (n / 16 at capacity 64; measured n / 22 on the merger).
int64_t job_bytes_cap(int64_t n, int32_t p, int32_t cap) {
    const int64_t S = (int64_t)(p + 1) * (p + 1);
    const int64_t boxes = 4 * n / std::max<int32_t>(cap, 1) + 64;
    return n * 500 + boxes * (16 * S + 1300);
}

// The worker of order p, created with its context on first use (caller holds
// sv->mu); null with *rc set when the context cannot be created.
Worker* worker_for(afmm_server* sv, int p, int* rc) {
    auto it = sv->workers.find(p);
    if (it != sv->workers.end()) return it->second.get();
    auto w = std::make_unique<Worker>();
    w->p = p;
    afmm_params prm{1.0, 0.0, p, sv->prm.leaf_cap, sv->prm.precision, sv->prm.device};
    *rc = afmm_create(&prm, &w->ctx);
    if (*rc) return nullptr;
    w->ctx->split_p2p = false;  // one stream per worker: the workers share the GPU
    Worker* raw = w.get();
    sv->workers.emplace(p, std::move(w));
    raw->th = std::thread(worker_loop, sv, raw);
    return raw;
}

// Moves pending jobs, in arrival order, to their order's worker while their
// footprint fits the budget (caller holds sv->mu).
void admit_locked(afmm_server* sv) {
    while (!sv->pending.empty()) {
        auto j = sv->pending.front();
        if (sv->hbm_used > 0 && sv->hbm_used + j->bytes > sv->prm.hbm_budget) break;
        int rc = AFMM_OK;
        Worker* w = worker_for(sv, j->d.p, &rc);
        sv->pending.pop_front();
        if (!w) {
            j->state = AFMM_JOB_FAILED;
            j->error = rc;
            j->msg = afmm_strerror(rc);
            j->t_done = sv->now();
            --sv->unfinished;
            sv->done_cv.notify_all();
            continue;
        }
        sv->hbm_used += j->bytes;
        j->state = AFMM_JOB_RUNNING;
        j->t_admit = sv->now();
        w->inbox.push_back(j);
    }
    sv->work_cv.notify_all();
}

void finish_locked(afmm_server* sv, Worker* w, const std::shared_ptr<Job>& j, int rc, const std::string& msg) {
    if (j->block) w->arena.release(j->block, j->block_bytes);
    j->block = nullptr;
    j->state = rc ? AFMM_JOB_FAILED : AFMM_JOB_DONE;
    j->error = rc;
    j->msg = msg;
    j->t_done = sv->now();
    sv->hbm_used -= j->bytes;
    --sv->unfinished;
    sv->done_cv.notify_all();
    admit_locked(sv);
}

// One batched step of the worker's active jobs; throws on HIP errors.
int batch_step(Worker* w) {
    afmm_ctx* c = w->ctx;
    hipStream_t s = c->stream;
    const int nj = (int)w->active.size();
    std::vector<int> boff(nj + 1, 0);
    std::vector<double> G(nj), eps(nj);
    std::vector<JobRef> refs(nj);
    for (int q = 0; q < nj; ++q) {
        Job& j = *w->active[q];
        boff[q + 1] = boff[q] + (int)j.d.n;
        G[q] = j.d.G;
        eps[q] = j.d.eps;
        refs[q] = JobRef{j.dpos, j.dvel, j.dacc, nullptr, j.dmass, j.d.dt, j.init ? 1 : 0, j.init ? 1 : 0};
    }
    const int n = boff[nj];
    JobRef* d_refs = c->pool.get<JobRef>("srv_jobs", nj);
    int* d_boff = c->pool.get<int>("srv_boff", nj + 1);
    double* bpos = c->pool.get<double>("srv_pos", 3 * (size_t)n);
    double* bmass = c->pool.get<double>("srv_mass", n);
    double* bacc = c->pool.get<double>("srv_acc", 3 * (size_t)n);
    AFMM_CHECK(hipMemcpyAsync(d_refs, refs.data(), sizeof(JobRef) * nj, hipMemcpyHostToDevice, s));
    AFMM_CHECK(hipMemcpyAsync(d_boff, boff.data(), sizeof(int) * (nj + 1), hipMemcpyHostToDevice, s));
    leapfrog_gather(s, n, nj, d_boff, d_refs, bpos, bmass);
    const int rc = eval_forest(c, nj, boff.data(), G.data(), eps.data(), bpos, bmass, bacc, nullptr);
    if (rc) return rc;
    leapfrog_kick(s, n, nj, d_boff, d_refs, bacc, nullptr);
    AFMM_CHECK(hipStreamSynchronize(s));
    return AFMM_OK;
}

void worker_loop(afmm_server* sv, Worker* w) {
    (void)hipSetDevice(sv->prm.device);
    for (;;) {
        std::vector<std::shared_ptr<Job>> fresh;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->work_cv.wait(lk, [&] { return sv->stop || !w->inbox.empty() || !w->active.empty(); });
            if (sv->stop && w->active.empty() && w->inbox.empty()) return;
            // bound a batch by max_batch jobs and by 2^30 bodies
            int64_t bodies = 0;
            for (auto& j : w->active) bodies += j->d.n;
            while (!w->inbox.empty() &&
                   (sv->prm.max_batch <= 0 || (int)(w->active.size() + fresh.size()) < sv->prm.max_batch) &&
                   bodies + w->inbox.front()->d.n < (int64_t(1) << 30)) {
                bodies += w->inbox.front()->d.n;
                fresh.push_back(w->inbox.front());
                w->inbox.pop_front();
            }
        }
        // upload the newcomers into the worker's arena
        for (auto& j : fresh) {
            const int rc = guarded(w->ctx, [&]() -> int {
                const size_t n = (size_t)j->d.n;
                j->block_bytes = sizeof(double) * 10 * n;
                j->block = w->arena.alloc(j->block_bytes);
                j->dpos = reinterpret_cast<double*>(j->block);
                j->dvel = j->dpos + 3 * n;
                j->dacc = j->dvel + 3 * n;
                j->dmass = j->dacc + 3 * n;
                hipStream_t s = w->ctx->stream;
                AFMM_CHECK(hipMemcpyAsync(j->dpos, j->pos.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
                AFMM_CHECK(hipMemcpyAsync(j->dvel, j->vel.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
                AFMM_CHECK(hipMemcpyAsync(j->dmass, j->mass.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
                AFMM_CHECK(hipStreamSynchronize(s));
                return AFMM_OK;
            });
            if (rc) {
                std::lock_guard<std::mutex> lk(sv->mu);
                finish_locked(sv, w, j, rc, afmm_last_error(w->ctx));
                continue;
            }
            w->active.push_back(j);
        }
        if (w->active.empty()) continue;
        const int rc = guarded(w->ctx, [&]() -> int { return batch_step(w); });
        // advance, collect finished jobs
        std::vector<std::shared_ptr<Job>> keep, done;
        for (auto& j : w->active) {
            ++j->batches;
            if (rc) {
                done.push_back(j);
                continue;
            }
            if (!j->init) j->init = true;
            else ++j->steps_done;
            if (j->steps_done >= j->d.steps) done.push_back(j);
            else keep.push_back(j);
        }
        int drc = AFMM_OK;
        if (!rc && !done.empty())
            drc = guarded(w->ctx, [&]() -> int {
                hipStream_t s = w->ctx->stream;
                for (auto& j : done) {
                    const size_t n = (size_t)j->d.n;
                    j->acc.resize(3 * n);
                    AFMM_CHECK(hipMemcpyAsync(j->pos.data(), j->dpos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
                    AFMM_CHECK(hipMemcpyAsync(j->vel.data(), j->dvel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
                    AFMM_CHECK(hipMemcpyAsync(j->acc.data(), j->dacc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
                }
                AFMM_CHECK(hipStreamSynchronize(s));
                return AFMM_OK;
            });
        const int frc = rc ? rc : drc;
        const std::string msg = frc ? afmm_last_error(w->ctx) : "";
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            for (auto& j : done) finish_locked(sv, w, j, frc, msg);
        }
        w->active.swap(keep);
    }
}

}  // namespace

extern "C" {

int64_t afmm_job_bytes(int64_t n, int32_t p) { return job_bytes_cap(n, p, 64); }

int afmm_server_create(const afmm_server_params* prm, afmm_server** out) {
    if (!prm || !out) return AFMM_ERR_ARG;
    *out = nullptr;
    if (prm->leaf_cap < 1) return AFMM_ERR_LEAF_CAP;
    if (prm->precision < AFMM_PRECISION_FP64 || prm->precision > AFMM_PRECISION_P2P_FP32)
        return AFMM_ERR_PRECISION;
    afmm_server* sv = new (std::nothrow) afmm_server();
    if (!sv) return AFMM_ERR_NO_MEMORY;
    sv->prm = *prm;
    if (sv->prm.hbm_budget <= 0) sv->prm.hbm_budget = int64_t(240) << 30;
    sv->t0 = Clock::now();
    *out = sv;
    return AFMM_OK;
}

int afmm_server_submit(afmm_server* sv, const afmm_job_desc* d, int64_t* id) {
    if (!sv || !d || !id || d->n < 1 || !d->pos_xyz || !d->mass || d->steps < 0) return AFMM_ERR_ARG;
    if (d->steps > 0 && !d->vel_xyz) return AFMM_ERR_ARG;
    if (d->p < 1 || d->p > kMaxOrder) return AFMM_ERR_ORDER;
    if (!std::isfinite(d->G) || !std::isfinite(d->eps) || d->eps < 0 || !std::isfinite(d->dt)) return AFMM_ERR_ARG;
    if (d->n >= (int64_t(1) << 30)) return AFMM_ERR_TOO_LARGE;
    auto j = std::make_shared<Job>();
    j->d = *d;
    const size_t n = (size_t)d->n;
    j->pos.assign(d->pos_xyz, d->pos_xyz + 3 * n);
    j->vel.assign(3 * n, 0.0);
    if (d->vel_xyz) j->vel.assign(d->vel_xyz, d->vel_xyz + 3 * n);
    j->mass.assign(d->mass, d->mass + n);
    j->d.pos_xyz = j->d.vel_xyz = j->d.mass = nullptr;
    j->bytes = job_bytes_cap(d->n, d->p, sv->prm.leaf_cap);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return AFMM_ERR_STATE;
    j->id = sv->next_id++;
    j->t_submit = sv->now();
    sv->jobs[j->id] = j;
    sv->pending.push_back(j);
    ++sv->unfinished;
    admit_locked(sv);
    *id = j->id;
    return AFMM_OK;
}

int afmm_server_warm(afmm_server* sv, int32_t p) {
    if (!sv) return AFMM_ERR_ARG;
    if (p < 1 || p > kMaxOrder) return AFMM_ERR_ORDER;
    std::lock_guard<std::mutex> lk(sv->mu);
    int rc = AFMM_OK;
    worker_for(sv, p, &rc);
    return rc;
}

int afmm_server_wait(afmm_server* sv, int64_t id) {
    if (!sv) return AFMM_ERR_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return AFMM_ERR_ARG;
    auto j = it->second;
    sv->done_cv.wait(lk, [&] { return j->state == AFMM_JOB_DONE || j->state == AFMM_JOB_FAILED; });
    return j->error;
}

int afmm_server_drain(afmm_server* sv) {
    if (!sv) return AFMM_ERR_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->done_cv.wait(lk, [&] { return sv->unfinished == 0; });
    return AFMM_OK;
}

int afmm_server_info(afmm_server* sv, int64_t id, afmm_job_info* info) {
    if (!sv || !info) return AFMM_ERR_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return AFMM_ERR_ARG;
    const Job& j = *it->second;
    info->state = j.state;
    info->error = j.error;
    info->hbm_bytes = j.bytes;
    info->steps_done = j.steps_done;
    info->batches = j.batches;
    info->t_submit = j.t_submit;
    info->t_admit = j.t_admit;
    info->t_done = j.t_done;
    return AFMM_OK;
}

int afmm_server_result(afmm_server* sv, int64_t id, double* pos, double* vel, double* acc) {
    if (!sv) return AFMM_ERR_ARG;
    std::shared_ptr<Job> j;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end()) return AFMM_ERR_ARG;
        j = it->second;
        if (j->state == AFMM_JOB_FAILED) return j->error;
        if (j->state != AFMM_JOB_DONE) return AFMM_ERR_STATE;
        sv->jobs.erase(it);
    }
    const size_t n = (size_t)j->d.n;
    if (pos) std::copy(j->pos.begin(), j->pos.begin() + 3 * n, pos);
    if (vel) std::copy(j->vel.begin(), j->vel.begin() + 3 * n, vel);
    if (acc) std::copy(j->acc.begin(), j->acc.begin() + 3 * n, acc);
    return AFMM_OK;
}

int64_t afmm_server_hbm_in_use(afmm_server* sv) {
    if (!sv) return -1;
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->hbm_used;
}

void afmm_server_destroy(afmm_server* sv) {
    if (!sv) return;
    {
        // queued jobs fail; each is counted finished and its waiters are woken
        // (a wait or drain must still return before the server is deleted:
        // callers must not let them race destroy)
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        auto fail_queued = [&](const std::shared_ptr<Job>& j, bool admitted) {
            j->state = AFMM_JOB_FAILED;
            j->error = AFMM_ERR_STATE;
            j->msg = "server destroyed before the job started";
            j->t_done = sv->now();
            if (admitted) sv->hbm_used -= j->bytes;
            --sv->unfinished;
        };
        for (auto& j : sv->pending) fail_queued(j, false);
        sv->pending.clear();
        for (auto& kv : sv->workers) {
            for (auto& j : kv.second->inbox) fail_queued(j, true);
            kv.second->inbox.clear();
        }
    }
    sv->done_cv.notify_all();
    sv->work_cv.notify_all();
    for (auto& kv : sv->workers)
        if (kv.second->th.joinable()) kv.second->th.join();
    for (auto& kv : sv->workers) afmm_destroy(kv.second->ctx);
    sv->workers.clear();
    delete sv;
}

}  // extern "C"
