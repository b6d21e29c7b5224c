// Serving runtime: a request loop over the batched step path.
//
// Jobs (one system and a step count each) are queued on submission.  A
// scheduler thread admits them, in arrival order, into batched contexts
// ("cohorts") while the estimated device bytes of everything resident stay
// under the budget and fewer than max_cohorts cohorts are running.  Each
// cohort has its own host thread and HIP streams, so cohorts admitted at
// different times overlap on the device, and every kernel launch of a cohort
// covers all of its jobs.  A cohort steps in chunks that end where one of its
// jobs reaches its step count; that job's results (observables, window sums,
// mean pressure, RDF, optionally the state) are copied out and the job is
// retired, so the other jobs continue bit-for-bit as if it were still there.
// When the last job of a cohort retires the context is destroyed and its
// bytes return to the budget.
#include <algorithm>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ibx.h"
#include "ibx_internal.h"

using namespace ibx;

namespace {

using Clock = std::chrono::steady_clock;

thread_local std::string g_server_err;

int serr(int code, const std::string &msg) {
    g_server_err = msg;
    return code;
}

struct Job {
    int64_t id = 0;
    HostSystem hs;              // moved into the cohort's context at admission
    int natoms = 0, nitems = 0, nsteps = 0, keep_state = 0, nupdate = 1;
    bool hybrid = false;
    size_t bytes = 0;           // admission estimate
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    int status = IBX_OK;
    std::string err;
    double obs[32] = {}, sums[32] = {};
    double mean_pressure = 0;
    std::vector<double> rdf, s, v;
    int cohort = -1, cohort_jobs = 0;
};

using JobPtr = std::shared_ptr<Job>;

}  // namespace

struct ibx_server {
    ibx_server_config cfg{};
    size_t budget = 0;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_sched;   // new job, freed bytes or cohort slot, shutdown
    std::condition_variable cv_done;    // a job finished
    std::deque<JobPtr> queue;
    std::map<int64_t, JobPtr> jobs;     // submitted and not yet collected by ibx_server_wait
    int64_t next_id = 1;
    size_t used = 0, peak_used = 0;
    int running = 0, peak_running = 0, ncohorts = 0, max_cohort = 0;
    int64_t nsubmitted = 0, ncompleted = 0, nfailed = 0, cohort_jobs_total = 0;
    double atom_steps = 0;
    bool stop = false;
    std::thread sched;
    std::map<int, std::thread> workers;
    std::vector<int> finished;          // cohorts whose thread has ended and is not yet joined

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void finish_job(ibx_server *S, Job &j, int status, const std::string &err) {
    std::lock_guard<std::mutex> lk(S->mu);
    j.status = status;
    j.err = err;
    j.t_done = S->now();
    j.done = true;
    if (status == IBX_OK) {
        ++S->ncompleted;
        S->atom_steps += (double)j.natoms * j.nsteps;
    } else {
        ++S->nfailed;
    }
    S->cv_done.notify_all();
}

// Results of job k of the cohort after its last step, then retirement.
int collect(ibx_server *S, ibx_ctx *ctx, int k, Job &j) {
    int rc = ibx_get_observables(ctx, k, j.obs);
    if (!rc) rc = ibx_get_averages(ctx, k, j.sums, 0);
    if (!rc && S->cfg.rdf_nbin > 0) {
        j.rdf.assign(S->cfg.rdf_nbin, 0.0);
        rc = ibx_rdf_get(ctx, k, j.rdf.data(), 0);
    }
    if (!rc && j.keep_state) {
        j.s.assign(3 * (size_t)j.nitems, 0.0);
        j.v.assign(3 * (size_t)j.natoms, 0.0);
        rc = ibx_get_state(ctx, k, j.s.data(), j.v.data());
    }
    if (!rc) rc = ibx_retire(ctx, k);
    // AVERAGE.f90:15: pressure = (PT11 + PT22 + PT33) / 3, averaged over the job's steps
    j.mean_pressure = j.sums[28] > 0 ? (j.sums[15] + j.sums[16] + j.sums[17]) / (3.0 * j.sums[28]) : 0.0;
    return rc;
}

void run_cohort(ibx_server *S, std::vector<JobPtr> jobs, size_t bytes, int cid) {
    const int n = (int)jobs.size();
    std::vector<HostSystem> hs;
    hs.reserve(n);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        const double ta = S->now();
        for (auto &j : jobs) {
            j->t_admit = ta;
            j->cohort = cid;
            j->cohort_jobs = n;
        }
    }
    for (auto &j : jobs) hs.push_back(std::move(j->hs));
    ibx_ctx *ctx = nullptr;
    std::string err;
    int rc = create_context(std::move(hs), &ctx);
    if (rc) err = ibx_last_error(nullptr);
    if (!rc && S->cfg.rdf_nbin > 0) {
        rc = ibx_rdf_enable(ctx, S->cfg.rdf_nbin, S->cfg.rdf_dr, 1);
        if (rc) err = ibx_last_error(ctx);
    }
    std::vector<int> live(n);
    std::iota(live.begin(), live.end(), 0);
    int done = 0;
    while (!rc && !live.empty()) {
        int target = INT_MAX;
        for (int k : live) target = std::min(target, jobs[k]->nsteps);
        const int nstep = std::min(target - done, S->cfg.chunk_steps);
        if (nstep > 0) {
            rc = ibx_step_graph(ctx, nstep, done);
            if (rc) {
                err = ibx_last_error(ctx);
                break;
            }
            done += nstep;
        }
        std::vector<int> keep;
        for (int k : live) {
            if (jobs[k]->nsteps > done) {
                keep.push_back(k);
                continue;
            }
            int r = collect(S, ctx, k, *jobs[k]);
            finish_job(S, *jobs[k], r, r ? ibx_last_error(ctx) : "");
        }
        live.swap(keep);
    }
    for (int k : live) finish_job(S, *jobs[k], rc ? rc : IBX_E_DEVICE, err);
    if (ctx) ibx_destroy(ctx);
    std::lock_guard<std::mutex> lk(S->mu);
    S->used -= bytes;
    --S->running;
    S->finished.push_back(cid);
    S->cv_sched.notify_all();
}

// Takes queued jobs, in arrival order, that fit the free bytes: jobs whose
// update interval differs from the first one's (or, with a step band, whose
// step count is too far from it) wait for a later cohort, and a hybrid system
// runs alone.  Called with the lock held.
std::vector<JobPtr> take_cohort(ibx_server *S, size_t &bytes) {
    std::vector<JobPtr> take;
    bytes = 0;
    const Job &head = *S->queue.front();
    for (auto it = S->queue.begin(); it != S->queue.end() && (int)take.size() < S->cfg.max_cohort_jobs;) {
        const Job &j = **it;
        const double band = S->cfg.step_band;
        const bool fits_steps = band <= 1.0 || (j.nsteps <= band * head.nsteps && head.nsteps <= band * j.nsteps);
        const bool fits_kind =
            j.nupdate == head.nupdate && fits_steps && (take.empty() || (!j.hybrid && !head.hybrid));
        if (!fits_kind) {
            ++it;
            continue;
        }
        if (S->used + bytes + j.bytes > S->budget) break;
        bytes += j.bytes;
        take.push_back(*it);
        it = S->queue.erase(it);
        if (head.hybrid) break;
    }
    return take;
}

void scheduler(ibx_server *S) {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (!S->finished.empty()) {
            std::vector<std::thread> ended;
            for (int cid : S->finished) {
                ended.push_back(std::move(S->workers[cid]));
                S->workers.erase(cid);
            }
            S->finished.clear();
            lk.unlock();
            for (auto &t : ended) t.join();
            lk.lock();
            continue;
        }
        if (S->stop && S->queue.empty() && S->running == 0) break;
        bool formed = false;
        if (!S->queue.empty() && S->running < S->cfg.max_cohorts) {
            JobPtr head = S->queue.front();
            if (head->bytes > S->budget) {
                // can never fit: fail it rather than block the queue
                S->queue.pop_front();
                head->status = IBX_E_NOMEM;
                head->err = "job needs more device bytes than the server budget";
                head->t_done = S->now();
                head->done = true;
                ++S->nfailed;
                S->cv_done.notify_all();
                continue;
            }
            size_t bytes = 0;
            std::vector<JobPtr> take = take_cohort(S, bytes);
            if (!take.empty()) {
                S->used += bytes;
                S->peak_used = std::max(S->peak_used, S->used);
                ++S->running;
                S->peak_running = std::max(S->peak_running, S->running);
                const int cid = S->ncohorts++;
                S->max_cohort = std::max(S->max_cohort, (int)take.size());
                S->cohort_jobs_total += (int64_t)take.size();
                S->workers[cid] = std::thread(run_cohort, S, std::move(take), bytes, cid);
                formed = true;
            }
        }
        if (!formed) S->cv_sched.wait(lk);
    }
}

}  // namespace

extern "C" {

const char *ibx_server_last_error(void) { return g_server_err.c_str(); }

int ibx_estimate_bytes(const ibx_system_desc *desc, int rdf_nbin, int64_t *bytes) {
    if (!desc || !bytes || rdf_nbin < 0) return serr(IBX_E_INVALID, "invalid arguments");
    HostSystem hs;
    std::string err;
    int rc = build_host_system(desc, hs, err);
    if (rc) return serr(rc, err);
    *bytes = (int64_t)estimate_device_bytes(hs, rdf_nbin);
    return IBX_OK;
}

int ibx_server_create(const ibx_server_config *cfg, ibx_server **out) {
    if (!out) return serr(IBX_E_INVALID, "invalid arguments");
    *out = nullptr;
    std::unique_ptr<ibx_server> S(new ibx_server());
    if (cfg) S->cfg = *cfg;
    ibx_server_config &c = S->cfg;
    if (c.hbm_budget <= 0) c.hbm_budget = (int64_t)200e9;
    if (c.max_cohorts <= 0) c.max_cohorts = 3;
    if (c.max_cohort_jobs <= 0) c.max_cohort_jobs = 256;
    if (c.chunk_steps <= 0) c.chunk_steps = 200;
    if (c.rdf_nbin < 0 || c.rdf_nbin > 8192 || (c.rdf_nbin > 0 && !(c.rdf_dr > 0)))
        return serr(IBX_E_INVALID, "invalid RDF settings");
    size_t free_b = 0;
    int rc = device_free_bytes(&free_b);
    if (rc) return serr(rc, "no usable HIP device");
    // never promise more than the device has free now (less a margin for the runtime)
    S->budget = std::min((size_t)c.hbm_budget, (size_t)(0.9 * (double)free_b));
    S->sched = std::thread(scheduler, S.get());
    *out = S.release();
    return IBX_OK;
}

int ibx_server_submit(ibx_server *S, const ibx_system_desc *desc, int nsteps, int keep_state, int64_t *job) {
    if (!S || !desc || !job || nsteps < 1) return serr(IBX_E_INVALID, "invalid arguments");
    auto j = std::make_shared<Job>();
    std::string err;
    int rc = build_host_system(desc, j->hs, err);
    if (rc) return serr(rc, err);
    j->natoms = j->hs.natoms;
    j->nitems = j->hs.nitems;
    j->nsteps = nsteps;
    j->keep_state = keep_state != 0;
    j->nupdate = j->hs.nupdate;
    j->hybrid = j->hs.hybrid;
    // served jobs run thousands of steps unattended: room for local density
    // twice the mean before a list would overflow
    j->hs.list_margin = 2.0;
    j->bytes = estimate_device_bytes(j->hs, S->cfg.rdf_nbin);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return serr(IBX_E_INVALID, "server is shutting down");
    j->id = S->next_id++;
    j->t_submit = S->now();
    S->queue.push_back(j);
    S->jobs[j->id] = j;
    ++S->nsubmitted;
    *job = j->id;
    S->cv_sched.notify_all();
    return IBX_OK;
}

int ibx_server_wait(ibx_server *S, int64_t job, double timeout_s, ibx_job_result *res, double *rdf, double *s,
                    double *v) {
    if (!S) return serr(IBX_E_INVALID, "invalid arguments");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return serr(IBX_E_INVALID, "unknown or already collected job");
    JobPtr j = it->second;
    auto ready = [&] { return j->done; };
    if (timeout_s < 0) {
        S->cv_done.wait(lk, ready);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return serr(IBX_E_PENDING, "job still running");
    }
    S->jobs.erase(it);
    lk.unlock();
    if (res) {
        std::memset(res, 0, sizeof *res);
        res->status = j->status;
        res->natoms = j->natoms;
        res->nsteps = j->nsteps;
        res->cohort = j->cohort;
        res->cohort_jobs = j->cohort_jobs;
        res->t_submit = j->t_submit;
        res->t_admit = j->t_admit;
        res->t_done = j->t_done;
        res->mean_pressure = j->mean_pressure;
        std::memcpy(res->obs, j->obs, sizeof j->obs);
        std::memcpy(res->sums, j->sums, sizeof j->sums);
        std::strncpy(res->error, j->err.c_str(), sizeof res->error - 1);
    }
    if (rdf && !j->rdf.empty()) std::memcpy(rdf, j->rdf.data(), j->rdf.size() * sizeof(double));
    if (s && !j->s.empty()) std::memcpy(s, j->s.data(), j->s.size() * sizeof(double));
    if (v && !j->v.empty()) std::memcpy(v, j->v.data(), j->v.size() * sizeof(double));
    if (j->status != IBX_OK) return serr(j->status, j->err);
    return IBX_OK;
}

int ibx_server_stats(ibx_server *S, double st[16]) {
    if (!S || !st) return serr(IBX_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    for (int k = 0; k < 16; ++k) st[k] = 0.0;
    st[0] = (double)S->nsubmitted;
    st[1] = (double)S->ncompleted;
    st[2] = (double)S->nfailed;
    st[3] = (double)S->queue.size();
    st[4] = S->running;
    st[5] = S->ncohorts;
    st[6] = S->peak_running;
    st[7] = (double)S->used;
    st[8] = (double)S->peak_used;
    st[9] = S->atom_steps;
    st[10] = S->now();
    st[11] = S->max_cohort;
    st[12] = S->ncohorts ? (double)S->cohort_jobs_total / S->ncohorts : 0.0;
    st[13] = (double)S->budget;
    return IBX_OK;
}

void ibx_server_destroy(ibx_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_sched.notify_all();
    }
    if (S->sched.joinable()) S->sched.join();
    for (auto &w : S->workers)
        if (w.second.joinable()) w.second.join();
    delete S;
}

}  // extern "C"
