// Serving runtime: a request loop over the batched step path, with continuous
// batching.
//
// Jobs (one system and a step count each) are queued on submission.  A
// scheduler thread gives each queued job, in arrival order, the smallest free
// place of a running batch ("cohort") that holds it; the cohort's thread hands
// the place over (ibx_assign) at its next graph-block boundary, so the job's
// own steps are in phase with the batch's list rebuilds.  A job that fits no
// free place starts a new cohort, with other queued jobs, while the estimated
// device bytes of all cohorts stay under the budget and fewer than max_cohorts
// cohorts run.  Every place is reserved for a size class (1024 items times a
// power of class_ratio; the other sizes scaled from its job), and a new cohort
// also holds spare places, per size class seen so far, for the jobs that
// arrive while it runs.  Each cohort has its own host thread and HIP stream,
// so cohorts overlap on the device, and every kernel launch of a cohort covers
// all of its places.  Every job stops by itself after its last step
// (ibx_set_last_step), so a cohort steps in whole graph blocks, in chunks that
// end at the block boundary after a job's last step or where a hand-over is
// due.  A finished job's results (observables, window sums, mean pressure, RDF,
// optionally the state) are then copied out, its place is retired and free
// again, and the other jobs continue bit for bit.  A cohort with no job left ends and its bytes return
// to the budget.  When the oldest queued job has waited for a place that no
// cohort has free and no new cohort can start, the cohort with the least work
// left stops taking jobs, so that it ends and makes room.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
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

constexpr int kClassBase = 1024;          // items of the smallest size class
constexpr double kDrainAfterS = 0.25;     // wait of the oldest queued job before a cohort is drained

struct Job {
    int64_t id = 0;
    HostSystem hs;              // moved into a context when the job starts
    SlotSizes need;             // what it needs of a place
    int list_need = 0;          // list capacity its density asks for
    int natoms = 0, nitems = 0, nsteps = 0, keep_state = 0, nupdate = 1, cls = 0;
    bool hybrid = false;
    bool movable = false;       // may take over a place (no virtual sites, one nonbonded group)
    size_t bytes = 0;           // admission estimate (its own sizes)
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    int status = IBX_OK;
    std::string err;
    double obs[32] = {}, sums[32] = {};
    double mean_pressure = 0;
    std::vector<double> rdf, s, v;
    int cohort = -1, cohort_jobs = 0;
    int place = -1, step0 = 0;  // place in its cohort and the cohort step its own step 0 was
};

using JobPtr = std::shared_ptr<Job>;
using SysPtr = std::shared_ptr<const HostSystem>;

struct Place {
    JobPtr job;                 // running here
    JobPtr pending;             // given this place: starts at the cohort's next block boundary
    SlotSizes cap;              // what the place holds
    bool dead = false;          // a hand-over failed here: not used again
};

struct Cohort {
    int id = 0;
    size_t bytes = 0;
    int nupdate = 1, period = 1, list_cap = 0;
    bool continuous = true;     // free places are given to queued jobs
    bool ready = false;         // context created, capacities known
    bool draining = false;      // takes no more jobs
    bool closing = false;       // its thread is ending
    int done = 0;               // steps run
    std::vector<Place> places;
};

using CohortPtr = std::shared_ptr<Cohort>;

int class_items(int cls, double ratio) {
    double c = kClassBase;
    for (int k = 0; k < cls; ++k) c = std::ceil(c * ratio);
    return (int)c;
}

int size_class(int nitems, double ratio) {
    int cls = 0;
    while (class_items(cls, ratio) < nitems) ++cls;
    return cls;
}

// Reservation of a place of size class cls for systems like h: items of the
// class, the other sizes scaled from h's with room for systems a little
// denser in terms, exclusions and cells than h; h's tables.
SlotSizes class_reserve(const HostSystem &h, int cls, double ratio) {
    const SlotSizes n = slot_needs(h);
    const int C = class_items(cls, ratio);
    const double f = (double)C / std::max(1, h.nitems);
    auto up = [&](int x, double extra, int slack) { return x > 0 ? (int)std::ceil(x * f * extra) + slack : 0; };
    SlotSizes r;
    r.items = C;
    r.terms = up(n.terms, 1.1, 64);
    r.contrib = up(n.contrib, 1.1, 256);
    r.excl = up(n.excl, 1.1, 256);
    for (int g = 0; g < 2; ++g) {
        r.members[g] = n.members[g] > 0 ? std::min(C, up(n.members[g], 1.1, 1)) : 0;
        r.cells[g] = up(n.cells[g], 1.3, 64);
    }
    r.nb_rows = n.nb_rows;
    r.nb_tabs = n.nb_tabs;
    r.nb_map = n.nb_map;
    r.bd_tabs = n.bd_tabs;
    r.bd_rows = n.bd_rows;
    return r;
}

}  // namespace

struct ibx_server {
    ibx_server_config cfg{};
    size_t budget = 0;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_sched;   // new job, freed place, bytes or cohort, shutdown
    std::condition_variable cv_done;    // a job finished
    std::deque<JobPtr> queue;
    std::map<int64_t, JobPtr> jobs;     // submitted and not yet collected by ibx_server_wait
    std::map<int, CohortPtr> cohorts;   // running
    std::map<int, SysPtr> templates;    // one system per size class seen (spare places are built from it)
    int64_t next_id = 1;
    size_t used = 0, peak_used = 0;
    int running = 0, peak_running = 0, ncohorts = 0, max_places = 0;
    int64_t nsubmitted = 0, ncompleted = 0, nfailed = 0, cohort_jobs_total = 0, nadmit_running = 0;
    double atom_steps = 0, atom_steps_run = 0;
    // host seconds of the cohort threads: stepping, hand-overs, result copies, context creation; chunks,
    // hand-overs, steps (summed over cohorts)
    double t_step = 0, t_assign = 0, t_collect = 0, t_create = 0;
    int64_t nchunks = 0, nassign = 0, nsteps_run = 0;
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

// Results of the job at place k after its last step, then retirement of the place.
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

int live_jobs(const Cohort &C) {
    int n = 0;
    for (const Place &p : C.places) n += (p.job ? 1 : 0) + (p.pending ? 1 : 0);
    return n;
}

// The thread of one cohort: jobs (places 0..n-1) and spare places (built from
// spare systems, retired at once) in one context, stepped until no job is left.
void run_cohort(ibx_server *S, CohortPtr C, std::vector<JobPtr> jobs, std::vector<HostSystem> spares) {
    const int n = (int)jobs.size();
    std::vector<HostSystem> hs;
    hs.reserve(n + spares.size());
    {
        std::lock_guard<std::mutex> lk(S->mu);
        const double ta = S->now();
        for (int k = 0; k < n; ++k) {
            jobs[k]->t_admit = ta;
            jobs[k]->cohort = C->id;
            jobs[k]->cohort_jobs = n;
            jobs[k]->place = k;
            jobs[k]->step0 = 0;
            C->places[k].job = jobs[k];
        }
    }
    for (auto &j : jobs) hs.push_back(std::move(j->hs));
    for (auto &h : spares) hs.push_back(std::move(h));
    const int nplace = (int)hs.size();
    ibx_ctx *ctx = nullptr;
    std::string err;
    const double tc0 = S->now();
    int rc = create_context(std::move(hs), &ctx);
    if (rc) err = ibx_last_error(nullptr);
    if (!rc && S->cfg.rdf_nbin > 0) {
        rc = ibx_rdf_enable(ctx, S->cfg.rdf_nbin, S->cfg.rdf_dr, 1);
        if (rc) err = ibx_last_error(ctx);
    }
    for (int k = n; !rc && k < nplace; ++k)
        if ((rc = ibx_retire(ctx, k))) err = ibx_last_error(ctx);
    // each job stops by itself after its last step, so the cohort steps in whole graph blocks
    for (int k = 0; !rc && k < n; ++k)
        if ((rc = ibx_set_last_step(ctx, k, jobs[k]->nsteps))) err = ibx_last_error(ctx);
    int64_t cap[16] = {};
    if (!rc && (rc = ibx_get_capacity(ctx, 0, cap))) err = ibx_last_error(ctx);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->t_create += S->now() - tc0;
        if (!rc) {
            for (int k = 0; k < nplace; ++k) C->places[k].cap = slot_capacity(ctx, k);
            C->period = (int)cap[14];
            C->list_cap = (int)cap[13];
            C->ready = true;
        } else {
            C->draining = true;
        }
        S->cv_sched.notify_all();
    }
    const int period = std::max(1, C->period);
    const int chunk = std::max(period, S->cfg.chunk_steps / period * period);
    int done = 0;   // always a multiple of the block period
    while (!rc) {
        // hand-overs due at this block boundary
        std::vector<std::pair<int, JobPtr>> due;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (int k = 0; k < nplace; ++k)
                if (C->places[k].pending) due.emplace_back(k, C->places[k].pending);
        }
        for (auto &d : due) {
            Job &j = *d.second;
            const double ta0 = S->now();
            int r = assign_system(ctx, d.first, std::move(j.hs), done);   // j.hs kept when it fails
            if (!r) r = ibx_set_last_step(ctx, d.first, j.nsteps);
            std::lock_guard<std::mutex> lk(S->mu);
            S->t_assign += S->now() - ta0;
            ++S->nassign;
            Place &p = C->places[d.first];
            p.pending.reset();
            if (!r) {
                p.job = d.second;
                j.t_admit = S->now();
                j.cohort = C->id;
                j.place = d.first;
                j.step0 = done;
                j.cohort_jobs = live_jobs(*C);
                ++S->cohort_jobs_total;
                ++S->nadmit_running;
            } else {
                p.dead = true;              // back to the queue, to start elsewhere
                S->queue.push_front(d.second);
                S->cv_sched.notify_all();
            }
        }
        // the next chunk ends at the block boundary at or after the first job's last step, or at the next
        // block boundary when a hand-over is due
        int target = INT_MAX;
        bool any = false, pending = false;
        double natoms = 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (const Place &p : C->places) {
                if (p.job) {
                    any = true;
                    target = std::min(target, p.job->step0 + p.job->nsteps);
                    natoms += p.job->natoms;
                }
                pending = pending || p.pending;
            }
            if (!any && !pending) {
                C->closing = true;          // nothing left: no job will be given to it now
                break;
            }
        }
        if (!any) continue;                 // a hand-over is due now
        target = (target + period - 1) / period * period;
        if (pending) target = std::min(target, done + period);
        const int nstep = std::min(target - done, chunk);
        const double ts0 = S->now();
        rc = ibx_step_graph(ctx, nstep, done);
        const double ts1 = S->now();
        if (rc) {
            err = ibx_last_error(ctx);
            break;
        }
        done += nstep;
        std::vector<std::pair<int, JobPtr>> ended;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            C->done = done;
            S->atom_steps_run += natoms * nstep;
            S->t_step += ts1 - ts0;
            ++S->nchunks;
            S->nsteps_run += nstep;
            for (int k = 0; k < nplace; ++k)   // stopped by themselves at their last steps
                if (C->places[k].job && C->places[k].job->step0 + C->places[k].job->nsteps <= done)
                    ended.emplace_back(k, C->places[k].job);
        }
        for (auto &e : ended) {
            const double tq0 = S->now();
            int r = collect(S, ctx, e.first, *e.second);
            {
                std::lock_guard<std::mutex> lk(S->mu);
                S->t_collect += S->now() - tq0;
            }
            finish_job(S, *e.second, r, r ? ibx_last_error(ctx) : "");
            std::lock_guard<std::mutex> lk(S->mu);
            C->places[e.first].job.reset();
            S->cv_sched.notify_all();       // a free place
        }
    }
    // an error ends the cohort: its jobs fail
    std::vector<JobPtr> left;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        C->closing = true;
        for (Place &p : C->places) {
            if (p.job) left.push_back(p.job);
            if (p.pending) left.push_back(p.pending);
            p.job.reset();
            p.pending.reset();
        }
    }
    for (auto &j : left) finish_job(S, *j, rc ? rc : IBX_E_DEVICE, err);
    if (ctx) ibx_destroy(ctx);
    std::lock_guard<std::mutex> lk(S->mu);
    S->used -= C->bytes;
    --S->running;
    S->cohorts.erase(C->id);
    S->finished.push_back(C->id);
    S->cv_sched.notify_all();
}

// Smallest free place of a running cohort that holds job j.  Lock held.
bool find_place(ibx_server *S, const Job &j, CohortPtr &best, int &best_k) {
    best.reset();
    best_k = -1;
    if (!j.movable) return false;
    for (auto &kv : S->cohorts) {
        const CohortPtr &C = kv.second;
        if (!C->continuous || !C->ready || C->draining || C->closing || C->nupdate != j.nupdate) continue;
        if (j.list_need > C->list_cap) continue;
        for (int k = 0; k < (int)C->places.size(); ++k) {
            const Place &p = C->places[k];
            if (p.job || p.pending || p.dead) continue;
            if (!slot_fits(p.cap, j.need) || j.need.items >= p.cap.items) continue;
            if (!best || p.cap.items < best->places[best_k].cap.items) {
                best = C;
                best_k = k;
            }
        }
    }
    return (bool)best;
}

// Takes queued jobs for a new cohort, in arrival order, while their reserved
// bytes fit the free budget: jobs whose update interval differs from the
// oldest one's (or, with a step band, whose step count is too far from it)
// wait, and a hybrid system runs alone.  Then spare places, per size class
// seen, while they fit.  Lock held.
std::vector<JobPtr> take_cohort(ibx_server *S, size_t &bytes, std::vector<SysPtr> &spare_sys,
                                std::vector<SlotSizes> &spare_res) {
    std::vector<JobPtr> take;
    bytes = 0;
    const double ratio = S->cfg.class_ratio;
    const Job &head = *S->queue.front();
    for (auto it = S->queue.begin(); it != S->queue.end() && (int)take.size() < S->cfg.max_cohort_jobs;) {
        Job &j = **it;
        const double band = S->cfg.step_band;
        const bool fits_steps = band <= 1.0 || (j.nsteps <= band * head.nsteps && head.nsteps <= band * j.nsteps);
        const bool fits_kind =
            j.nupdate == head.nupdate && fits_steps && (take.empty() || (!j.hybrid && !head.hybrid));
        if (!fits_kind) {
            ++it;
            continue;
        }
        if (j.movable && S->cfg.spare_places >= 0) j.hs.reserve = class_reserve(j.hs, j.cls, ratio);
        const size_t b = estimate_device_bytes(j.hs, S->cfg.rdf_nbin);
        if (S->used + bytes + b > S->budget) {
            j.hs.reserve = SlotSizes();
            break;
        }
        bytes += b;
        take.push_back(*it);
        it = S->queue.erase(it);
        if (head.hybrid) break;
    }
    for (const JobPtr &j : take)
        if (!j->movable) return take;   // places of such a batch are not handed over
    if (take.empty() || S->cfg.spare_places < 0) return take;
    for (auto &kv : S->templates) {
        if (kv.second->nupdate != take[0]->nupdate) continue;
        const SlotSizes res = class_reserve(*kv.second, kv.first, ratio);
        const size_t b = estimate_device_bytes(*kv.second, res, S->cfg.rdf_nbin);
        for (int r = 0; r < S->cfg.spare_places; ++r) {
            if (S->used + bytes + b > S->budget) return take;
            bytes += b;
            spare_sys.push_back(kv.second);
            spare_res.push_back(res);
        }
    }
    return take;
}

// The running cohort with the least work left (atom-steps of its jobs), which then takes no more jobs.  Lock held.
void drain_one(ibx_server *S) {
    CohortPtr pick;
    double least = 0;
    for (auto &kv : S->cohorts) {
        const CohortPtr &C = kv.second;
        if (C->draining) return;            // one at a time
        if (C->closing || !C->ready) continue;
        double w = 0;
        for (const Place &p : C->places) {
            if (p.job) w += (double)p.job->natoms * (p.job->step0 + p.job->nsteps - C->done);
            if (p.pending) w += (double)p.pending->natoms * p.pending->nsteps;
        }
        if (!pick || w < least) {
            pick = C;
            least = w;
        }
    }
    if (pick) pick->draining = true;
}

void fail_queued(ibx_server *S, const JobPtr &j, int code, const char *msg) {
    j->status = code;
    j->err = msg;
    j->t_done = S->now();
    j->done = true;
    ++S->nfailed;
    S->cv_done.notify_all();
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
        // queued jobs, in arrival order, into free places of running cohorts
        for (auto it = S->queue.begin(); it != S->queue.end();) {
            CohortPtr C;
            int k = -1;
            if (find_place(S, **it, C, k)) {
                C->places[k].pending = *it;
                it = S->queue.erase(it);
            } else {
                ++it;
            }
        }
        bool formed = false;
        if (!S->queue.empty() && S->running < S->cfg.max_cohorts) {
            JobPtr head = S->queue.front();
            if (head->bytes > S->budget) {
                // can never fit: fail it rather than block the queue
                S->queue.pop_front();
                fail_queued(S, head, IBX_E_NOMEM, "job needs more device bytes than the server budget");
                continue;
            }
            size_t bytes = 0;
            std::vector<SysPtr> spare_sys;
            std::vector<SlotSizes> spare_res;
            std::vector<JobPtr> take = take_cohort(S, bytes, spare_sys, spare_res);
            if (!take.empty()) {
                auto C = std::make_shared<Cohort>();
                C->id = S->ncohorts++;
                C->bytes = bytes;
                C->nupdate = take[0]->nupdate;
                C->continuous = S->cfg.spare_places >= 0;
                for (const JobPtr &j : take) C->continuous = C->continuous && j->movable;
                C->places.resize(take.size() + spare_sys.size());
                S->cohorts[C->id] = C;
                S->used += bytes;
                S->peak_used = std::max(S->peak_used, S->used);
                ++S->running;
                S->peak_running = std::max(S->peak_running, S->running);
                S->max_places = std::max(S->max_places, (int)C->places.size());
                S->cohort_jobs_total += (int64_t)take.size();
                // the spare systems are copied on the cohort's thread (templates are never changed)
                S->workers[C->id] = std::thread([S, C, take, spare_sys, spare_res]() mutable {
                    std::vector<HostSystem> spares;
                    spares.reserve(spare_sys.size());
                    for (size_t q = 0; q < spare_sys.size(); ++q) {
                        spares.push_back(*spare_sys[q]);
                        spares.back().reserve = spare_res[q];
                    }
                    run_cohort(S, C, std::move(take), std::move(spares));
                });
                formed = true;
            }
        }
        if (formed) continue;
        if (!S->queue.empty()) {
            // the oldest job waits for a place: after a while, drain a cohort to make room
            const double waited = S->now() - S->queue.front()->t_submit;
            if (waited >= kDrainAfterS) {
                drain_one(S);
                S->cv_sched.wait_for(lk, std::chrono::milliseconds(50));
            } else {
                S->cv_sched.wait_for(lk, std::chrono::duration<double>(kDrainAfterS - waited));
            }
            continue;
        }
        S->cv_sched.wait(lk);
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
    if (c.spare_places == 0) c.spare_places = 1;
    if (!(c.class_ratio > 1.0)) c.class_ratio = 1.25;
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
    j->movable = !j->hs.hybrid && j->hs.nvs == 0 && j->hs.ngroup == 1;
    j->need = slot_needs(j->hs);
    j->list_need = list_capacity_needed(j->hs);
    j->cls = size_class(j->nitems + 1, S->cfg.class_ratio);   // one padding item
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop) return serr(IBX_E_INVALID, "server is shutting down");
    const bool new_class = j->movable && S->cfg.spare_places > 0 && !S->templates.count(j->cls);
    lk.unlock();
    SysPtr tmpl = new_class ? std::make_shared<const HostSystem>(j->hs) : nullptr;   // copied without the lock
    lk.lock();
    if (tmpl && !S->templates.count(j->cls)) S->templates[j->cls] = tmpl;
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
    st[11] = S->max_places;
    st[12] = S->ncohorts ? (double)S->cohort_jobs_total / S->ncohorts : 0.0;
    st[13] = (double)S->budget;
    st[14] = (double)S->nadmit_running;
    st[15] = S->atom_steps_run;
    return IBX_OK;
}

int ibx_server_timing(ibx_server *S, double t[8]) {
    if (!S || !t) return serr(IBX_E_INVALID, "invalid arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    const double v[8] = {S->t_step, S->t_assign, S->t_collect, S->t_create,
                         (double)S->nchunks, (double)S->nassign, (double)S->nsteps_run, 0.0};
    for (int k = 0; k < 8; ++k) t[k] = v[k];
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
