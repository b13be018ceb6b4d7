// Continuous-batching request loop for G0W0 jobs on one device.
//
// Worker threads each own a context (its own HIP stream), so jobs admitted
// while others run overlap on the GPU.  A worker that takes a job also takes
// every other admitted job that can share its frequency batches (same
// padded particle-hole order and form, eta == 0, not the decomposed form; up
// to mgw_hip_server_set_grouping jobs) and runs them through
// mgw_hip_g0w0_group: one launch sequence factors all their frequencies.  Admission is by an HBM estimate:
// a job starts when the bytes reserved by running jobs plus its estimate
// (resident ERI + V + frequency batch + contraction scratch,
// mgw_hip_estimate_bytes, plus the error judge's FP64 panel for the mixed
// forms) fit the server's budget; jobs start in arrival order.  A job whose
// estimate exceeds the whole budget fails with MGW_ERR_MEMORY instead of
// waiting forever.  After each job a worker frees the device buffers of its
// grouped members; its lead context keeps its own (reused by the next job
// when large enough) while they hold under a quarter of the budget's
// per-worker share, so the reservation exceeds the running jobs' estimates
// by at most that much per worker.  The ERI storage of a job
// (mgw_hip_server_set_eri_storage) enters its estimate: a streamed job
// reserves the held (pk|ia) block and one upload slab instead of the nmo^4
// tensor, and under MGW_ERI_AUTO a job that does not fit resident starts
// streamed when that fits (the degrade path).
//
// Mixed-precision jobs run under the context's error judge: a sample of their
// frequencies is recomputed in FP64, and a failed judge or a factorization
// breakdown reruns the job in FP64 (mgw_hip_set_judge).  A job can be
// cancelled before it starts or between two of its frequency batches.
// Finished jobs keep their record (status, error, judge outcome) until
// mgw_hip_server_release.  Destroying the server cancels jobs that have not
// started and waits for the running ones.
#include "mgw_hip.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct ServeJob {
    int64_t ticket = 0;
    mgw_g0w0_job job{};
    std::vector<double> mo, vxc;
    std::vector<int32_t> states;
    const double* eri = nullptr;
    const double* factor = nullptr;
    int32_t kdim = 0;
    mgw_g0w0_out out{};
    size_t estimate = 0;      // reserved while running (of the storage taken)
    size_t est_resident = 0, est_streamed = 0;
    int32_t storage_req = MGW_ERI_RESIDENT;
    int32_t storage = MGW_ERI_AUTO;     // taken at start
    int status = -1;          // -1 pending, -2 running, else final mgw_status
    int32_t cancel = 0;       // polled by the running job between frequency batches
    double judge_err = -1.0;
    int32_t fell_back = 0;
    std::string error;
    Clock::time_point t_submit, t_start, t_end;
};

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Device bytes of the error judge of a mixed job: its FP64 panel of the
// contraction columns (mpad x np) and one row of quadratic forms.
size_t judge_bytes(const mgw_g0w0_job& g) {
    const size_t ns = g.states && g.nstates > 0 ? (size_t)g.nstates : (size_t)g.nmo;
    const size_t m = ns * (size_t)g.nmo;
    const size_t n = (size_t)g.nocc * (size_t)(g.nmo - g.nocc);
    return round_up(m, 256) * round_up(n, 256) * 8 + m * 8;
}

}  // namespace

struct mgw_hip_server {
    int device = 0;
    size_t budget = 0;
    int32_t batch_freqs = 32;
    // error judge of mixed jobs, applied to each worker's context
    int32_t judge_n = 2;
    double judge_tol = 0.0;
    int32_t judge_fallback = 1;
    double budget_scale = 1.0;        // fault-injection hook: per-job device budget scale
    int32_t group_max = 16;           // jobs one worker runs together (1: no grouping)
    int32_t eri_storage = MGW_ERI_RESIDENT;   // storage of jobs submitted from now on
    int32_t graphs = 1;               // workers' contexts replay frequency batches as hipGraphs
    // A worker's lead context keeps its device buffers from one job to the
    // next (reused when large enough) while they stay under a quarter of the
    // budget's per-worker share: allocating and freeing every job's ERI and
    // batch buffers slowed the loop down with the age of the process
    // (MGW_SERVE_KEEP_BUFFERS=0: free them after every job).
    int32_t keep_buffers = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeJob*> pending;
    std::map<int64_t, std::unique_ptr<ServeJob>> jobs;
    int64_t next_ticket = 1;
    size_t reserved = 0;
    int running = 0, peak_running = 0;
    int64_t completed = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    // a job that ends before it starts (lock held)
    void finish_unstarted(ServeJob* j, int status, const char* why) {
        j->status = status;
        j->error = why;
        j->t_start = j->t_end = Clock::now();
        ++completed;
    }

    // Jobs that can share frequency batches (mgw_hip_g0w0_group): eta == 0,
    // not the decomposed form, same precision and padded particle-hole order.
    static bool groupable(const ServeJob* j) {
        return j->job.eta == 0.0 && j->job.precision != MGW_PREC_FP64_CD;
    }
    static int padded_order(const ServeJob* j) {
        const int n = j->job.nocc * (j->job.nmo - j->job.nocc);
        return (n + 255) / 256 * 256;
    }
    static bool same_group(const ServeJob* a, const ServeJob* b) {
        return groupable(b) && a->job.precision == b->job.precision &&
               padded_order(a) == padded_order(b);
    }

    // The storage a job would start in now and the bytes it would reserve
    // (lock held): its requested storage, or under MGW_ERI_AUTO resident when
    // that fits the budget left, else streamed.
    size_t start_bytes(const ServeJob* j, int32_t* storage) const {
        int32_t st = j->storage_req;
        if (st == MGW_ERI_AUTO)
            st = reserved + j->est_resident <= budget ? MGW_ERI_RESIDENT : MGW_ERI_STREAMED;
        if (storage) *storage = st;
        return st == MGW_ERI_RESIDENT ? j->est_resident : j->est_streamed;
    }
    // the smallest reservation the job can ever start with
    static size_t least_bytes(const ServeJob* j) {
        return j->storage_req == MGW_ERI_RESIDENT ? j->est_resident : j->est_streamed;
    }
    bool can_start(const ServeJob* j) const {
        return reserved + start_bytes(j, nullptr) <= budget;
    }
    void start(ServeJob* j) {
        int32_t st = MGW_ERI_RESIDENT;
        j->estimate = start_bytes(j, &st);
        j->storage = st;
        reserved += j->estimate;
    }

    void work() {
        // pool[0] leads a group (its stream and batch buffers run the shared
        // batches); pool[k] holds the k-th grouped job's resident ERI
        std::vector<mgw_hip_ctx*> pool(1, nullptr);
        const int rc = mgw_hip_create(device, &pool[0]);
        if (rc == MGW_OK) mgw_hip_set_graphs(pool[0], graphs);   // batches of a job replay one graph
        for (;;) {
            std::vector<ServeJob*> group;
            int32_t jn = 0, jf = 0;
            double jt = 0.0, scale = 1.0;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (pending.empty()) return false;
                    ServeJob* h = pending.front();
                    return h->cancel || least_bytes(h) > budget || can_start(h);
                });
                if (stop) {
                    while (!pending.empty()) {
                        finish_unstarted(pending.front(), MGW_ERR_STATE,
                                         "server stopped before the job started");
                        pending.pop_front();
                    }
                    cv_done.notify_all();
                    break;
                }
                ServeJob* j = pending.front();
                pending.pop_front();
                if (j->cancel || least_bytes(j) > budget || rc != MGW_OK) {
                    if (j->cancel)
                        finish_unstarted(j, MGW_ERR_STATE, "cancelled before the job started");
                    else if (rc != MGW_OK)
                        finish_unstarted(j, rc, "worker could not create a device context");
                    else
                        finish_unstarted(j, MGW_ERR_MEMORY, "job estimate exceeds the server's HBM budget");
                    cv_done.notify_all();
                    cv_work.notify_all();
                    continue;
                }
                group.push_back(j);
                start(j);
                // every other admitted job that shares the head job's batches
                // joins it (continuous batching: later arrivals that fit the
                // budget start with it)
                if (group_max > 1 && groupable(j)) {
                    for (auto it = pending.begin();
                         it != pending.end() && (int)group.size() < group_max;) {
                        ServeJob* c = *it;
                        if (!c->cancel && same_group(j, c) && can_start(c)) {
                            start(c);
                            group.push_back(c);
                            it = pending.erase(it);
                        } else {
                            ++it;
                        }
                    }
                }
                running += (int)group.size();
                peak_running = std::max(peak_running, running);
                for (ServeJob* g : group) {
                    g->status = -2;
                    g->t_start = Clock::now();
                }
                jn = judge_n;
                jt = judge_tol;
                jf = judge_fallback;
                scale = budget_scale;
            }
            const int k = (int)group.size();
            while ((int)pool.size() < k) {
                mgw_hip_ctx* c = nullptr;
                if (mgw_hip_create(device, &c) != MGW_OK) break;
                mgw_hip_set_graphs(c, graphs);
                pool.push_back(c);
            }
            std::vector<int> st(k, MGW_OK);
            std::vector<std::string> err(k);
            if ((int)pool.size() < k) {
                // no context for the whole group: the jobs beyond the pool fail
                for (int i = (int)pool.size(); i < k; ++i) {
                    st[i] = MGW_ERR_HIP;
                    err[i] = "worker could not create a device context for a grouped job";
                }
            }
            // the lead's budget and batch limit cover the group's stacked batches
            size_t sum_est = 0;
            int32_t sum_nb = 0;
            for (ServeJob* g : group) {
                sum_est += g->estimate;
                sum_nb += std::min(g->job.nfreq, batch_freqs);
            }
            std::vector<mgw_hip_ctx*> ctxs;
            std::vector<mgw_g0w0_job> jobsv;
            std::vector<mgw_g0w0_out> outsv;
            std::vector<int> slot;
            for (int i = 0; i < k; ++i) {
                if (st[i] != MGW_OK) continue;
                ServeJob* j = group[i];
                mgw_hip_ctx* ctx = pool[i];
                const size_t est = i == 0 ? sum_est : j->estimate;
                // a lead that keeps buffers holds some of an earlier job's: no per-job
                // cap then (admission still reserves the estimate), unless the
                // fault-injection hook scales the budget
                const bool uncapped = keep_buffers && i == 0 && scale == 1.0;
                int r = mgw_hip_set_memory_budget(
                    ctx, uncapped ? 0 : std::max<size_t>(1, (size_t)((double)est * scale)));
                if (r == MGW_OK) r = mgw_hip_set_batch_limit(ctx, i == 0 ? sum_nb : batch_freqs);
                if (r == MGW_OK) r = mgw_hip_set_judge(ctx, jn, jt, jf);
                if (r == MGW_OK) r = mgw_hip_set_cancel_flag(ctx, &j->cancel);
                if (r == MGW_OK) r = mgw_hip_set_eri_storage(ctx, j->storage);
                if (r == MGW_OK) {
                    r = j->factor ? mgw_hip_upload_eri_factor(ctx, j->factor, j->job.nmo, j->job.nocc, j->kdim)
                                  : mgw_hip_upload_eri(ctx, j->eri, j->job.nmo, j->job.nocc);
                }
                if (r != MGW_OK) {
                    st[i] = r;
                    err[i] = mgw_hip_last_error(ctx);
                    continue;
                }
                ctxs.push_back(ctx);
                jobsv.push_back(j->job);
                outsv.push_back(j->out);
                slot.push_back(i);
            }
            if (!ctxs.empty()) {
                std::vector<int32_t> gst(ctxs.size(), MGW_OK);
                if (ctxs.size() == 1)
                    gst[0] = mgw_hip_g0w0(ctxs[0], &jobsv[0], &outsv[0]);
                else
                    mgw_hip_g0w0_group(ctxs.data(), jobsv.data(), outsv.data(), (int32_t)ctxs.size(),
                                       gst.data());
                for (size_t q = 0; q < ctxs.size(); ++q) {
                    st[slot[q]] = gst[q];
                    if (gst[q] != MGW_OK) err[slot[q]] = mgw_hip_last_error(ctxs[q]);
                }
            }
            std::vector<double> jerr(k, -1.0);
            std::vector<int32_t> fell(k, 0);
            for (int i = 0; i < k && i < (int)pool.size(); ++i) {
                if (st[i] == MGW_OK) mgw_hip_last_judge(pool[i], &jerr[i], nullptr, &fell[i]);
                mgw_hip_set_cancel_flag(pool[i], nullptr);
                size_t held = 0;
                const bool keep = keep_buffers && i == 0 && st[i] == MGW_OK &&
                                  mgw_hip_device_bytes(pool[i], &held) == MGW_OK &&
                                  held <= budget / (4 * workers.size());
                if (!keep) mgw_hip_release_buffers(pool[i]);
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                for (int i = 0; i < k; ++i) {
                    ServeJob* j = group[i];
                    j->status = st[i];
                    j->error = err[i];
                    j->judge_err = jerr[i];
                    j->fell_back = fell[i];
                    j->t_end = Clock::now();
                    reserved -= j->estimate;
                    --running;
                    ++completed;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
        for (mgw_hip_ctx* c : pool)
            if (c) mgw_hip_destroy(c);
    }

    ServeJob* find(int64_t ticket) {
        auto it = jobs.find(ticket);
        return it == jobs.end() ? nullptr : it->second.get();
    }
};

extern "C" {

int mgw_hip_server_create(int device, int32_t workers, size_t hbm_budget, int32_t batch_freqs,
                          mgw_hip_server** out) {
    if (!out || workers <= 0 || hbm_budget == 0 || batch_freqs <= 0) return MGW_ERR_ARG;
    auto* s = new mgw_hip_server();
    s->device = device;
    s->budget = hbm_budget;
    s->batch_freqs = batch_freqs;
    if (const char* env = std::getenv("MGW_SERVE_GRAPHS")) s->graphs = env[0] != '0';
    if (const char* env = std::getenv("MGW_SERVE_KEEP_BUFFERS")) s->keep_buffers = env[0] != '0';
    for (int i = 0; i < workers; ++i) s->workers.emplace_back([s] { s->work(); });
    *out = s;
    return MGW_OK;
}

int mgw_hip_server_destroy(mgw_hip_server* s) {
    if (!s) return MGW_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
    return MGW_OK;
}

int mgw_hip_server_set_judge(mgw_hip_server* s, int32_t nsample, double tol, int32_t fallback) {
    if (!s || nsample < 0 || nsample > 64 || !(tol >= 0.0)) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->judge_n = nsample;
    s->judge_tol = tol;
    s->judge_fallback = fallback != 0;
    return MGW_OK;
}

int mgw_hip_server_set_grouping(mgw_hip_server* s, int32_t max_jobs) {
    if (!s || max_jobs < 1 || max_jobs > 64) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->group_max = max_jobs;
    return MGW_OK;
}

int mgw_hip_server_set_budget_scale(mgw_hip_server* s, double scale) {
    if (!s || !(scale > 0.0 && scale <= 1.0)) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->budget_scale = scale;
    return MGW_OK;
}

int mgw_hip_server_set_eri_storage(mgw_hip_server* s, int32_t storage) {
    if (!s || (storage != MGW_ERI_AUTO && storage != MGW_ERI_RESIDENT && storage != MGW_ERI_STREAMED))
        return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->eri_storage = storage;
    return MGW_OK;
}

int mgw_hip_server_job_eri_storage(mgw_hip_server* s, int64_t ticket, int32_t* storage) {
    if (!s || !storage) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    ServeJob* j = s->find(ticket);
    if (!j) return MGW_ERR_ARG;
    *storage = j->storage;
    return MGW_OK;
}

int mgw_hip_server_submit(mgw_hip_server* s, const mgw_serve_job* job, const mgw_g0w0_out* out,
                          int64_t* ticket) {
    if (!s || !job || !out || !ticket || !job->job.mo_energy || !job->job.vxc_diag || !out->qp)
        return MGW_ERR_ARG;
    if ((job->eri == nullptr) == (job->eri_factor == nullptr)) return MGW_ERR_ARG;
    const mgw_g0w0_job& g = job->job;
    if (g.nmo <= 1 || g.nocc <= 0 || g.nocc >= g.nmo || g.nfreq <= 0) return MGW_ERR_ARG;
    if (g.states && g.nstates > g.nmo) return MGW_ERR_ARG;
    auto j = std::make_unique<ServeJob>();
    j->job = g;
    j->mo.assign(g.mo_energy, g.mo_energy + g.nmo);
    j->vxc.assign(g.vxc_diag, g.vxc_diag + g.nmo);
    j->job.mo_energy = j->mo.data();
    j->job.vxc_diag = j->vxc.data();
    if (g.states && g.nstates > 0) {
        j->states.assign(g.states, g.states + g.nstates);
        j->job.states = j->states.data();
    } else {
        j->job.states = nullptr;
        j->job.nstates = 0;
    }
    j->eri = job->eri;
    j->factor = job->eri_factor;
    j->kdim = job->kdim;
    j->out = *out;
    const int nb = std::min(g.nfreq, s->batch_freqs);
    size_t est = mgw_hip_estimate_bytes(g.nmo, g.nocc, g.nfreq, g.nstates, nb, g.precision);
    if (job->eri_factor) {
        // the factor on the device during the upload (64 more rows when streamed)
        const size_t n2 = (size_t)g.nmo * g.nmo;
        est += (round_up(n2, 64) + 64) * round_up((size_t)job->kdim, 16) * 8;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->judge_n > 0 && (g.precision == MGW_PREC_MIXED || g.precision == MGW_PREC_MIXED_INT8 ||
                           g.precision == MGW_PREC_MIXED_INT8_4))
        est += judge_bytes(g);
    // the decomposed form factors the whole ERI pair matrix: resident only
    j->storage_req = g.precision == MGW_PREC_FP64_CD ? MGW_ERI_RESIDENT : s->eri_storage;
    j->est_resident = est;
    j->est_streamed = est - mgw_hip_eri_bytes(g.nmo, g.nocc, MGW_ERI_RESIDENT) +
                      mgw_hip_eri_bytes(g.nmo, g.nocc, MGW_ERI_STREAMED);
    j->t_submit = Clock::now();
    j->ticket = s->next_ticket++;
    *ticket = j->ticket;
    s->pending.push_back(j.get());
    s->jobs[j->ticket] = std::move(j);
    s->cv_work.notify_one();
    return MGW_OK;
}

int mgw_hip_server_cancel(mgw_hip_server* s, int64_t ticket) {
    if (!s) return MGW_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        ServeJob* j = s->find(ticket);
        if (!j) return MGW_ERR_ARG;
        if (j->status >= 0) return MGW_OK;                 // already finished
        __atomic_store_n(&j->cancel, 1, __ATOMIC_RELEASE);
        if (j->status == -1) {                             // pending: finish it here
            s->pending.erase(std::find(s->pending.begin(), s->pending.end(), j));
            s->finish_unstarted(j, MGW_ERR_STATE, "cancelled before the job started");
        }
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
    return MGW_OK;
}

int mgw_hip_server_wait(mgw_hip_server* s, int64_t ticket, double timeout_s, int32_t* status,
                        double* latency_ms, double* run_ms) {
    if (!s || !status) return MGW_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    ServeJob* j = s->find(ticket);
    if (!j) return MGW_ERR_ARG;
    auto done = [&] { return j->status >= 0; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        *status = -1;
        return MGW_OK;
    }
    *status = j->status;
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(j->t_end - j->t_submit).count();
    if (run_ms) *run_ms = std::chrono::duration<double, std::milli>(j->t_end - j->t_start).count();
    return MGW_OK;
}

const char* mgw_hip_server_job_error(mgw_hip_server* s, int64_t ticket) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    ServeJob* j = s->find(ticket);
    return j ? j->error.c_str() : "unknown ticket";
}

int mgw_hip_server_job_judge(mgw_hip_server* s, int64_t ticket, double* max_rel_err,
                             int32_t* fell_back) {
    if (!s) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    ServeJob* j = s->find(ticket);
    if (!j || j->status < 0) return MGW_ERR_ARG;
    if (max_rel_err) *max_rel_err = j->judge_err;
    if (fell_back) *fell_back = j->fell_back;
    return MGW_OK;
}

int mgw_hip_server_release(mgw_hip_server* s, int64_t ticket) {
    if (!s) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end() || it->second->status < 0) return MGW_ERR_ARG;
    s->jobs.erase(it);
    return MGW_OK;
}

int mgw_hip_server_stats(mgw_hip_server* s, int64_t* completed, int32_t* running,
                         int32_t* peak_running, size_t* reserved) {
    if (!s) return MGW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->completed;
    if (running) *running = s->running;
    if (peak_running) *peak_running = s->peak_running;
    if (reserved) *reserved = s->reserved;
    return MGW_OK;
}

}  // extern "C"
