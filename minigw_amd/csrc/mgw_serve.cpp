// Continuous-batching request loop for G0W0 jobs on one device.
//
// Worker threads each own a context (its own HIP stream), so jobs admitted
// while others run overlap on the GPU.  Admission is by an HBM estimate:
// a job starts when the bytes reserved by running jobs plus its estimate
// (resident ERI + V + frequency batch + contraction scratch,
// mgw_hip_estimate_bytes) fit the server's budget; jobs start in arrival
// order.  A job whose estimate exceeds the whole budget fails with
// MGW_ERR_MEMORY instead of waiting forever.  After each job a worker frees
// its device buffers, so the reservation is the job's footprint only.
// Destroying the server cancels jobs that have not started and waits for the
// running ones.
#include "mgw_hip.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
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
    size_t estimate = 0;
    int status = -1;          // -1 pending, -2 running, else final mgw_status
    std::string error;
    Clock::time_point t_submit, t_start, t_end;
};

}  // namespace

struct mgw_hip_server {
    int device = 0;
    size_t budget = 0;
    int32_t batch_freqs = 32;
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
    std::string error;

    void work() {
        mgw_hip_ctx* ctx = nullptr;
        const int rc = mgw_hip_create(device, &ctx);
        if (rc == MGW_OK) mgw_hip_set_graphs(ctx, 1);   // batches of a job replay one graph
        for (;;) {
            ServeJob* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (pending.empty()) return false;
                    ServeJob* h = pending.front();
                    return h->estimate > budget || reserved + h->estimate <= budget;
                });
                if (stop) {
                    while (!pending.empty()) {
                        ServeJob* c = pending.front();
                        pending.pop_front();
                        c->status = MGW_ERR_STATE;
                        c->error = "server stopped before the job started";
                        c->t_start = c->t_end = Clock::now();
                    }
                    cv_done.notify_all();
                    break;
                }
                j = pending.front();
                pending.pop_front();
                if (j->estimate > budget || rc != MGW_OK) {
                    j->status = rc != MGW_OK ? rc : MGW_ERR_MEMORY;
                    j->error = rc != MGW_OK ? "worker could not create a device context"
                                            : "job estimate exceeds the server's HBM budget";
                    j->t_start = j->t_end = Clock::now();
                    ++completed;
                    cv_done.notify_all();
                    cv_work.notify_all();
                    continue;
                }
                reserved += j->estimate;
                ++running;
                peak_running = std::max(peak_running, running);
                j->status = -2;
                j->t_start = Clock::now();
            }
            int st = mgw_hip_set_memory_budget(ctx, j->estimate);
            if (st == MGW_OK) {
                st = j->factor ? mgw_hip_upload_eri_factor(ctx, j->factor, j->job.nmo, j->job.nocc, j->kdim)
                               : mgw_hip_upload_eri(ctx, j->eri, j->job.nmo, j->job.nocc);
            }
            if (st == MGW_OK) st = mgw_hip_g0w0(ctx, &j->job, &j->out);
            std::string err = st == MGW_OK ? "" : mgw_hip_last_error(ctx);
            mgw_hip_release_buffers(ctx);
            {
                std::lock_guard<std::mutex> lk(mu);
                j->status = st;
                j->error = err;
                j->t_end = Clock::now();
                reserved -= j->estimate;
                --running;
                ++completed;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
        if (ctx) mgw_hip_destroy(ctx);
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

int mgw_hip_server_submit(mgw_hip_server* s, const mgw_serve_job* job, const mgw_g0w0_out* out,
                          int64_t* ticket) {
    if (!s || !job || !out || !ticket || !job->job.mo_energy || !job->job.vxc_diag || !out->qp)
        return MGW_ERR_ARG;
    if ((job->eri == nullptr) == (job->eri_factor == nullptr)) return MGW_ERR_ARG;
    const mgw_g0w0_job& g = job->job;
    if (g.nmo <= 1 || g.nocc <= 0 || g.nocc >= g.nmo || g.nfreq <= 0) return MGW_ERR_ARG;
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
    j->estimate = mgw_hip_estimate_bytes(g.nmo, g.nocc, g.nfreq, g.nstates, nb, g.precision);
    if (job->eri_factor) {
        const size_t n2 = (size_t)g.nmo * g.nmo;
        j->estimate += ((n2 + 63) / 64 * 64) * ((size_t)(job->kdim + 15) / 16 * 16) * 8;
    }
    j->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    j->ticket = s->next_ticket++;
    *ticket = j->ticket;
    s->pending.push_back(j.get());
    s->jobs[j->ticket] = std::move(j);
    s->cv_work.notify_one();
    return MGW_OK;
}

int mgw_hip_server_wait(mgw_hip_server* s, int64_t ticket, double timeout_s, int32_t* status,
                        double* latency_ms, double* run_ms) {
    if (!s || !status) return MGW_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return MGW_ERR_ARG;
    ServeJob* j = it->second.get();
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
    if (j->status != MGW_OK) s->error = j->error;
    return MGW_OK;
}

const char* mgw_hip_server_job_error(mgw_hip_server* s, int64_t ticket) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    return it == s->jobs.end() ? "unknown ticket" : it->second->error.c_str();
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
