// Host-only harness of the request loop (minigw_amd/csrc/mgw_serve.cpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitizers.py.  The device layer it calls (mgw_hip_create,
// mgw_hip_g0w0, ...) is replaced here by a stand-in that does no GPU work:
// it checks the server's calls (budget, batch cap, judge, cancel flag), sleeps
// per frequency batch, polls the cancel flag between batches, fails an
// allocation beyond the context's budget, and writes outputs that identify
// the job.  Failure conditions are forced from the host: a cut budget, a
// cancelled job, a judge that reports a fallback, a server destroyed with
// jobs queued.  Exit status 0 = every scenario passed.
#include "mgw_hip.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

struct FakeCtx {
    size_t budget = 0;
    int32_t batch_limit = 0, judge_n = 0, judge_fallback = 0;
    double judge_tol = 0.0;
    const int32_t* cancel = nullptr;
    int32_t storage = MGW_ERI_AUTO;
    int nmo = 0, nocc = 0;
    std::string error;
    double judge_err = -1.0;
    int32_t fell_back = 0;
};

std::atomic<int> g_contexts{0};
std::atomic<int> g_running{0}, g_peak{0};

int ctx_fail(FakeCtx* c, int code, const char* msg) {
    c->error = msg;
    return code;
}

size_t eri_bytes(int nmo) { return (size_t)nmo * nmo * nmo * nmo * 8; }

#define CHECK(cond)                                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::fprintf(stderr, "harness check failed at line %d: %s\n", __LINE__, #cond); \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

}  // namespace

// ------------------------------------------------------- stand-in device layer
extern "C" {

const char* mgw_hip_last_error(const mgw_hip_ctx* ctx) {
    return ctx ? reinterpret_cast<const FakeCtx*>(ctx)->error.c_str() : "";
}

int mgw_hip_create(int device, mgw_hip_ctx** out) {
    if (device != 0 || !out) return MGW_ERR_ARG;
    *out = reinterpret_cast<mgw_hip_ctx*>(new FakeCtx());
    ++g_contexts;
    return MGW_OK;
}

int mgw_hip_destroy(mgw_hip_ctx* ctx) {
    delete reinterpret_cast<FakeCtx*>(ctx);
    --g_contexts;
    return MGW_OK;
}

int mgw_hip_set_graphs(mgw_hip_ctx*, int32_t) { return MGW_OK; }

int mgw_hip_set_memory_budget(mgw_hip_ctx* ctx, size_t bytes) {
    reinterpret_cast<FakeCtx*>(ctx)->budget = bytes;
    return MGW_OK;
}

int mgw_hip_set_batch_limit(mgw_hip_ctx* ctx, int32_t n) {
    if (n < 0) return MGW_ERR_ARG;
    reinterpret_cast<FakeCtx*>(ctx)->batch_limit = n;
    return MGW_OK;
}

int mgw_hip_set_judge(mgw_hip_ctx* ctx, int32_t nsample, double tol, int32_t fallback) {
    auto* c = reinterpret_cast<FakeCtx*>(ctx);
    c->judge_n = nsample;
    c->judge_tol = tol;
    c->judge_fallback = fallback;
    return MGW_OK;
}

int mgw_hip_set_cancel_flag(mgw_hip_ctx* ctx, const int32_t* flag) {
    reinterpret_cast<FakeCtx*>(ctx)->cancel = flag;
    return MGW_OK;
}

int mgw_hip_set_eri_storage(mgw_hip_ctx* ctx, int32_t storage) {
    if (storage != MGW_ERI_AUTO && storage != MGW_ERI_RESIDENT && storage != MGW_ERI_STREAMED)
        return MGW_ERR_ARG;
    reinterpret_cast<FakeCtx*>(ctx)->storage = storage;
    return MGW_OK;
}

// the held ph block plus one slab row when streamed
size_t mgw_hip_eri_bytes(int32_t nmo, int32_t nocc, int32_t storage) {
    if (nmo <= 1 || nocc <= 0 || nocc >= nmo) return 0;
    if (storage == MGW_ERI_RESIDENT) return eri_bytes(nmo);
    if (storage == MGW_ERI_STREAMED)
        return ((size_t)nmo * nmo * nocc * (nmo - nocc) + (size_t)nmo * nmo * nmo) * 8;
    return 0;
}

static int upload(mgw_hip_ctx* ctx, int32_t nmo, int32_t nocc) {
    auto* c = reinterpret_cast<FakeCtx*>(ctx);
    if (c->storage != MGW_ERI_RESIDENT && c->storage != MGW_ERI_STREAMED)
        return ctx_fail(c, MGW_ERR_ARG, "the server must set the job's ERI storage");
    if (c->budget && mgw_hip_eri_bytes(nmo, nocc, c->storage) > c->budget)
        return ctx_fail(c, MGW_ERR_MEMORY, "ERI: allocation would exceed the device budget");
    c->nmo = nmo;
    c->nocc = nocc;
    return MGW_OK;
}

int mgw_hip_upload_eri(mgw_hip_ctx* ctx, const double* eri, int32_t nmo, int32_t nocc) {
    if (!eri) return MGW_ERR_ARG;
    return upload(ctx, nmo, nocc);
}

int mgw_hip_upload_eri_factor(mgw_hip_ctx* ctx, const double* b, int32_t nmo, int32_t nocc,
                              int32_t kdim) {
    if (!b || kdim <= 0) return MGW_ERR_ARG;
    return upload(ctx, nmo, nocc);
}

int mgw_hip_release_buffers(mgw_hip_ctx* ctx) {
    reinterpret_cast<FakeCtx*>(ctx)->nmo = 0;
    return MGW_OK;
}

// bytes held: the resident ERI of the last upload, until released
int mgw_hip_device_bytes(const mgw_hip_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) return MGW_ERR_ARG;
    const FakeCtx* c = reinterpret_cast<const FakeCtx*>(ctx);
    *bytes = c->nmo ? eri_bytes(c->nmo) : 0;
    return MGW_OK;
}

// Outputs: qp[p] = 1000 * nmo + p (+ 0.5 when rerun in FP64), sigma = nfreq.
// The judge and outputs of one finished job.
static void finish_job(FakeCtx* c, const mgw_g0w0_job* job, mgw_g0w0_out* out) {
    const bool mixed = job->precision == MGW_PREC_MIXED || job->precision == MGW_PREC_MIXED_INT8 ||
                       job->precision == MGW_PREC_MIXED_INT8_4;
    c->judge_err = mixed && c->judge_n > 0 ? 1e-9 : -1.0;
    // a tolerance below the stand-in's error is a failed judge
    c->fell_back = mixed && c->judge_n > 0 && c->judge_fallback && c->judge_tol > 0.0 &&
                           c->judge_tol < 1e-9
                       ? 1
                       : 0;
    const int ns = job->nstates > 0 && job->states ? job->nstates : job->nmo;
    for (int p = 0; p < job->nmo; ++p) out->qp[p] = 1000.0 * job->nmo + p + (c->fell_back ? 0.5 : 0.0);
    if (out->sigma_c_re)
        for (int i = 0; i < ns * job->nfreq; ++i) out->sigma_c_re[i] = job->nfreq;
}

static void note_running(int count) {
    const int now = (g_running += count);
    for (int p = g_peak.load(); now > p && !g_peak.compare_exchange_weak(p, now);) {
    }
}

int mgw_hip_g0w0(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, mgw_g0w0_out* out) {
    auto* c = reinterpret_cast<FakeCtx*>(ctx);
    if (c->nmo != job->nmo) return ctx_fail(c, MGW_ERR_STATE, "no resident ERI for the job");
    note_running(1);
    const int nb = c->batch_limit > 0 ? c->batch_limit : 64;
    int status = MGW_OK;
    for (int f0 = 0; f0 < job->nfreq; f0 += nb) {
        if (c->cancel && __atomic_load_n(c->cancel, __ATOMIC_ACQUIRE) != 0) {
            status = ctx_fail(c, MGW_ERR_STATE, "mgw_hip_g0w0: cancelled");
            break;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200 * job->nmo / 8 + 100));
    }
    --g_running;
    if (status != MGW_OK) return status;
    finish_job(c, job, out);
    return MGW_OK;
}

// Groups: the stacked batches of all jobs (one sleep per batch of the lead's
// limit), then each job finished on its own context; a job cancelled while its
// group ran ends with MGW_ERR_STATE.  g_groups counts calls with two or more
// jobs, g_group_peak the largest.
std::atomic<int> g_groups{0}, g_group_peak{0};

int mgw_hip_g0w0_group(mgw_hip_ctx* const* ctxs, const mgw_g0w0_job* jobs, mgw_g0w0_out* outs,
                       int32_t count, int32_t* status) {
    if (!ctxs || !jobs || !outs || count < 1) return MGW_ERR_ARG;
    int total = 0, maxnmo = 0;
    for (int i = 0; i < count; ++i) {
        for (int k = 0; k < i; ++k)
            if (ctxs[k] == ctxs[i]) return MGW_ERR_ARG;
        auto* c = reinterpret_cast<FakeCtx*>(ctxs[i]);
        if (c->nmo != jobs[i].nmo) {
            if (status) status[i] = ctx_fail(c, MGW_ERR_STATE, "no resident ERI for the job");
            return MGW_ERR_STATE;
        }
        total += jobs[i].nfreq;
        maxnmo = std::max(maxnmo, jobs[i].nmo);
    }
    ++g_groups;
    for (int p = g_group_peak.load(); count > p && !g_group_peak.compare_exchange_weak(p, count);) {
    }
    note_running(count);
    const auto* lead = reinterpret_cast<FakeCtx*>(ctxs[0]);
    const int nb = lead->batch_limit > 0 ? lead->batch_limit : 64;
    for (int f0 = 0; f0 < total; f0 += nb)
        std::this_thread::sleep_for(std::chrono::microseconds(200 * maxnmo / 8 + 100));
    g_running -= count;
    int first = MGW_OK;
    for (int i = 0; i < count; ++i) {
        auto* c = reinterpret_cast<FakeCtx*>(ctxs[i]);
        int st = MGW_OK;
        if (c->cancel && __atomic_load_n(c->cancel, __ATOMIC_ACQUIRE) != 0)
            st = ctx_fail(c, MGW_ERR_STATE, "mgw_hip_g0w0: cancelled during its group's batches");
        else
            finish_job(c, &jobs[i], &outs[i]);
        if (status) status[i] = st;
        if (first == MGW_OK) first = st;
    }
    return first;
}

int mgw_hip_last_judge(const mgw_hip_ctx* ctx, double* err, int32_t* samples, int32_t* fell) {
    const auto* c = reinterpret_cast<const FakeCtx*>(ctx);
    if (err) *err = c->judge_err;
    if (samples) *samples = c->judge_err >= 0 ? c->judge_n : 0;
    if (fell) *fell = c->fell_back;
    return MGW_OK;
}

size_t mgw_hip_estimate_bytes(int32_t nmo, int32_t nocc, int32_t nfreq, int32_t, int32_t nbatch,
                              int32_t) {
    if (nmo <= 1 || nocc <= 0 || nocc >= nmo || nfreq <= 0) return 0;
    return eri_bytes(nmo) + (size_t)nbatch * 4096;
}

}  // extern "C"

// ------------------------------------------------------------------ scenarios
namespace {

struct Job {
    int nmo, nocc, nfreq, precision;
    std::vector<double> mo, vxc, b, qp, sre;
    int64_t ticket = 0;
};

Job make_job(int nmo, int nfreq, int precision = MGW_PREC_FP64) {
    Job j;
    j.nmo = nmo;
    j.nocc = nmo / 5 > 0 ? nmo / 5 : 1;
    j.nfreq = nfreq;
    j.precision = precision;
    j.mo.assign(nmo, -0.5);
    j.vxc.assign(nmo, 0.1);
    j.b.assign((size_t)nmo * nmo * 3, 0.01);
    j.qp.assign(nmo, std::nan(""));
    j.sre.assign((size_t)nmo * nfreq, 0.0);
    return j;
}

int submit(mgw_hip_server* s, Job& j) {
    mgw_serve_job sj{};
    sj.job.nmo = j.nmo;
    sj.job.nocc = j.nocc;
    sj.job.nfreq = j.nfreq;
    sj.job.npar = 16;
    sj.job.mo_energy = j.mo.data();
    sj.job.vxc_diag = j.vxc.data();
    sj.job.fermi_energy = 0.0;
    sj.job.precision = j.precision;
    sj.eri_factor = j.b.data();
    sj.kdim = 3;
    mgw_g0w0_out out{};
    out.qp = j.qp.data();
    out.sigma_c_re = j.sre.data();
    return mgw_hip_server_submit(s, &sj, &out, &j.ticket);
}

int wait_status(mgw_hip_server* s, const Job& j) {
    int32_t st = -1;
    double lat = 0, run = 0;
    CHECK(mgw_hip_server_wait(s, j.ticket, -1.0, &st, &lat, &run) == MGW_OK);
    CHECK(st >= 0 && lat >= run && run >= 0.0);
    return st;
}

void check_outputs(const Job& j, double extra = 0.0) {
    for (int p = 0; p < j.nmo; ++p) CHECK(j.qp[p] == 1000.0 * j.nmo + p + extra);
    for (double v : j.sre) CHECK(v == j.nfreq);
}

void scenario_mix() {
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 4, (size_t)1 << 40, 16, &s) == MGW_OK);
    std::vector<Job> jobs;
    for (int i = 0; i < 24; ++i) jobs.push_back(make_job(8 + (i * 7) % 30, 40 + i, i % 3));
    for (auto& j : jobs) CHECK(submit(s, j) == MGW_OK);
    for (auto& j : jobs) {
        CHECK(wait_status(s, j) == MGW_OK);
        check_outputs(j);
        double err = 0;
        int32_t fell = -1;
        CHECK(mgw_hip_server_job_judge(s, j.ticket, &err, &fell) == MGW_OK);
        const bool mixed = j.precision != MGW_PREC_FP64;
        CHECK(fell == 0 && (mixed ? err == 1e-9 : err == -1.0));
        CHECK(mgw_hip_server_release(s, j.ticket) == MGW_OK);
        CHECK(mgw_hip_server_release(s, j.ticket) == MGW_ERR_ARG);     // already forgotten
    }
    int64_t done = 0;
    int32_t running = -1, peak = 0;
    size_t reserved = 1;
    CHECK(mgw_hip_server_stats(s, &done, &running, &peak, &reserved) == MGW_OK);
    CHECK(done == 24 && running == 0 && reserved == 0 && peak >= 2);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);
}

void scenario_admission_and_faults() {
    // the budget admits one nmo=20 job at a time (ERI 1.28 MB)
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 3, 2000000, 8, &s) == MGW_OK);
    g_peak = 0;
    std::vector<Job> jobs;
    for (int i = 0; i < 5; ++i) jobs.push_back(make_job(20, 24));
    for (auto& j : jobs) CHECK(submit(s, j) == MGW_OK);
    for (auto& j : jobs) CHECK(wait_status(s, j) == MGW_OK);
    CHECK(g_peak.load() == 1);
    Job big = make_job(40, 24);                          // 20 MB: larger than the budget
    CHECK(submit(s, big) == MGW_OK);
    CHECK(wait_status(s, big) == MGW_ERR_MEMORY);
    CHECK(std::strstr(mgw_hip_server_job_error(s, big.ticket), "exceeds") != nullptr);
    // injected HBM shortfall: the job's own allocation fails, its reservation is released
    CHECK(mgw_hip_server_set_budget_scale(s, 0.01) == MGW_OK);
    Job cut = make_job(20, 24);
    CHECK(submit(s, cut) == MGW_OK);
    CHECK(wait_status(s, cut) == MGW_ERR_MEMORY);
    CHECK(mgw_hip_server_set_budget_scale(s, 1.0) == MGW_OK);
    CHECK(mgw_hip_server_set_budget_scale(s, 0.0) == MGW_ERR_ARG);
    Job ok = make_job(20, 24);
    CHECK(submit(s, ok) == MGW_OK);
    CHECK(wait_status(s, ok) == MGW_OK);
    check_outputs(ok);
    size_t reserved = 1;
    CHECK(mgw_hip_server_stats(s, nullptr, nullptr, nullptr, &reserved) == MGW_OK && reserved == 0);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);
}

// ERI storage: the default keeps every job resident; under MGW_ERI_AUTO a
// job that does not fit resident beside a running one starts streamed; a job
// too large even streamed fails; a streamed server streams every job.
void scenario_eri_storage() {
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 3, 2000000, 8, &s) == MGW_OK);
    CHECK(mgw_hip_server_set_eri_storage(s, 7) == MGW_ERR_ARG);
    Job first = make_job(20, 24);
    CHECK(submit(s, first) == MGW_OK);
    CHECK(wait_status(s, first) == MGW_OK);
    int32_t st = -1;
    CHECK(mgw_hip_server_job_eri_storage(s, first.ticket, &st) == MGW_OK && st == MGW_ERI_RESIDENT);
    CHECK(mgw_hip_server_set_eri_storage(s, MGW_ERI_AUTO) == MGW_OK);
    g_peak = 0;
    std::vector<Job> jobs;
    for (int i = 0; i < 6; ++i) jobs.push_back(make_job(20, 400));     // 30 ms each
    for (auto& j : jobs) CHECK(submit(s, j) == MGW_OK);
    int resident = 0, streamed = 0;
    for (auto& j : jobs) {
        CHECK(wait_status(s, j) == MGW_OK);
        check_outputs(j);
        CHECK(mgw_hip_server_job_eri_storage(s, j.ticket, &st) == MGW_OK);
        resident += st == MGW_ERI_RESIDENT;
        streamed += st == MGW_ERI_STREAMED;
    }
    // one resident job (1.4 MB) leaves room for one streamed job (0.4 MB) only
    CHECK(resident + streamed == 6 && resident >= 1 && streamed >= 1 && g_peak.load() == 2);
    Job big = make_job(40, 24);                      // 3.8 MB even streamed
    CHECK(submit(s, big) == MGW_OK);
    CHECK(wait_status(s, big) == MGW_ERR_MEMORY);
    CHECK(mgw_hip_server_set_eri_storage(s, MGW_ERI_STREAMED) == MGW_OK);
    Job one = make_job(20, 24);
    CHECK(submit(s, one) == MGW_OK);
    CHECK(wait_status(s, one) == MGW_OK);
    CHECK(mgw_hip_server_job_eri_storage(s, one.ticket, &st) == MGW_OK && st == MGW_ERI_STREAMED);
    size_t reserved = 1;
    CHECK(mgw_hip_server_stats(s, nullptr, nullptr, nullptr, &reserved) == MGW_OK && reserved == 0);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);
}

void scenario_cancel_and_fallback() {
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 1, (size_t)1 << 40, 4, &s) == MGW_OK);
    Job longjob = make_job(64, 400);                     // 100 batches of 4 frequencies
    Job queued = make_job(10, 8);
    CHECK(submit(s, longjob) == MGW_OK);
    CHECK(submit(s, queued) == MGW_OK);
    CHECK(mgw_hip_server_cancel(s, queued.ticket) == MGW_OK);
    CHECK(wait_status(s, queued) == MGW_ERR_STATE);
    CHECK(std::strstr(mgw_hip_server_job_error(s, queued.ticket), "before the job started"));
    int32_t running = 0;
    while (running == 0) mgw_hip_server_stats(s, nullptr, &running, nullptr, nullptr);
    CHECK(mgw_hip_server_cancel(s, longjob.ticket) == MGW_OK);
    CHECK(wait_status(s, longjob) == MGW_ERR_STATE);
    CHECK(std::strstr(mgw_hip_server_job_error(s, longjob.ticket), "cancelled"));
    CHECK(mgw_hip_server_cancel(s, longjob.ticket) == MGW_OK);      // finished: unchanged
    CHECK(mgw_hip_server_cancel(s, 999999) == MGW_ERR_ARG);
    // a judge tolerance below the form's error: the job reports its FP64 rerun
    CHECK(mgw_hip_server_set_judge(s, 2, 1e-12, 1) == MGW_OK);
    Job mixed = make_job(12, 16, MGW_PREC_MIXED_INT8);
    CHECK(submit(s, mixed) == MGW_OK);
    CHECK(wait_status(s, mixed) == MGW_OK);
    check_outputs(mixed, 0.5);
    int32_t fell = 0;
    CHECK(mgw_hip_server_job_judge(s, mixed.ticket, nullptr, &fell) == MGW_OK && fell == 1);
    CHECK(mgw_hip_server_set_judge(s, 65, 0.0, 1) == MGW_ERR_ARG);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);
}

void scenario_grouping() {
    // one worker: a long job runs while jobs of one padded order queue up;
    // they start together as one group, a cancelled member ends alone
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 1, (size_t)1 << 40, 8, &s) == MGW_OK);
    CHECK(mgw_hip_server_set_grouping(s, 0) == MGW_ERR_ARG);
    CHECK(mgw_hip_server_set_grouping(s, 4) == MGW_OK);
    g_groups = 0;
    g_group_peak = 0;
    Job first = make_job(64, 200);
    CHECK(submit(s, first) == MGW_OK);
    int32_t running = 0;
    while (running == 0) mgw_hip_server_stats(s, nullptr, &running, nullptr, nullptr);
    // nmo 16..20 with nocc = nmo / 5: particle-hole orders 39..64, all padded to 256
    std::vector<Job> jobs;
    for (int i = 0; i < 6; ++i) jobs.push_back(make_job(16 + (i % 5), 24 + i, MGW_PREC_MIXED_INT8_4));
    for (auto& j : jobs) CHECK(submit(s, j) == MGW_OK);
    CHECK(mgw_hip_server_cancel(s, jobs[1].ticket) == MGW_OK);   // pending: ends before it starts
    CHECK(wait_status(s, first) == MGW_OK);
    for (size_t i = 0; i < jobs.size(); ++i) {
        const int st = wait_status(s, jobs[i]);
        if (i == 1) {
            CHECK(st == MGW_ERR_STATE);
            continue;
        }
        CHECK(st == MGW_OK);
        check_outputs(jobs[i]);
    }
    CHECK(g_groups.load() >= 1 && g_group_peak.load() == 4);     // capped at 4 per group
    size_t reserved = 1;
    CHECK(mgw_hip_server_stats(s, nullptr, &running, nullptr, &reserved) == MGW_OK);
    CHECK(running == 0 && reserved == 0);
    // grouping off: every job alone
    CHECK(mgw_hip_server_set_grouping(s, 1) == MGW_OK);
    g_groups = 0;
    std::vector<Job> more;
    for (int i = 0; i < 4; ++i) more.push_back(make_job(16, 24));
    for (auto& j : more) CHECK(submit(s, j) == MGW_OK);
    for (auto& j : more) CHECK(wait_status(s, j) == MGW_OK);
    CHECK(g_groups.load() == 0);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);
}

void scenario_destroy_with_queue() {
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 1, (size_t)1 << 40, 8, &s) == MGW_OK);
    std::vector<Job> jobs;
    for (int i = 0; i < 6; ++i) jobs.push_back(make_job(30, 64));
    for (auto& j : jobs) CHECK(submit(s, j) == MGW_OK);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);          // queued jobs cancelled, records freed
}

void scenario_arguments() {
    mgw_hip_server* s = nullptr;
    CHECK(mgw_hip_server_create(0, 0, 1, 1, &s) == MGW_ERR_ARG);
    CHECK(mgw_hip_server_create(0, 1, 0, 1, &s) == MGW_ERR_ARG);
    CHECK(mgw_hip_server_create(0, 1, 1 << 20, 1, &s) == MGW_OK);
    Job bad = make_job(10, 8);
    bad.nocc = 10;                                      // nocc == nmo
    CHECK(submit(s, bad) == MGW_ERR_ARG);
    int32_t st = 0;
    CHECK(mgw_hip_server_wait(s, 12345, 0.0, &st, nullptr, nullptr) == MGW_ERR_ARG);
    CHECK(mgw_hip_server_job_judge(s, 12345, nullptr, nullptr) == MGW_ERR_ARG);
    CHECK(std::strcmp(mgw_hip_server_job_error(s, 12345), "unknown ticket") == 0);
    CHECK(mgw_hip_server_destroy(s) == MGW_OK);
    CHECK(mgw_hip_server_destroy(nullptr) == MGW_OK);
}

}  // namespace

int main() {
    scenario_mix();
    scenario_admission_and_faults();
    scenario_eri_storage();
    scenario_cancel_and_fallback();
    scenario_grouping();
    scenario_destroy_with_queue();
    scenario_arguments();
    CHECK(g_contexts.load() == 0);                      // every worker context destroyed
    std::printf("serve harness: all scenarios passed\n");
    return 0;
}
