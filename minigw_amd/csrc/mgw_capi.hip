// C ABI of the MI355X screening + Sigma_c backend (include/mgw_hip.h) and the
// host runtime behind it: device buffers, the frequency grid, batching of
// imaginary frequencies, and the hoisted G0W0 driver.
#include "mgw_hip.h"
#include "mgw_hip_diag.h"

#include "mgw_kernels.h"

#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_thread_error;

// Largest slab of whole ERI rows p that a streamed upload moves at once:
// MGW_ERI_SLAB_BYTES, default 512 MiB.
size_t env_slab_bytes() {
    if (const char* env = std::getenv("MGW_ERI_SLAB_BYTES"))
        return (size_t)std::max(1ll, std::atoll(env));
    return (size_t)512 << 20;
}

struct MgwError : std::runtime_error {
    int code;
    MgwError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw MgwError(code, buf);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(MGW_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
}

// Named host ranges for rocprofv3 --marker-trace, with the phase names of the
// reference's profiling table (src/profiling.cpp:12-32): "screening setup",
// "build W_c", "sigma_c", "sigma_c wall", "exchange", "pade/output pade".
// The ROCTx library is loaded at run time (rocprofiler-sdk's, else the
// roctracer one); without it the ranges cost a branch.  A range spans the
// host's enqueue of a phase and, where the driver waits (after each frequency
// batch and at the end of a molecule), the device work too.
struct Roctx {
    int (*push)(const char*) = nullptr;
    int (*pop)() = nullptr;
    Roctx() {
        for (const char* lib : {"librocprofiler-sdk-roctx.so.1", "libroctx64.so.4",
                                "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"}) {
            void* h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
            if (!h) continue;
            push = reinterpret_cast<int (*)(const char*)>(dlsym(h, "roctxRangePushA"));
            pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
            if (push && pop) return;
            push = nullptr;
            pop = nullptr;
        }
    }
};

Roctx& roctx() {
    static Roctx r;
    return r;
}

struct Range {
    explicit Range(const char* name) {
        if (roctx().push) roctx().push(name);
    }
    ~Range() {
        if (roctx().pop) roctx().pop();
    }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

}  // namespace

// ------------------------------------------------------- fixed device scratch
// Scratch of mgw_hip_quadratic_forms_noalloc: the panel's ERI row offsets, the
// packed panel X (np x mpad doubles), the contraction's per-row-tile partial
// sums and the batch offset, for np <= kNoAllocMaxNph and width <= 256.  Each
// device holds one copy for the life of the process; host code uses only the
// device address that hipGetSymbolAddress returned for it (noalloc_base).
constexpr int kNoAllocMaxNph = 8192, kNoAllocMaxWidth = 256;
constexpr size_t kNoAllocPk = (size_t)kNoAllocMaxWidth * 8;
constexpr size_t kNoAllocX = (size_t)kNoAllocMaxNph * kNoAllocMaxWidth * 8;
constexpr size_t kNoAllocQpart = (size_t)(kNoAllocMaxNph / mgw::kCT) * kNoAllocMaxWidth * 8;
constexpr size_t kNoAllocBytes = kNoAllocPk + kNoAllocX + kNoAllocQpart + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch;        // device -> address of its scratch

// The scratch array's address on the current device (looked up once per device).
void* noalloc_base(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return it->second;
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[device] = ptr;
    return ptr;
}
}  // namespace

// ------------------------------------------------------------------ context
// Everything a captured frequency batch bakes into its kernel arguments: the
// shapes (true particle-hole count n of each molecule, whose rows past it get
// S = 0; padded and factored orders; column counts; batch and matrix offsets),
// the form and every buffer address, as one signature.
struct GraphKey {
    std::vector<uint64_t> sig;
    bool operator==(const GraphKey& o) const { return sig == o.sig; }
};

// most matrices (molecules x frequencies) one stacked batch of a group holds
constexpr int kMaxStack = 2048;

struct mgw_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the factorization runs parts of a frequency batch on `stream` and the
    // side streams (forked and joined with events): each part's chain of
    // small, latency-bound launches then runs beside the others' block products
    static constexpr int kSides = 3;
    hipStream_t side[kSides] = {nullptr, nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[kSides] = {nullptr, nullptr, nullptr};
    int factor_streams = 2;           // MGW_FACTOR_STREAMS: 1, 2 or 4 parts
    // most frequencies factored at once (MGW_FREQ_BATCH_CAP); at nmo = 200 a
    // batch of 100 takes 97 GB in the int8 form
    int batch_cap = 100;
    int split_order = 4096;           // MGW_FACTOR_SPLIT_ORDER: smallest order that splits
    std::string error;
    size_t budget = 0;     // 0 = no cap
    size_t held = 0;
    double timings[15] = {0};

    // grow-only device buffer with budget accounting
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    std::vector<Buf*> all;

    // Stream-ordered allocation from the device's default memory pool
    // (MGW_ASYNC_ALLOC=1; off by default): a free does not wait for the whole
    // device as a plain hipFree does, which stalls the other workers of the
    // request loop at each job's end.  Off by default because a context then
    // keeps device memory past the release of its buffers until it is
    // destroyed (2 GiB after an nph = 4096 job, outside the pool's own
    // accounting), which the released-memory test catches.
    static bool async_alloc() {
        static const bool on = [] {
            const char* env = std::getenv("MGW_ASYNC_ALLOC");
            return env != nullptr && env[0] == '1';
        }();
        return on;
    }
    void* get(Buf& b, size_t bytes, const char* name) {
        if (bytes <= b.bytes && b.p) return b.p;
        release(b);
        if (bytes == 0) return nullptr;
        if (budget && held + bytes > budget)
            fail(MGW_ERR_MEMORY, "%s: %zu bytes would exceed the device budget (%zu held of %zu)",
                 name, bytes, held, budget);
        void* p = nullptr;
        const hipError_t e = async_alloc() && stream ? hipMallocAsync(&p, bytes, stream)
                                                     : hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(MGW_ERR_MEMORY, "hipMalloc(%s, %zu bytes) failed: %s", name, bytes,
                 hipGetErrorString(e));
        }
        b.p = p;
        b.bytes = bytes;
        held += bytes;
        return p;
    }
    void release(Buf& b) {
        if (b.p && !graphs.empty()) drop_graphs();   // a freed address may be handed out again
        if (b.p) {
            if (async_alloc() && stream)
                (void)hipFreeAsync(b.p, stream);
            else
                (void)hipFree(b.p);
            held -= b.bytes;
        }
        b.p = nullptr;
        b.bytes = 0;
    }

    // resident inputs
    Buf eri, v, colidx, rowbase, delta;
    int nmo = 0, nocc = 0;        // of the resident ERI (0 = none)
    bool eri_pk_sym = false;      // (pk|ia) == (kp|ia) exactly for the resident ERI
    // ERI storage: requested for the next upload, taken by the held one
    // (MGW_ERI_RESIDENT: eri = nmo^4 tensor; MGW_ERI_STREAMED: eri = the
    // (pk|ia) block, rows of nph values, and exsd = (pi|pi)), values per row
    // (p, k) of what eri holds
    int eri_req = MGW_ERI_AUTO;
    int eri_storage = 0;
    int64_t eri_rs = 0;
    Buf exsd, eslab;
    size_t slab_bytes = 0;            // largest streamed slab (env_slab_bytes at creation)
    Buf flag, hcolmap, wxp, hxp, f0buf;

    // frequency batch forced by the autotuner while it times candidates (0 = auto)
    int forced_nb = 0;
    // caller's cap on the frequencies factored at once (mgw_hip_set_batch_limit, 0 = none)
    int max_nb = 0;
    // caller's stop flag, polled between frequency batches (mgw_hip_set_cancel_flag)
    const int32_t* cancel = nullptr;
    // frequencies per 32-tile unit of the mixed contractions: MGW_SPLIT_UNIT_FREQS
    // (0 = the autotuned value, else 4); forced_uf while the autotuner times one
    int split_uf = 0;
    int forced_uf = 0;

    // captured frequency-batch graphs (hipGraph), keyed by shape and buffers
    bool use_graphs = false;
    struct GraphEntry {
        GraphKey key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    hipGraphExec_t find_graph(const GraphKey& k) const {
        for (const auto& g : graphs)
            if (g.key == k) return g.exec;
        return nullptr;
    }
    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    int n = 0, np = 0;            // ph dimension of the resident V
    bool v_ready = false;
    // single-frequency workspace (compute_wc / quadratic forms)
    Buf wA, wS, wdinv, wtbuf, winfo, wpkrow, wqpart, wq, wsrc, wcol;
    bool wc_ready = false;
    // hoisted path
    Buf hA, hS, hdinv, htbuf, hinfo, hpkrow, hqpart, hq, homega, hweight, hmo, hsig_re, hsig_im,
        hsx, hvxc, hstates, hqp, hiters;
    Buf xhi, xlo, xscale, mhi, mlo, mscale;
    Buf hqim, hY, wY, wW;             // complex path (eta != 0)
    Buf hfr, hfc;                     // factor block slices (mixed forms)
    Buf hinfob;                       // breakdown flags of one frequency batch
    // Cholesky-decomposed ERI form: pair tables, decomposition workspace,
    // factor columns, and B^T / pair vectors / their norms
    Buf cdidx, cdwork, cdl, cdb, cdx, cdn;
    Buf hsg;                          // Sigma_c quadrature factors g(n, f, k)
    Buf hpade;                        // Pade tables beyond kMaxPade points
    int eri_pair_sym = -1;            // symmetric to 1e-10 for the decomposed form: -1 unchecked
    double eri_asym = 0.0, eri_absmax = 0.0;
    double cd_tol = 1e-12;            // stop when the residual diagonal <= cd_tol * first pivot
    bool int8_factor = true;          // mixed forms: block products on int8 MFMA (MGW_INT8_FACTOR=0: FP64)
    bool wc_complex = false;
    // error judge of the mixed forms (mgw_hip_set_judge) and its last outcome
    int judge_n = 0;                  // sampled frequencies (0 = off)
    double judge_tol = 0.0;           // <= 0: the form's default
    bool judge_fallback = false;      // rerun in FP64 when the judge fails
    double judge_err = -1.0;          // largest relative q error of the last judged run (-1: none)
    int judge_used = 0;               // frequencies the last run judged
    int fell_back = 0;                // the last mgw_hip_g0w0 reran in FP64 (1: judge, 2: numeric)
    Buf hqj, hjom, hjinfo;            // the judge's q rows, sampled grid and breakdown flags
    // pivoted-LU screening of a dielectric matrix that is not positive definite
    Buf luz, ludl, ludu, lupiv, wz;   // W = E^-1 (batch path / workspace), block inverses, pivots
    bool wc_dense = false;            // the workspace's W_c is held as dense E^-1 (wz)
    int lu_freqs = 0;                 // frequencies of the last mgw_hip_g0w0 screened by the LU
    Buf ppk, panel;                   // the last filled device panel and its row offsets

    mgw_hip_ctx() {
        all = {&flag, &hcolmap, &wxp, &hxp, &f0buf, &eri, &v, &colidx, &rowbase, &delta, &wA, &wS, &wdinv, &wtbuf, &winfo, &wpkrow,
               &wqpart, &wq, &wsrc, &wcol, &hA, &hS, &hdinv, &htbuf, &hinfo, &hpkrow, &hqpart,
               &hq, &homega, &hweight, &hmo, &hsig_re, &hsig_im, &hsx, &hvxc, &hstates, &hqp,
               &hiters, &xhi, &xlo, &xscale, &mhi, &mlo, &mscale, &hqim, &hY, &wY, &wW, &hfr, &hfc,
               &cdidx, &cdwork, &cdl, &cdb, &cdx, &cdn, &hsg, &hqj, &luz, &ludl, &ludu,
               &lupiv, &wz, &ppk, &panel, &hinfob, &hjom, &hjinfo, &exsd, &eslab, &hpade};
    }
    ~mgw_hip_ctx() {
        (void)hipSetDevice(device);
        drop_graphs();
        for (Buf* b : all) release(*b);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (int i = 0; i < kSides; ++i) {
            if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
            if (side[i]) (void)hipStreamDestroy(side[i]);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
    void sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }
    template <class T>
    void h2d(void* dst, const T* src, size_t count) {
        hip_check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync H2D");
    }
    template <class T>
    void d2h(T* dst, const void* src, size_t count) {
        hip_check(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync D2H");
    }
};

namespace {

template <class F>
int guarded(mgw_hip_ctx* ctx, F&& body) {
    try {
        if (ctx) hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        body();
        if (ctx) ctx->error.clear();
        return MGW_OK;
    } catch (const MgwError& e) {
        (ctx ? ctx->error : g_thread_error) = e.what();
        if (ctx) g_thread_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        (ctx ? ctx->error : g_thread_error) = "host allocation failed";
        return MGW_ERR_MEMORY;
    } catch (const std::exception& e) {
        (ctx ? ctx->error : g_thread_error) = e.what();
        return MGW_ERR_HIP;
    }
}

void need(bool cond, const char* fmt, ...) {
    if (cond) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw MgwError(MGW_ERR_ARG, buf);
}

template <class T>
T* as(void* p) {
    return static_cast<T*>(p);
}

constexpr double kPi = 3.141592653589793238462643383279502884;

// Numerical-Recipes Gauss-Legendre (src/gw/frequency_grids.cpp:12-44) and the
// (1+y)/(2(1-y)) transform, sorted ascending, first node clamped >= 1e-12
// (src/gw/frequency_grids.cpp:62-93).
void transformed_legendre(int nfreq, double* omega, double* weight) {
    const int n = nfreq;
    std::vector<double> x(n), w(n);
    const double eps = 3e-14;
    const int m = (n + 1) / 2;
    const double xm = 0.0, xl = 1.0;
    for (int i = 0; i < m; ++i) {
        double z = std::cos(kPi * ((double)(i + 1) - 0.25) / ((double)n + 0.5));
        double pp = 0.0;
        for (int iter = 0; iter < 100; ++iter) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * (double)j - 1.0) * z * p2 - ((double)j - 1.0) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            const double z_old = z;
            z = z_old - p1 / pp;
            if (std::fabs(z - z_old) <= eps) break;
        }
        x[i] = xm - xl * z;
        x[n - 1 - i] = xm + xl * z;
        w[i] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - 1 - i] = w[i];
    }
    std::vector<double> tv(n), tw(n);
    for (int i = 0; i < n; ++i) {
        tv[i] = 0.5 * (1.0 + x[i]) / (1.0 - x[i]);
        tw[i] = w[i] / ((1.0 - x[i]) * (1.0 - x[i]));
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tv[a] < tv[b]; });
    for (int i = 0; i < n; ++i) {
        omega[i] = tv[order[i]];
        weight[i] = tw[order[i]];
    }
    if (n > 0 && omega[0] < 1e-12) omega[0] = 1e-12;
}

struct PhLayout {
    int nmo, nocc, nvir, n, np;
    std::vector<int> colidx;        // np: i*nmo + nocc + a (0 in the padding)
    std::vector<int64_t> rowbase;   // np: (j*nmo + nocc + b) * nmo^2 (0 in the padding)
};

// values per (p, k) row of the ERI as a context holds it
int64_t eri_row_values(int nmo, int nocc, int storage) {
    return storage == MGW_ERI_STREAMED ? (int64_t)nocc * (nmo - nocc) : (int64_t)nmo * nmo;
}

PhLayout ph_layout(int nmo, int nocc, int storage = MGW_ERI_RESIDENT) {
    PhLayout L;
    L.nmo = nmo;
    L.nocc = nocc;
    L.nvir = nmo - nocc;
    L.n = nocc * L.nvir;
    L.np = mgw::round_up(std::max(L.n, 1), mgw::kPad);
    L.colidx.assign(L.np, 0);
    L.rowbase.assign(L.np, 0);
    // the streamed storage holds only the ph columns, in ph order
    const bool held_ph = storage == MGW_ERI_STREAMED;
    const int64_t rs = eri_row_values(nmo, nocc, storage);
    for (int i = 0; i < nocc; ++i)
        for (int a = 0; a < L.nvir; ++a) {
            const int ph = i * L.nvir + a;
            L.colidx[ph] = held_ph ? ph : i * nmo + nocc + a;
            L.rowbase[ph] = ((int64_t)i * nmo + nocc + a) * rs;
        }
    return L;
}

// the forms whose contraction runs on int8 MFMA, and their digit scheme
bool int8_form(int precision) {
    return precision == MGW_PREC_MIXED_INT8 || precision == MGW_PREC_MIXED_INT8_4;
}
// the forms whose factorization's block products run on int8 MFMA
bool mixed_form(int precision) { return precision == MGW_PREC_MIXED || int8_form(precision); }
int digit_scheme(int precision) {
    return precision == MGW_PREC_MIXED_INT8_4 ? mgw::kDigits4x8 : mgw::kDigits5x7;
}

size_t batch_bytes(int np, int mpad, int precision, bool cplx = false) {
    const size_t nblk = np / mgw::kNB;
    if (cplx)                                // E/L, S, dinv, Y/Z complex; qpart unused
        return (size_t)np * np * 16 + (size_t)np * 16 + nblk * 64 * 64 * 16 +
               (size_t)np * mpad * 16 + 4;
    size_t b = (size_t)np * np * 8           // A
               + (size_t)np * 8              // S
               + nblk * 64 * 64 * 8          // dinv
               + (size_t)np * 64 * 8         // tbuf
               + (size_t)(np / mgw::kCT) * mpad * 8  // qpart
               + 4;                          // info
    if (precision == MGW_PREC_MIXED) b += (size_t)np * np * 4 + (size_t)np * 8;  // M planes
    if (int8_form(precision))
        b += (size_t)np * np * mgw::i8_digits(digit_scheme(precision)) + (size_t)np * 8;
    if (mixed_form(precision))               // factor slices, 2nd trtri slot
        b += 2 * mgw::fact_slice_bytes(np, digit_scheme(precision)) + (size_t)np * 64 * 8;
    return b;
}

// Factor + invert nb symmetrised dielectric matrices already in A (lower).
// With slice buffers (mixed forms) the block products run on int8 MFMA.
// Blocked Cholesky and inverse of nb matrices (frequencies) of order np; with
// slice buffers, the block products on int8 MFMA.  Every buffer holds one
// record per frequency at a fixed stride (the inverse's workspace tbuf one per
// slot, 2 slots with int8 products), so the batch splits into parts of
// consecutive frequencies that run on their own streams.
void factor_batch(mgw_hip_ctx* ctx, double* A, double* dinv, int* info, double* tbuf, int np,
                  int nb, char* rslices = nullptr, char* cslices = nullptr,
                  int scheme = mgw::kDigits5x7) {
    const size_t nblk = (size_t)(np / mgw::kNB);
    const bool i8 = rslices && cslices;
    const size_t digits = (size_t)mgw::i8_digits(scheme);
    const size_t planes = (size_t)nb * nblk * nblk * digits * 4096;
    auto run = [&](int f0, int count, hipStream_t s) {
        double* Af = A + (size_t)f0 * np * np;
        double* Df = dinv + (size_t)f0 * nblk * 4096;
        int* If = info + f0;
        double* Tf = tbuf + (size_t)f0 * (i8 ? 2 : 1) * np * mgw::kNB;
        if (i8) {
            int8_t* RP = reinterpret_cast<int8_t*>(rslices) + (size_t)f0 * nblk * nblk * digits * 4096;
            int8_t* CP = reinterpret_cast<int8_t*>(cslices) + (size_t)f0 * nblk * nblk * digits * 4096;
            // scale exponents: one int per row of each 64 x 64 block
            double* RS = reinterpret_cast<double*>(rslices + planes + (size_t)f0 * nblk * nblk * 64 * 4);
            double* CS = reinterpret_cast<double*>(cslices + planes + (size_t)f0 * nblk * nblk * 64 * 4);
            hip_check(mgw::launch_potrf_i8(Af, Df, If, RP, RS, CP, CS, np, count, scheme, s),
                      "potrf (int8 products) launch");
            hip_check(mgw::launch_trtri_i8(Af, Df, Tf, RP, RS, CP, CS, np, count, scheme, s),
                      "trtri (int8 products) launch");
        } else {
            hip_check(mgw::launch_potrf(Af, Df, If, np, count, s), "potrf launch");
            hip_check(mgw::launch_trtri(Af, Df, Tf, np, count, s), "trtri launch");
        }
    };
    // small orders gain nothing from the split (and the request loop already
    // runs several jobs at once): one stream below split_order
    const int parts = np >= ctx->split_order ? std::min(nb, ctx->factor_streams) : 1;
    if (parts < 2) {
        run(0, nb, ctx->stream);
        return;
    }
    hip_check(hipEventRecord(ctx->fork_ev, ctx->stream), "hipEventRecord");
    for (int i = 1; i < parts; ++i)
        hip_check(hipStreamWaitEvent(ctx->side[i - 1], ctx->fork_ev, 0), "hipStreamWaitEvent");
    for (int i = 0; i < parts; ++i) {
        const int f0 = (int)((long)nb * i / parts), f1 = (int)((long)nb * (i + 1) / parts);
        run(f0, f1 - f0, i == 0 ? ctx->stream : ctx->side[i - 1]);
    }
    for (int i = 1; i < parts; ++i) {
        hip_check(hipEventRecord(ctx->join_ev[i - 1], ctx->side[i - 1]), "hipEventRecord");
        hip_check(hipStreamWaitEvent(ctx->stream, ctx->join_ev[i - 1], 0), "hipStreamWaitEvent");
    }
}

void check_info(mgw_hip_ctx* ctx, const int* dinfo, int count, const char* what) {
    std::vector<int> info(count);
    ctx->d2h(info.data(), dinfo, count);
    ctx->sync();
    for (int i = 0; i < count; ++i)
        if (info[i] != 0)
            fail(MGW_ERR_NUMERIC,
                 "%s: symmetrised dielectric matrix %d is not positive definite (pivot %d)", what, i,
                 info[i]);
}

// E = I + S V S of `count` matrices (S: count x np, V: np x np) factored by
// pivoted LU and inverted into W (count x np x np) -- the screening of a
// dielectric matrix the Cholesky path cannot factor.  A is the matrices'
// workspace (count x np x np, overwritten).  A singular E is MGW_ERR_NUMERIC.
void lu_screen(mgw_hip_ctx* ctx, const double* V, const double* S, double* A, double* W, int np,
               int count) {
    hipStream_t st = ctx->stream;
    const size_t nblk = np / mgw::kNB;
    double* dl = as<double>(ctx->get(ctx->ludl, (size_t)count * nblk * 4096 * 8, "LU block inverses"));
    double* du = as<double>(ctx->get(ctx->ludu, (size_t)count * nblk * 4096 * 8, "LU block inverses"));
    int* piv = as<int>(ctx->get(ctx->lupiv, ((size_t)2 * count * np + count) * 4, "LU pivots"));
    int* perm = piv + (size_t)count * np;
    int* info = perm + (size_t)count * np;
    std::vector<int> zero(count, 0);
    ctx->h2d(info, zero.data(), count);
    hip_check(mgw::launch_dielectric_full(V, S, A, np, count, st), "dielectric (full)");
    hip_check(mgw::launch_lu_factor(A, piv, info, dl, du, np, count, st), "LU");
    std::vector<int> hp((size_t)count * np), hi(count);
    ctx->d2h(hp.data(), piv, hp.size());
    ctx->d2h(hi.data(), info, count);
    ctx->sync();
    for (int w = 0; w < count; ++w)
        if (hi[w] != 0)
            fail(MGW_ERR_NUMERIC, "getrf: the dielectric matrix is singular (zero pivot %d)", hi[w]);
    // rows of P^T I: the interchanges applied in order to the identity's rows
    std::vector<int> hperm((size_t)count * np);
    for (int w = 0; w < count; ++w) {
        int* pm = hperm.data() + (size_t)w * np;
        std::iota(pm, pm + np, 0);
        for (int c = 0; c < np; ++c) std::swap(pm[c], pm[hp[(size_t)w * np + c]]);
    }
    ctx->h2d(perm, hperm.data(), hperm.size());
    hip_check(mgw::launch_lu_inverse(A, dl, du, perm, W, np, count, st), "LU inverse");
    ctx->sync();       // the copy reads hperm, which ends with this function
}

struct CdResult {
    int rank = 0, rpad = 0;
    double *bt = nullptr, *lx = nullptr, *nrm = nullptr;
};

// Pivoted Cholesky decomposition of the resident ERI over unique pairs
// (M[u][v] = (p_u q_u | p_v q_v), u = (p <= q)) on the device, stopped at
// residual diagonal <= cd_tol * first pivot; then B^T (rpad x np: the rows of
// the particle-hole pairs) and the pair vectors of the contraction columns
// (mpad x rpad) with their squared norms.  The factor grows by doubling when
// the rank passes its capacity; the host polls the rank every 64 pivots.
CdResult cd_decompose(mgw_hip_ctx* ctx, const PhLayout& L,
                      const std::vector<std::pair<int, int>>& cols, int mpad) {
    const int nmo = L.nmo, nocc = L.nocc, np = L.np, mu = (int)cols.size();
    const int N = nmo * (nmo + 1) / 2;
    hipStream_t st = ctx->stream;
    std::vector<int> uidx((size_t)nmo * nmo);
    std::vector<int> idx((size_t)N + np + mpad, 0);       // pq | phu | colu
    for (int p = 0, u = 0; p < nmo; ++p)
        for (int q = p; q < nmo; ++q, ++u) {
            idx[u] = p * nmo + q;
            uidx[(size_t)p * nmo + q] = uidx[(size_t)q * nmo + p] = u;
        }
    for (int i = 0; i < nocc; ++i)
        for (int a = 0; a < L.nvir; ++a)
            idx[(size_t)N + i * L.nvir + a] = uidx[(size_t)i * nmo + nocc + a];
    for (int c = 0; c < mu; ++c)
        idx[(size_t)N + np + c] = uidx[(size_t)cols[c].first * nmo + cols[c].second];
    int* didx = as<int>(ctx->get(ctx->cdidx, idx.size() * 4, "pair tables"));
    ctx->h2d(didx, idx.data(), idx.size());
    const int* dpq = didx;
    const double* eri = as<double>(ctx->eri.p);

    // The form uses one ordering of every pair: the ERI must be symmetric
    // under pq <-> rs and p <-> k (on the (pk|ia) columns) to 1e-10 of its
    // largest element (checked once per resident ERI).
    if (ctx->eri_pair_sym < 0) {
        unsigned long long* dout =
            as<unsigned long long>(ctx->get(ctx->cdwork, 3 * sizeof(unsigned long long), "flag"));
        hip_check(hipMemsetAsync(dout, 0, 3 * sizeof(unsigned long long), st), "hipMemsetAsync");
        hip_check(mgw::launch_cd_asymmetry(eri, dpq, nmo, nocc, N, dout, st), "symmetry check");
        double a[3];
        ctx->d2h(a, dout, 3);
        ctx->sync();
        ctx->eri_asym = std::max(a[0], a[1]);
        ctx->eri_pair_sym = ctx->eri_asym <= 1e-10 * a[2] ? 1 : 0;
        ctx->eri_absmax = a[2];
    }
    if (!ctx->eri_pair_sym)
        fail(MGW_ERR_UNSUPPORTED,
             "MGW_PREC_FP64_CD: the resident ERI is not symmetric under pq <-> rs / p <-> k "
             "(largest difference %g, largest element %g); use MGW_PREC_FP64",
             ctx->eri_asym, ctx->eri_absmax);

    const int nblk = mgw::cd_blocks(N), Npad = mgw::cd_npad(N);
    const size_t wbytes = (size_t)Npad * 8 + (size_t)2 * nblk * 16 + (size_t)N * 8 + 16 +
                          (size_t)N * 4 + 16;
    char* wk = as<char>(ctx->get(ctx->cdwork, wbytes, "decomposition workspace"));
    double* dD = reinterpret_cast<double*>(wk);
    double2* dbm = reinterpret_cast<double2*>(wk + (size_t)Npad * 8);
    double* dpval = reinterpret_cast<double*>(wk + (size_t)Npad * 8 + (size_t)2 * nblk * 16);
    double* dres = dpval + N;
    int* dpiv = reinterpret_cast<int*>(dres + 2);
    int* drank = dpiv + N;

    const size_t col_bytes = (size_t)Npad * 8;
    int rcap = std::min(N, mgw::round_up(std::max(256, 4 * nmo), 64));
    rcap = std::min(N, std::max<int>(rcap, (int)std::min<size_t>(ctx->cdl.bytes / col_bytes, N)));
    double* Lc = as<double>(ctx->get(ctx->cdl, (size_t)rcap * col_bytes, "decomposition factor"));
    hip_check(mgw::launch_cd_init(eri, dpq, nmo, N, dD, dbm, drank, st), "decomposition init");
    int k = 0, rank = -1;
    for (;;) {
        const int k1 = std::min(rcap, k + 64);
        hip_check(mgw::launch_cd_steps(eri, dpq, nmo, N, Lc, dD, dbm, dpiv, dpval, drank, k, k1,
                                       ctx->cd_tol, st),
                  "decomposition steps");
        ctx->d2h(&rank, drank, 1);
        ctx->sync();
        k = k1;
        if (rank >= 0) break;
        if (k >= N) {               // every pair is a pivot: full rank
            rank = N;
            break;
        }
        if (k >= rcap) {            // grow the factor, keeping its columns
            const int ncap = std::min(N, 2 * rcap);
            mgw_hip_ctx::Buf grown;
            double* Ln = as<double>(ctx->get(grown, (size_t)ncap * col_bytes, "decomposition factor"));
            std::swap(grown, ctx->cdl);
            hip_check(hipMemcpyAsync(Ln, Lc, (size_t)rcap * col_bytes, hipMemcpyDeviceToDevice, st),
                      "hipMemcpyAsync D2D");
            ctx->sync();
            ctx->release(grown);
            Lc = Ln;
            rcap = ncap;
        }
    }
    hip_check(mgw::launch_cd_residual(dD, N, dres, st), "decomposition residual");
    double res[2] = {0.0, 0.0}, p0 = 0.0;
    ctx->d2h(res, dres, 2);
    if (rank > 0) ctx->d2h(&p0, dpval, 1);
    ctx->sync();
    if (res[1] != 0.0)
        fail(MGW_ERR_NUMERIC, "ERI decomposition: %g non-finite residual diagonal entries", res[1]);
    if (res[0] < -1e-8 * p0)
        fail(MGW_ERR_UNSUPPORTED,
             "ERI decomposition: the pair matrix is not positive semidefinite (residual diagonal "
             "%g, first pivot %g); use MGW_PREC_FP64",
             res[0], p0);

    CdResult r;
    r.rank = rank;
    r.rpad = mgw::round_up(std::max(rank, 1), mgw::kCT);
    r.bt = as<double>(ctx->get(ctx->cdb, (size_t)r.rpad * np * 8, "B^T"));
    r.lx = as<double>(ctx->get(ctx->cdx, (size_t)mpad * r.rpad * 8, "pair vectors"));
    r.nrm = as<double>(ctx->get(ctx->cdn, (size_t)mpad * 8, "pair vector norms"));
    hip_check(mgw::launch_cd_gather(Lc, N, didx + N, didx + N + np, rank, L.n, mu, r.bt, r.lx,
                                    r.nrm, r.rpad, np, mpad, st),
              "gather pair vectors");
    return r;
}

}  // namespace

// Frequency-batch sizes picked by mgw_hip_autotune, per (nmo, nocc, nstates,
// precision), shared by all contexts of the process.
std::mutex g_tune_mu;
// autotuned frequency batch and (mixed forms) contraction unit shape
struct Tuned {
    int nb = 0;     // 0: not tuned
    int uf = 0;     // frequencies per 32-tile unit of the mixed contractions, 0: default
};
std::map<std::array<int, 4>, Tuned> g_tune;

Tuned tuned(int nmo, int nocc, int ns, int precision) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find({nmo, nocc, ns, precision});
    return it == g_tune.end() ? Tuned{} : it->second;
}

// ================================================================ C ABI
extern "C" {

const char* mgw_hip_version(void) { return "minigw-amd 0.1 (gfx950)"; }

const char* mgw_hip_last_error(const mgw_hip_ctx* ctx) {
    return ctx ? ctx->error.c_str() : g_thread_error.c_str();
}

int mgw_hip_device_count(int* count) {
    return guarded(nullptr, [&] {
        need(count != nullptr, "mgw_hip_device_count: count is NULL");
        hip_check(hipGetDeviceCount(count), "hipGetDeviceCount");
    });
}

int mgw_hip_device_for_local_rank(size_t local_rank, size_t tasks_per_gpu, int* device) {
    return guarded(nullptr, [&] {
        need(device != nullptr, "mgw_hip_device_for_local_rank: device is NULL");
        need(tasks_per_gpu > 0, "mgw_hip_device_for_local_rank: tasks_per_gpu must be positive");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (count <= 0) fail(MGW_ERR_HIP, "no HIP devices are visible");
        *device = (int)((local_rank / tasks_per_gpu) % (size_t)count);
    });
}

int mgw_hip_create(int device, mgw_hip_ctx** out) {
    return guarded(nullptr, [&] {
        need(out != nullptr, "mgw_hip_create: out is NULL");
        *out = nullptr;
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        need(device >= 0 && device < count, "mgw_hip_create: device %d out of range (%d visible)",
             device, count);
        hip_check(hipSetDevice(device), "hipSetDevice");
        (void)noalloc_base(device);          // the device's scratch address, once per device
        auto* ctx = new mgw_hip_ctx();
        ctx->device = device;
        if (const char* env = std::getenv("MGW_INT8_FACTOR")) ctx->int8_factor = env[0] != '0';
        if (const char* env = std::getenv("MGW_SPLIT_UNIT_FREQS")) {
            const int uf = std::atoi(env);
            if (uf == 1 || uf == 2 || uf == 4 || uf == 8 || uf == 16 || uf == 32) ctx->split_uf = uf;
        }
        if (const char* env = std::getenv("MGW_FACTOR_STREAMS"))
            ctx->factor_streams = env[0] == '1' ? 1 : env[0] == '4' ? 4 : 2;
        if (const char* env = std::getenv("MGW_FACTOR_SPLIT_ORDER")) ctx->split_order = std::max(0, std::atoi(env));
        ctx->slab_bytes = env_slab_bytes();
        if (const char* env = std::getenv("MGW_FREQ_BATCH_CAP"))
            ctx->batch_cap = std::min(1024, std::max(1, std::atoi(env)));
        if (mgw_hip_ctx::async_alloc()) {
            // keep freed memory in the pool (the default threshold returns it at every sync)
            hipMemPool_t pool = nullptr;
            if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
                uint64_t keep = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
            (void)hipGetLastError();
        }
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming);
        // only the side streams the split uses: idle streams still take hardware queues,
        // which the contexts of the request loop's workers share
        for (int i = 0; i < ctx->factor_streams - 1 && e == hipSuccess; ++i) {
            e = hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            delete ctx;
            hip_check(e, "stream / event creation");
        }
        *out = ctx;
    });
}

int mgw_hip_destroy(mgw_hip_ctx* ctx) {
    if (!ctx) return MGW_OK;
    delete ctx;
    return MGW_OK;
}

int mgw_hip_device_bytes(const mgw_hip_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) {
        g_thread_error = "mgw_hip_device_bytes: NULL argument";
        return MGW_ERR_ARG;
    }
    *bytes = ctx->held;
    return MGW_OK;
}

int mgw_hip_set_memory_budget(mgw_hip_ctx* ctx, size_t bytes) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_memory_budget: ctx is NULL");
        ctx->budget = bytes;
    });
}

int mgw_hip_set_graphs(mgw_hip_ctx* ctx, int32_t enable) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_graphs: ctx is NULL");
        ctx->use_graphs = enable != 0;
        if (!ctx->use_graphs) ctx->drop_graphs();
    });
}

int mgw_hip_set_cd_tolerance(mgw_hip_ctx* ctx, double tol) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_cd_tolerance: ctx is NULL");
        need(tol >= 0.0 && tol < 1.0, "mgw_hip_set_cd_tolerance: tol must be in [0, 1)");
        ctx->cd_tol = tol;
    });
}

int mgw_hip_set_cancel_flag(mgw_hip_ctx* ctx, const int32_t* flag) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_cancel_flag: ctx is NULL");
        ctx->cancel = flag;
    });
}

int mgw_hip_set_batch_limit(mgw_hip_ctx* ctx, int32_t max_freqs) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_batch_limit: ctx is NULL");
        need(max_freqs >= 0, "mgw_hip_set_batch_limit: max_freqs must be >= 0");
        ctx->max_nb = max_freqs;
    });
}

int mgw_hip_set_judge(mgw_hip_ctx* ctx, int32_t nsample, double tol, int32_t fallback) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_judge: ctx is NULL");
        need(nsample >= 0 && nsample <= 64, "mgw_hip_set_judge: nsample must be in [0, 64]");
        need(tol >= 0.0 && std::isfinite(tol), "mgw_hip_set_judge: tol must be finite and >= 0");
        ctx->judge_n = nsample;
        ctx->judge_tol = tol;
        ctx->judge_fallback = fallback != 0;
    });
}

int mgw_hip_last_judge(const mgw_hip_ctx* ctx, double* max_rel_err, int32_t* samples,
                       int32_t* fell_back) {
    if (!ctx) {
        g_thread_error = "mgw_hip_last_judge: ctx is NULL";
        return MGW_ERR_ARG;
    }
    if (max_rel_err) *max_rel_err = ctx->judge_err;
    if (samples) *samples = ctx->judge_used;
    if (fell_back) *fell_back = ctx->fell_back;
    return MGW_OK;
}

int mgw_grid_transformed_legendre(int32_t nfreq, double* omega, double* weight) {
    return guarded(nullptr, [&] {
        need(nfreq > 0, "generate_transformed_legendre_grid: num_freq_points must be positive");
        need(omega && weight, "mgw_grid_transformed_legendre: NULL output");
        transformed_legendre(nfreq, omega, weight);
    });
}

// ---------------------------------------------------- DeviceScreeningWorkspace
int mgw_hip_upload_vph(mgw_hip_ctx* ctx, const double* v, int32_t n) {
    return guarded(ctx, [&] {
        need(ctx && v, "mgw_hip_upload_vph: NULL argument");
        need(n > 0, "DeviceScreeningWorkspace: V_ph must be non-empty and square");
        const int np = mgw::round_up(n, mgw::kPad);
        double* tmp = as<double>(ctx->get(ctx->wsrc, (size_t)n * n * 8, "V_ph staging"));
        double* dv = as<double>(ctx->get(ctx->v, (size_t)np * np * 8, "V_ph"));
        ctx->h2d(tmp, v, (size_t)n * n);
        hip_check(mgw::launch_pad_square(tmp, dv, n, np, ctx->stream), "pad V_ph");
        ctx->sync();
        ctx->release(ctx->wsrc);
        ctx->n = n;
        ctx->np = np;
        ctx->v_ready = true;
        ctx->wc_ready = false;
    });
}

int mgw_hip_build_vph_from_eri(mgw_hip_ctx* ctx) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_build_vph_from_eri: ctx is NULL");
        if (!ctx->nmo) fail(MGW_ERR_STATE, "mgw_hip_build_vph_from_eri: no resident ERI");
        const PhLayout L = ph_layout(ctx->nmo, ctx->nocc);
        double* dv = as<double>(ctx->get(ctx->v, (size_t)L.np * L.np * 8, "V_ph"));
        hip_check(mgw::launch_build_v(as<double>(ctx->eri.p), as<int64_t>(ctx->rowbase.p),
                                      as<int>(ctx->colidx.p), dv, L.n, L.np, ctx->stream),
                  "build V_ph");
        ctx->sync();
        ctx->n = L.n;
        ctx->np = L.np;
        ctx->v_ready = true;
        ctx->wc_ready = false;
    });
}

int mgw_hip_compute_wc(mgw_hip_ctx* ctx, const double* pi0_re, const double* pi0_im, int32_t n) {
    return guarded(ctx, [&] {
        need(ctx && pi0_re, "mgw_hip_compute_wc: NULL argument");
        if (!ctx->v_ready) fail(MGW_ERR_STATE, "compute_w_c before V_ph was uploaded");
        need(n == ctx->n, "DeviceScreeningWorkspace::compute_w_c: inconsistent pi0 dimension");
        bool cplx = false;
        for (int i = 0; i < n; ++i) {
            if (pi0_im && pi0_im[i] != 0.0) cplx = true;
            if (!(pi0_re[i] <= 0.0))
                fail(MGW_ERR_UNSUPPORTED,
                     "compute_w_c: Pi0[%d] = %g > 0 (non-positive particle-hole gap)", i, pi0_re[i]);
        }
        const int np = ctx->np;
        ctx->wc_ready = false;
        if (cplx) {
            // complex symmetric E = I + S V S, S = sqrt(-Pi0): L L^T and the
            // diagonal-block inverses stay resident for the quadratic forms
            double* A = as<double>(ctx->get(ctx->wA, (size_t)np * np * 16, "dielectric"));
            double* S = as<double>(ctx->get(ctx->wS, (size_t)np * 16 + (size_t)np * 16, "scale"));
            double* dinv = as<double>(ctx->get(ctx->wdinv, (size_t)np * 64 * 16, "dinv"));
            int* info = as<int>(ctx->get(ctx->winfo, sizeof(int), "info"));
            double* pre = S + 2 * np;
            double* pim = pre + np;
            ctx->h2d(pre, pi0_re, n);
            ctx->h2d(pim, pi0_im, n);
            hip_check(hipMemsetAsync(info, 0, sizeof(int), ctx->stream), "hipMemsetAsync");
            hip_check(mgw::launch_scale_pi0_c(pre, pim, S, n, np, ctx->stream), "scale (complex)");
            hip_check(mgw::launch_dielectric_c(as<double>(ctx->v.p), S, A, np, 1, ctx->stream),
                      "dielectric (complex)");
            hip_check(mgw::launch_potrf_c(A, dinv, info, np, 1, ctx->stream), "potrf (complex)");
            check_info(ctx, info, 1, "compute_w_c");
            ctx->wc_complex = true;
            ctx->wc_dense = false;
            ctx->wc_ready = true;
            return;
        }
        double* A = as<double>(ctx->get(ctx->wA, (size_t)np * np * 8, "dielectric"));
        double* S = as<double>(ctx->get(ctx->wS, (size_t)np * 8 * 2, "scale"));
        double* dinv = as<double>(ctx->get(ctx->wdinv, (size_t)np * 64 * 8, "dinv"));
        double* tbuf = as<double>(ctx->get(ctx->wtbuf, (size_t)np * 64 * 8, "trtri scratch"));
        int* info = as<int>(ctx->get(ctx->winfo, sizeof(int), "info"));
        double* pi0 = S + np;
        ctx->h2d(pi0, pi0_re, n);
        hip_check(hipMemsetAsync(info, 0, sizeof(int), ctx->stream), "hipMemsetAsync");
        hip_check(mgw::launch_scale_from_pi0(pi0, S, n, np, ctx->stream), "scale");
        hip_check(mgw::launch_dielectric(as<double>(ctx->v.p), S, A, np, 1, ctx->stream),
                  "dielectric");
        factor_batch(ctx, A, dinv, info, tbuf, np, 1);
        int bad = 0;
        ctx->d2h(&bad, info, 1);
        ctx->sync();
        ctx->wc_dense = bad != 0;
        if (bad) {      // not positive definite: pivoted LU, W = E^-1 kept dense
            double* W = as<double>(ctx->get(ctx->wz, (size_t)np * np * 8, "E^-1"));
            lu_screen(ctx, as<double>(ctx->v.p), S, A, W, np, 1);
        }
        ctx->wc_complex = false;
        ctx->wc_ready = true;
    });
}

int mgw_hip_download_wc(mgw_hip_ctx* ctx, double* wc_re, double* wc_im, int32_t n) {
    return guarded(ctx, [&] {
        need(ctx && wc_re, "mgw_hip_download_wc: NULL argument");
        if (!ctx->wc_ready) fail(MGW_ERR_STATE, "W_c requested before compute_w_c");
        need(n == ctx->n, "mgw_hip_download_wc: n = %d, workspace has %d", n, ctx->n);
        if (ctx->wc_complex) {
            const int np = ctx->np;
            double* wc = as<double>(ctx->get(ctx->wsrc, (size_t)n * n * 16, "W_c staging"));
            double* Y = as<double>(ctx->get(ctx->wY, (size_t)np * np * 16, "complex scratch"));
            double* W = as<double>(ctx->get(ctx->wW, (size_t)np * np * 16, "complex W_c"));
            hip_check(mgw::launch_wc_explicit_c(as<double>(ctx->wA.p), as<double>(ctx->wdinv.p),
                                                as<double>(ctx->wS.p), Y, W, wc, wc + (size_t)n * n, n,
                                                np, ctx->stream),
                      "W_c (complex)");
            ctx->d2h(wc_re, wc, (size_t)n * n);
            if (wc_im) ctx->d2h(wc_im, wc + (size_t)n * n, (size_t)n * n);
            ctx->sync();
            ctx->release(ctx->wY);
            ctx->release(ctx->wW);
            return;
        }
        double* wc = as<double>(ctx->get(ctx->wsrc, (size_t)n * n * 8, "W_c staging"));
        if (ctx->wc_dense)
            hip_check(mgw::launch_wc_from_dense(as<double>(ctx->wz.p), as<double>(ctx->wS.p), wc, n,
                                                ctx->np, ctx->stream),
                      "W_c (dense)");
        else
            hip_check(mgw::launch_wc_explicit(as<double>(ctx->wA.p), as<double>(ctx->wS.p), wc, n,
                                              ctx->np, ctx->stream),
                      "W_c");
        ctx->d2h(wc_re, wc, (size_t)n * n);
        ctx->sync();
        if (wc_im) std::fill(wc_im, wc_im + (size_t)n * n, 0.0);
    });
}

namespace {
// q for columns described by (src, pkrow[mpad], colidx[np]) with the resident M.
void contract_single(mgw_hip_ctx* ctx, const double* src, const int64_t* hpkrow, int m,
                     const int* hcolidx, double* q_re, double* q_im) {
    const int np = ctx->np;
    const int mpad = mgw::round_up(std::max(m, 1), mgw::kPad);
    std::vector<int64_t> pk(mpad, 0);
    std::copy(hpkrow, hpkrow + m, pk.begin());
    int64_t* dpk = as<int64_t>(ctx->get(ctx->wpkrow, (size_t)mpad * 8, "pkrow"));
    int* dcol = as<int>(ctx->get(ctx->wcol, (size_t)np * 4, "colidx"));
    double* qpart =
        as<double>(ctx->get(ctx->wqpart, (size_t)(np / mgw::kCT) * mpad * 8, "qpart"));
    double* q = as<double>(ctx->get(ctx->wq, (size_t)mpad * 8, "q"));
    ctx->h2d(dpk, pk.data(), mpad);
    ctx->h2d(dcol, hcolidx, np);
    double* xp = as<double>(ctx->get(ctx->wxp, (size_t)mpad * np * 8, "packed panel"));
    hip_check(mgw::launch_pack_x(src, dpk, dcol, xp, ctx->n, np, mpad, ctx->stream), "pack X");
    int* f0p = as<int>(ctx->get(ctx->f0buf, sizeof(int), "batch offset"));
    hip_check(mgw::launch_set_int(f0p, 0, ctx->stream), "set offset");
    if (ctx->wc_complex) {
        double* Y = as<double>(ctx->get(ctx->wY, (size_t)np * mpad * 16, "complex panel"));
        double* qi = as<double>(ctx->get(ctx->wW, (size_t)mpad * 8, "q imag"));
        hip_check(mgw::launch_make_y(as<double>(ctx->wS.p), xp, Y, np, mpad, 1, ctx->stream), "S X");
        hip_check(mgw::launch_fsub_c(as<double>(ctx->wA.p), as<double>(ctx->wdinv.p), Y, np, mpad, 1,
                                     ctx->stream),
                  "L^-1 S X");
        hip_check(mgw::launch_sumsq_c(Y, nullptr, q, qi, np, m, mpad, f0p, 1, ctx->stream), "q");
        ctx->d2h(q_re, q, m);
        if (q_im) ctx->d2h(q_im, qi, m);
        ctx->sync();
        return;
    }
    if (ctx->wc_dense)
        hip_check(mgw::launch_contract_dense(as<double>(ctx->wz.p), as<double>(ctx->wS.p), xp, qpart,
                                             np, mpad, 1, ctx->stream),
                  "contract (dense)");
    else
        hip_check(mgw::launch_contract_f64(as<double>(ctx->wA.p), as<double>(ctx->wS.p), xp, qpart,
                                           np, mpad, 1, ctx->stream),
                  "contract");
    hip_check(mgw::launch_reduce_q(qpart, nullptr, q, np / mgw::kCT, m, mpad, f0p, 1, ctx->stream),
              "reduce q");
    ctx->d2h(q_re, q, m);
    ctx->sync();
    if (q_im) std::fill(q_im, q_im + m, 0.0);
}
}  // namespace

int mgw_hip_quadratic_forms_host(mgw_hip_ctx* ctx, const double* x, int32_t n, int32_t nvec,
                                 double* q_re, double* q_im) {
    return guarded(ctx, [&] {
        need(ctx && x && q_re, "mgw_hip_quadratic_forms_host: NULL argument");
        if (!ctx->wc_ready) fail(MGW_ERR_STATE, "quadratic_forms_panel before compute_w_c");
        need(n == ctx->n && nvec >= 0,
             "DeviceScreeningWorkspace::quadratic_forms_panel: inconsistent dimensions");
        if (nvec == 0) return;
        // transpose the row-major n x nvec panel into columns of length np
        const int np = ctx->np;
        std::vector<double> xt((size_t)nvec * np, 0.0);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < nvec; ++c) xt[(size_t)c * np + r] = x[(size_t)r * nvec + c];
        double* dsrc = as<double>(ctx->get(ctx->wsrc, xt.size() * 8, "panel"));
        ctx->h2d(dsrc, xt.data(), xt.size());
        std::vector<int64_t> pk(nvec);
        for (int c = 0; c < nvec; ++c) pk[c] = (int64_t)c * np;
        std::vector<int> col(np);
        std::iota(col.begin(), col.end(), 0);
        contract_single(ctx, dsrc, pk.data(), nvec, col.data(), q_re, q_im);
    });
}

// ---------------------------------------------------- DevicePqPhPanelView
namespace {
// ERI bookkeeping shared by both upload paths: ph tables of the storage and
// the (pk|ia) == (kp|ia) check.
void finish_eri_upload(mgw_hip_ctx* ctx, int nmo, int nocc, int storage) {
    const PhLayout L = ph_layout(nmo, nocc, storage);
    int* col = as<int>(ctx->get(ctx->colidx, (size_t)L.np * 4, "colidx"));
    int64_t* rb = as<int64_t>(ctx->get(ctx->rowbase, (size_t)L.np * 8, "rowbase"));
    ctx->h2d(col, L.colidx.data(), L.np);
    ctx->h2d(rb, L.rowbase.data(), L.np);
    int* dflag = as<int>(ctx->get(ctx->flag, sizeof(int), "flag"));
    hip_check(hipMemsetAsync(dflag, 0, sizeof(int), ctx->stream), "hipMemsetAsync");
    const bool held_ph = storage == MGW_ERI_STREAMED;
    hip_check(mgw::launch_check_pk_symmetry(as<double>(ctx->eri.p), nmo, nocc,
                                            eri_row_values(nmo, nocc, storage),
                                            held_ph ? L.nvir : nmo, held_ph ? 0 : nocc, dflag,
                                            ctx->stream),
              "symmetry check");
    int asym = 1;
    ctx->d2h(&asym, dflag, 1);
    ctx->sync();
    ctx->eri_pk_sym = (asym == 0);
    ctx->eri_pair_sym = -1;
    ctx->eri_storage = storage;
    ctx->eri_rs = eri_row_values(nmo, nocc, storage);
    ctx->nmo = nmo;
    ctx->nocc = nocc;
}


size_t resident_eri_bytes(int nmo) { return (size_t)nmo * nmo * nmo * nmo * 8; }

// rows p per slab: as many as fit slab_bytes, at least one
int slab_rows(int nmo, size_t slab_bytes) {
    const size_t row = (size_t)nmo * nmo * nmo * 8;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)nmo, slab_bytes / row));
}

size_t streamed_eri_bytes(int nmo, int nocc, size_t slab_bytes) {
    const size_t held = ((size_t)nmo * nmo * nocc * (nmo - nocc) + (size_t)nmo * nocc) * 8;
    return held + (size_t)slab_rows(nmo, slab_bytes) * nmo * nmo * nmo * 8;
}

// Free device memory as the allocator sees it: the driver's free bytes plus
// what the memory pool holds unused (freed buffers it keeps for reuse).
hipError_t device_free_bytes(int device, size_t* free_b, size_t* total_b) {
    hipError_t e = hipMemGetInfo(free_b, total_b);
    if (e != hipSuccess || !mgw_hip_ctx::async_alloc()) return e;
    hipMemPool_t pool = nullptr;
    uint64_t reserved = 0, used = 0;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess &&
        hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) == hipSuccess &&
        hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess &&
        reserved > used)
        *free_b += (size_t)(reserved - used);
    (void)hipGetLastError();
    return hipSuccess;
}

// The storage an upload takes: the requested one, or for MGW_ERI_AUTO the
// reference's rule (src/workspace/device_pq_ph_panel.cu:168-190): resident
// only while the tensor takes less than a quarter of the free device memory
// (of the context's budget left, when that is smaller).  Under MGW_ERI_AUTO
// the context's previous ERI is released first so that it does not count
// against the new one; a requested storage keeps the held ERI buffer, which
// the upload reuses when it is large enough.
int choose_storage(mgw_hip_ctx* ctx, int nmo) {
    ctx->nmo = 0;
    ctx->eri_storage = 0;
    if (ctx->eri_req != MGW_ERI_AUTO) {
        if (ctx->eri_req == MGW_ERI_RESIDENT) ctx->release(ctx->exsd);   // streamed-only
        return ctx->eri_req;
    }
    ctx->release(ctx->eri);
    ctx->release(ctx->exsd);
    size_t free_b = 0, total_b = 0;
    if (device_free_bytes(ctx->device, &free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return MGW_ERI_STREAMED;
    }
    if (ctx->budget) free_b = std::min(free_b, ctx->budget > ctx->held ? ctx->budget - ctx->held : 0);
    return resident_eri_bytes(nmo) < free_b / 4 ? MGW_ERI_RESIDENT : MGW_ERI_STREAMED;
}

// Streamed storage: slabs of whole rows p, each copied from the host ERI or
// formed from the factor (db: Np + 64 rows of Kp, zero padded), gathered into
// the held (pk|ia) block and (pi|pi), one slab buffer reused in stream order.
void upload_streamed(mgw_hip_ctx* ctx, int nmo, int nocc, const double* host, const double* db,
                     int Kp) {
    const int n = nocc * (nmo - nocc), N = nmo * nmo;
    const size_t row = (size_t)nmo * nmo * nmo;
    const int rows = slab_rows(nmo, ctx->slab_bytes);
    double* eph = as<double>(ctx->get(ctx->eri, (size_t)N * n * 8, "ERI ph block"));
    double* sxd = as<double>(ctx->get(ctx->exsd, (size_t)nmo * nocc * 8, "exchange diagonal"));
    double* slab = as<double>(ctx->get(ctx->eslab, (size_t)rows * row * 8, "ERI slab"));
    for (int p0 = 0; p0 < nmo; p0 += rows) {
        const int r = std::min(rows, nmo - p0);
        if (host)
            ctx->h2d(slab, host + (size_t)p0 * row, (size_t)r * row);
        else
            hip_check(mgw::launch_eri_rows_from_factor(db, N, Kp, p0 * nmo, r * nmo, slab,
                                                       ctx->stream),
                      "ERI rows from factor");
        hip_check(mgw::launch_gather_ph(slab, p0, r, nmo, nocc, eph, sxd, ctx->stream),
                  "gather ph block");
    }
    ctx->sync();
    ctx->release(ctx->eslab);
}
}  // namespace

int mgw_hip_set_eri_storage(mgw_hip_ctx* ctx, int32_t storage) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_set_eri_storage: ctx is NULL");
        need(storage == MGW_ERI_AUTO || storage == MGW_ERI_RESIDENT || storage == MGW_ERI_STREAMED,
             "mgw_hip_set_eri_storage: unknown storage %d", storage);
        ctx->eri_req = storage;
    });
}

int mgw_hip_eri_storage(const mgw_hip_ctx* ctx, int32_t* storage, size_t* device_bytes) {
    if (!ctx || !storage) {
        g_thread_error = "mgw_hip_eri_storage: NULL argument";
        return MGW_ERR_ARG;
    }
    if (!ctx->nmo) {
        g_thread_error = "mgw_hip_eri_storage: no ERI uploaded";
        return MGW_ERR_STATE;
    }
    *storage = ctx->eri_storage;
    if (device_bytes) *device_bytes = ctx->eri.bytes + ctx->exsd.bytes;
    return MGW_OK;
}

size_t mgw_hip_eri_bytes(int32_t nmo, int32_t nocc, int32_t storage) {
    if (nmo <= 1 || nocc <= 0 || nocc >= nmo) return 0;
    if (storage == MGW_ERI_RESIDENT) return resident_eri_bytes(nmo);
    if (storage == MGW_ERI_STREAMED) return streamed_eri_bytes(nmo, nocc, env_slab_bytes());
    return 0;
}

int mgw_hip_upload_eri(mgw_hip_ctx* ctx, const double* eri, int32_t nmo, int32_t nocc) {
    return guarded(ctx, [&] {
        need(ctx && eri, "mgw_hip_upload_eri: NULL argument");
        need(nmo > 0, "MolecularIntegrals: ERI tensor must have shape [nmo,nmo,nmo,nmo]");
        need(nocc > 0 && nocc < nmo, "OrbitalSpace: invalid number of occupied orbitals");
        const int storage = choose_storage(ctx, nmo);
        if (storage == MGW_ERI_STREAMED) {
            upload_streamed(ctx, nmo, nocc, eri, nullptr, 0);
        } else {
            const size_t count = (size_t)nmo * nmo * nmo * nmo;
            double* d = as<double>(ctx->get(ctx->eri, count * 8, "ERI"));
            ctx->h2d(d, eri, count);
        }
        finish_eri_upload(ctx, nmo, nocc, storage);
    });
}

int mgw_hip_upload_eri_factor(mgw_hip_ctx* ctx, const double* b, int32_t nmo, int32_t nocc,
                              int32_t kdim) {
    return guarded(ctx, [&] {
        need(ctx && b, "mgw_hip_upload_eri_factor: NULL argument");
        need(nmo > 0 && kdim > 0, "mgw_hip_upload_eri_factor: nmo and kdim must be positive");
        need(nocc > 0 && nocc < nmo, "OrbitalSpace: invalid number of occupied orbitals");
        const int N = nmo * nmo;
        const int Np = mgw::round_up(N, 64), Kp = mgw::round_up(kdim, 16);
        const int storage = choose_storage(ctx, nmo);
        // 64 zero rows past Np: the row-slab product reads whole 64-row tiles
        const int rows_b = storage == MGW_ERI_STREAMED ? Np + 64 : Np;
        std::vector<double> bp((size_t)rows_b * Kp, 0.0);
        for (int r = 0; r < N; ++r)
            std::copy(b + (size_t)r * kdim, b + (size_t)(r + 1) * kdim, bp.begin() + (size_t)r * Kp);
        double* db = as<double>(ctx->get(ctx->wsrc, bp.size() * 8, "ERI factor"));
        ctx->h2d(db, bp.data(), bp.size());
        if (storage == MGW_ERI_STREAMED) {
            upload_streamed(ctx, nmo, nocc, nullptr, db, Kp);
        } else {
            double* d = as<double>(ctx->get(ctx->eri, (size_t)N * N * 8, "ERI"));
            hip_check(mgw::launch_eri_from_factor(db, N, Kp, d, ctx->stream), "ERI from factor");
            ctx->sync();
        }
        ctx->release(ctx->wsrc);
        finish_eri_upload(ctx, nmo, nocc, storage);
    });
}

int mgw_hip_download_eri(mgw_hip_ctx* ctx, double* eri) {
    return guarded(ctx, [&] {
        need(ctx && eri, "mgw_hip_download_eri: NULL argument");
        if (!ctx->nmo) fail(MGW_ERR_STATE, "mgw_hip_download_eri: no resident ERI");
        if (ctx->eri_storage != MGW_ERI_RESIDENT)
            fail(MGW_ERR_UNSUPPORTED,
                 "mgw_hip_download_eri: the streamed storage holds only the (pk|ia) block");
        const size_t count = (size_t)ctx->nmo * ctx->nmo * ctx->nmo * ctx->nmo;
        ctx->d2h(eri, ctx->eri.p, count);
        ctx->sync();
    });
}

int mgw_hip_release_buffers(mgw_hip_ctx* ctx) {
    return guarded(ctx, [&] {
        need(ctx != nullptr, "mgw_hip_release_buffers: ctx is NULL");
        ctx->sync();
        for (auto* b : ctx->all) ctx->release(*b);
        ctx->sync();                 // stream-ordered frees have returned the memory to the pool
        ctx->nmo = ctx->nocc = 0;
        ctx->eri_storage = 0;
        ctx->n = ctx->np = 0;
        ctx->v_ready = ctx->wc_ready = false;
    });
}

int mgw_hip_quadratic_forms(mgw_hip_ctx* ctx, int32_t p, int32_t k0, int32_t width, double* q_re,
                            double* q_im) {
    return guarded(ctx, [&] {
        need(ctx && q_re, "mgw_hip_quadratic_forms: NULL argument");
        if (!ctx->nmo) fail(MGW_ERR_STATE, "fill_panel before the ERI was uploaded");
        if (!ctx->wc_ready) fail(MGW_ERR_STATE, "quadratic_forms_panel before compute_w_c");
        const int nmo = ctx->nmo;
        need(p >= 0 && p < nmo, "DevicePqPhPanelView::fill_panel: p index out of range");
        need(k0 >= 0 && width >= 0 && k0 <= nmo && width <= nmo - k0,
             "DevicePqPhPanelView::fill_panel: k panel exceeds nmo");
        const PhLayout L = ph_layout(nmo, ctx->nocc, ctx->eri_storage);
        need(L.n == ctx->n, "quadratic forms: V_ph dimension %d does not match the ERI (%d)",
             ctx->n, L.n);
        if (width == 0) return;
        std::vector<int64_t> pk(width);
        for (int c = 0; c < width; ++c) pk[c] = ((int64_t)p * nmo + k0 + c) * ctx->eri_rs;
        contract_single(ctx, as<double>(ctx->eri.p), pk.data(), width, L.colidx.data(), q_re,
                        q_im);
    });
}

int mgw_hip_quadratic_forms_noalloc(mgw_hip_ctx* ctx, int32_t p, int32_t k0, int32_t width,
                                    double* q_dev) {
    return guarded(ctx, [&] {
        need(ctx && q_dev, "mgw_hip_quadratic_forms_noalloc: NULL argument");
        if (!ctx->nmo) fail(MGW_ERR_STATE, "fill_panel before the ERI was uploaded");
        if (!ctx->wc_ready) fail(MGW_ERR_STATE, "quadratic_forms_panel before compute_w_c");
        if (ctx->wc_complex || ctx->wc_dense)
            fail(MGW_ERR_UNSUPPORTED,
                 "mgw_hip_quadratic_forms_noalloc: needs the Cholesky form of a real W_c (eta == 0, "
                 "positive definite dielectric matrix)");
        const int nmo = ctx->nmo, np = ctx->np;
        need(p >= 0 && p < nmo, "DevicePqPhPanelView::fill_panel: p index out of range");
        need(k0 >= 0 && width >= 0 && k0 <= nmo && width <= nmo - k0,
             "DevicePqPhPanelView::fill_panel: k panel exceeds nmo");
        need(ctx->n == ctx->nocc * (nmo - ctx->nocc),
             "quadratic forms: V_ph dimension %d does not match the ERI", ctx->n);
        need(width <= kNoAllocMaxWidth,
             "mgw_hip_quadratic_forms_noalloc: width %d exceeds the scratch limit %d", width,
             kNoAllocMaxWidth);
        need(np <= kNoAllocMaxNph,
             "mgw_hip_quadratic_forms_noalloc: padded nph %d exceeds the scratch limit %d", np,
             kNoAllocMaxNph);
        if (width == 0) return;
        const int mpad = mgw::round_up(width, mgw::kPad);
        char* base = static_cast<char*>(noalloc_base(ctx->device));
        int64_t* dpk = reinterpret_cast<int64_t*>(base);
        double* xp = reinterpret_cast<double*>(base + kNoAllocPk);
        double* qpart = reinterpret_cast<double*>(base + kNoAllocPk + kNoAllocX);
        int* f0p = reinterpret_cast<int*>(base + kNoAllocPk + kNoAllocX + kNoAllocQpart);
        hipStream_t st = ctx->stream;
        hip_check(mgw::launch_pk_rows(dpk, p, k0, width, nmo, mpad, ctx->eri_rs, st), "panel rows");
        hip_check(mgw::launch_pack_x(as<double>(ctx->eri.p), dpk, as<int>(ctx->colidx.p), xp, ctx->n,
                                     np, mpad, st),
                  "pack X");
        hip_check(mgw::launch_set_int(f0p, 0, st), "set offset");
        hip_check(mgw::launch_contract_f64(as<double>(ctx->wA.p), as<double>(ctx->wS.p), xp, qpart,
                                           np, mpad, 1, st),
                  "contract");
        hip_check(mgw::launch_reduce_q(qpart, nullptr, q_dev, np / mgw::kCT, width, mpad, f0p, 1, st),
                  "reduce q");
        ctx->sync();
    });
}

int mgw_hip_fill_panel(mgw_hip_ctx* ctx, int32_t p, int32_t k0, int32_t width, const double** x_dev,
                       int32_t* ld) {
    return guarded(ctx, [&] {
        need(ctx && x_dev && ld, "mgw_hip_fill_panel: NULL argument");
        if (!ctx->nmo) fail(MGW_ERR_STATE, "fill_panel before the ERI was uploaded");
        const int nmo = ctx->nmo;
        need(p >= 0 && p < nmo, "DevicePqPhPanelView::fill_panel: p index out of range");
        need(k0 >= 0 && width >= 0 && k0 <= nmo && width <= nmo - k0,
             "DevicePqPhPanelView::fill_panel: k panel exceeds nmo");
        const PhLayout L = ph_layout(nmo, ctx->nocc);
        const int mpad = mgw::round_up(std::max(width, 1), mgw::kPad);
        int64_t* dpk = as<int64_t>(ctx->get(ctx->ppk, (size_t)mpad * 8, "panel rows"));
        double* xp = as<double>(ctx->get(ctx->panel, (size_t)mpad * L.np * 8, "device panel"));
        hip_check(mgw::launch_pk_rows(dpk, p, k0, width, nmo, mpad, ctx->eri_rs, ctx->stream),
                  "panel rows");
        hip_check(mgw::launch_pack_x(as<double>(ctx->eri.p), dpk, as<int>(ctx->colidx.p), xp, L.n,
                                     L.np, mpad, ctx->stream),
                  "pack X");
        ctx->sync();   // the panel may be read on another context's stream
        *x_dev = xp;
        *ld = L.np;
    });
}

int mgw_hip_quadratic_forms_device(mgw_hip_ctx* ctx, const double* x_dev, int32_t n, int32_t ld,
                                   int32_t nvec, double* q_re, double* q_im) {
    return guarded(ctx, [&] {
        need(ctx && x_dev && q_re, "mgw_hip_quadratic_forms_device: NULL argument");
        if (!ctx->wc_ready) fail(MGW_ERR_STATE, "quadratic_forms_panel before compute_w_c");
        need(n == ctx->n && nvec >= 0 && ld >= n,
             "DeviceScreeningWorkspace::quadratic_forms_panel: inconsistent dimensions");
        hipPointerAttribute_t attr{};
        const hipError_t e = hipPointerGetAttributes(&attr, x_dev);
        if (e != hipSuccess) (void)hipGetLastError();
        need(e == hipSuccess && attr.type == hipMemoryTypeDevice && attr.device == ctx->device,
             "mgw_hip_quadratic_forms_device: the panel is not device memory of device %d",
             ctx->device);
        if (nvec == 0) return;
        std::vector<int64_t> pk(nvec);
        for (int c = 0; c < nvec; ++c) pk[c] = (int64_t)c * ld;
        std::vector<int> col(ctx->np);
        std::iota(col.begin(), col.end(), 0);
        contract_single(ctx, x_dev, pk.data(), nvec, col.data(), q_re, q_im);
    });
}

// ---------------------------------------------------- hoisted G0W0
namespace {
// One molecule's G0W0 in three phases: mol_prepare (its layout, contraction
// columns and resident per-molecule data, on its own context), run_batches
// (the frequency batches, on the lead context's stream; molecules of one
// padded order share every batch launch, their frequencies stacked), and
// mol_finish (LU re-screening of broken-down frequencies, error judge,
// Sigma_c, Pade/QP and downloads, on its own context).  A molecule run alone
// is a group of one on its own context.
struct MolRun {
    mgw_hip_ctx* ctx = nullptr;
    const mgw_g0w0_job* job = nullptr;
    int nmo = 0, nocc = 0, nfreq = 0, npar = 0, ns = 0, n = 0, np = 0, m = 0, mu = 0, mpad = 0;
    int fp = 0, cd_rank = 0, scheme = mgw::kDigits5x7;
    bool cplx = false, cd = false, mixed = false, i8 = false;
    size_t digits = 5;
    std::vector<int> states, colmap;
    std::vector<std::pair<int, int>> cols;
    std::vector<double> delta, omega, weight;
    std::vector<int64_t> pkrow;
    double cd_ms = 0.0;
    double *dV = nullptr, *dom = nullptr, *dwt = nullptr, *dmo = nullptr, *ddelta = nullptr;
    double *dvxc = nullptr, *dq = nullptr, *dqi = nullptr;
    int *dstates = nullptr, *dcolmap = nullptr;
    int64_t* dpk = nullptr;
    double *cdBt = nullptr, *cdLx = nullptr, *cdnrm = nullptr;
    uint16_t *xh = nullptr, *xl = nullptr;
    double *xs = nullptr, *xp = nullptr, *xj = nullptr;
    double *sre = nullptr, *sim = nullptr, *sx = nullptr, *qp = nullptr, *pade = nullptr;
    int* iters = nullptr;
    double2* gq = nullptr;
    int *info = nullptr, *f0p = nullptr;
    std::vector<int> failed;           // frequencies whose factorization broke down
    // timings (ms) of the frequency batches the molecule took part in
    float t_setup = 0, t_diel = 0, t_fact = 0, t_contr = 0, t_aux = 0, t_graph = 0;
    int n_contr = 0, nb = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};   // setup start / end on the molecule's stream

    MolRun() = default;
    MolRun(const MolRun&) = delete;
    MolRun& operator=(const MolRun&) = delete;
    ~MolRun() {
        // the host vectors above feed asynchronous copies on the context's
        // stream: on an error path the run can end while they are pending
        if (ctx && ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Phase 1: validate the job, lay out the particle-hole space and the
// contraction columns, and build the molecule's resident data on its context
// (V_ph, the grid, X in the form's representation, output buffers).
void mol_prepare(MolRun& r, mgw_hip_ctx* ctx, const mgw_g0w0_job* job, mgw_g0w0_out* out) {
        need(ctx && job && out, "mgw_hip_g0w0: NULL argument");
        r.ctx = ctx;
        r.job = job;
        const int nmo = job->nmo, nocc = job->nocc, nfreq = job->nfreq, npar = job->npar;
        need(nmo > 1 && nocc > 0 && nocc < nmo, "OrbitalSpace: invalid number of occupied orbitals");
        if (ctx->nmo != nmo || ctx->nocc != nocc)
            fail(MGW_ERR_STATE, "mgw_hip_g0w0: resident ERI has nmo=%d nocc=%d, job nmo=%d nocc=%d",
                 ctx->nmo, ctx->nocc, nmo, nocc);
        need(nfreq > 0, "generate_transformed_legendre_grid: num_freq_points must be positive");
        need(npar > 0, "Pade: npar must be positive");
        need(job->mo_energy && job->vxc_diag, "mgw_hip_g0w0: mo_energy / vxc_diag are NULL");
        need(out->qp != nullptr, "mgw_hip_g0w0: out->qp is NULL");
        const bool cplx = job->eta != 0.0;
        if (cplx && job->precision != MGW_PREC_FP64)
            fail(MGW_ERR_UNSUPPORTED,
                 "eta != 0 (complex screening) runs in MGW_PREC_FP64 only on the hip backend");
        need(job->precision == MGW_PREC_FP64 || job->precision == MGW_PREC_MIXED ||
                 job->precision == MGW_PREC_MIXED_INT8 || job->precision == MGW_PREC_FP64_CD ||
                 job->precision == MGW_PREC_MIXED_INT8_4,
             "mgw_hip_g0w0: unknown precision %d", job->precision);
        const bool cd = job->precision == MGW_PREC_FP64_CD;
        if (cd && ctx->eri_storage != MGW_ERI_RESIDENT)
            fail(MGW_ERR_UNSUPPORTED,
                 "MGW_PREC_FP64_CD decomposes the whole ERI pair matrix: it needs MGW_ERI_RESIDENT "
                 "storage");
        r.nmo = nmo;
        r.nocc = nocc;
        r.nfreq = nfreq;
        r.npar = npar;
        r.cplx = cplx;
        r.cd = cd;
        r.mixed = job->precision == MGW_PREC_MIXED;
        r.i8 = int8_form(job->precision);
        r.scheme = digit_scheme(job->precision);
        r.digits = (size_t)mgw::i8_digits(r.scheme);
        std::vector<int>& states = r.states;
        if (job->nstates <= 0 || !job->states) {
            states.resize(nmo);
            std::iota(states.begin(), states.end(), 0);
        } else {
            states.assign(job->states, job->states + job->nstates);
            for (size_t i = 0; i < states.size(); ++i) {
                need(states[i] >= 0 && states[i] < nmo, "Selected state is out of range");
                need(i == 0 || states[i] > states[i - 1], "states must be ascending and unique");
            }
        }
        const int ns = (int)states.size();
        const PhLayout L = ph_layout(nmo, nocc);
        const int n = L.n, np = L.np;
        const int m = ns * nmo;
        r.ns = ns;
        r.n = n;
        r.np = np;
        r.m = m;
        // Contraction columns.  With (pk|ia) == (kp|ia) exactly, q(p,k) = q(k,p):
        // each unordered pair {p,k} with p or k among the states is contracted once
        // and colmap sends every (state, k) to its column.  The decomposed form
        // has one vector per unordered pair (its ERI symmetry is checked to a
        // tolerance) and always contracts unordered pairs.
        std::vector<int>& colmap = r.colmap;
        std::vector<std::pair<int, int>>& cols = r.cols;
        colmap.resize(m);
        {
            std::vector<int> slot(nmo, -1);
            for (int s = 0; s < ns; ++s) slot[states[s]] = s;
            cols.reserve(m);
            for (int s = 0; s < ns; ++s)
                for (int k = 0; k < nmo; ++k) {
                    const int p = states[s];
                    if ((ctx->eri_pk_sym || cd) && k < p && slot[k] >= 0) {
                        colmap[(size_t)s * nmo + k] = colmap[(size_t)slot[k] * nmo + p];
                    } else {
                        colmap[(size_t)s * nmo + k] = (int)cols.size();
                        cols.emplace_back(p, k);
                    }
                }
        }
        const int mu = (int)cols.size(), mpad = mgw::round_up(mu, mgw::kPad);
        r.mu = mu;
        r.mpad = mpad;
        std::vector<double>& delta = r.delta;
        delta.assign(np, 0.0);
        for (int i = 0; i < nocc; ++i)
            for (int a = 0; a < L.nvir; ++a) {
                const double d = job->mo_energy[nocc + a] - job->mo_energy[i];
                if (!(d > 0.0))
                    fail(MGW_ERR_UNSUPPORTED,
                         "particle-hole gap e_%d - e_%d = %g is not positive", nocc + a, i, d);
                delta[i * L.nvir + a] = d;
            }
        r.omega.resize(nfreq);
        r.weight.resize(nfreq);
        transformed_legendre(nfreq, r.omega.data(), r.weight.data());
        r.pkrow.assign(mpad, 0);
        for (int c = 0; c < mu; ++c)
            r.pkrow[c] = ((int64_t)cols[c].first * nmo + cols[c].second) * ctx->eri_rs;

        hipStream_t st = ctx->stream;
        for (auto& e : r.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventRecord(r.ev[0], st), "hipEventRecord");
        Range setup_range("screening setup");

        // resident per-molecule data (the decomposed form needs no V_ph)
        r.dV = cd ? nullptr : as<double>(ctx->get(ctx->v, (size_t)np * np * 8, "V_ph"));
        r.dom = as<double>(ctx->get(ctx->homega, (size_t)nfreq * 8, "omega"));
        r.dwt = as<double>(ctx->get(ctx->hweight, (size_t)nfreq * 8, "weight"));
        r.dmo = as<double>(ctx->get(ctx->hmo, (size_t)(nmo + np) * 8, "mo/delta"));
        r.ddelta = r.dmo + nmo;
        r.dvxc = as<double>(ctx->get(ctx->hvxc, (size_t)nmo * 8, "vxc"));
        r.dstates = as<int>(ctx->get(ctx->hstates, (size_t)ns * 4, "states"));
        r.dpk = as<int64_t>(ctx->get(ctx->hpkrow, (size_t)mpad * 8, "pkrow"));
        r.dq = as<double>(ctx->get(ctx->hq, (size_t)nfreq * m * 8, "q"));
        r.dqi = cplx ? as<double>(ctx->get(ctx->hqim, (size_t)nfreq * m * 8, "q imag")) : nullptr;
        ctx->h2d(r.dom, r.omega.data(), nfreq);
        ctx->h2d(r.dwt, r.weight.data(), nfreq);
        ctx->h2d(r.dmo, job->mo_energy, nmo);
        ctx->h2d(r.ddelta, delta.data(), np);
        ctx->h2d(r.dvxc, job->vxc_diag, nmo);
        ctx->h2d(r.dstates, states.data(), ns);
        ctx->h2d(r.dpk, r.pkrow.data(), mpad);
        r.dcolmap = as<int>(ctx->get(ctx->hcolmap, (size_t)m * 4, "colmap"));
        ctx->h2d(r.dcolmap, colmap.data(), m);
        if (!cd) {
            hip_check(mgw::launch_build_v(as<double>(ctx->eri.p), as<int64_t>(ctx->rowbase.p),
                                          as<int>(ctx->colidx.p), r.dV, n, np, st),
                      "build V_ph");
            ctx->n = n;
            ctx->np = np;
            ctx->v_ready = true;
        }
        ctx->wc_ready = false;
        // Decomposed form: M = L L^T over unique pairs, then B^T (ph rows of L)
        // and the pair vectors of the contraction columns; fp = order of the
        // factored matrices (the rank, padded) instead of np.
        r.fp = np;
        if (cd) {
            const auto c0 = std::chrono::steady_clock::now();
            const CdResult c = cd_decompose(ctx, L, cols, mpad);
            r.cd_rank = c.rank;
            r.fp = c.rpad;
            r.cdBt = c.bt;
            r.cdLx = c.lx;
            r.cdnrm = c.nrm;
            r.cd_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0)
                          .count();
        }
        const mgw::Split16 fmt = mgw::Split16::FP16;
        if (cd) {
            r.xp = r.cdLx;
            r.xs = r.cdnrm;
            r.dV = r.cdBt;
        } else if (!r.mixed && !r.i8) {
            r.xp = as<double>(ctx->get(ctx->hxp, (size_t)mpad * np * 8, "packed X"));
            hip_check(mgw::launch_pack_x(as<double>(ctx->eri.p), r.dpk, as<int>(ctx->colidx.p), r.xp,
                                         n, np, mpad, st),
                      "pack X");
        }
        if (r.mixed) {
            r.xh = as<uint16_t>(ctx->get(ctx->xhi, (size_t)mpad * np * 2, "X hi"));
            r.xl = as<uint16_t>(ctx->get(ctx->xlo, (size_t)mpad * np * 2, "X lo"));
            r.xs = as<double>(ctx->get(ctx->xscale, (size_t)mpad * 8, "X scale"));
            hip_check(mgw::launch_split_x(fmt, as<double>(ctx->eri.p), r.dpk,
                                          as<int>(ctx->colidx.p), r.xh, r.xl, r.xs, n, np, mpad, st),
                      "split X");
        }
        if (r.i8) {
            need(np <= mgw::i8_max_k(r.scheme),
                 "mgw_hip_g0w0: nph %d exceeds the int8 contraction's limit %d", n,
                 mgw::i8_max_k(r.scheme));
            r.xh = as<uint16_t>(ctx->get(ctx->xhi, r.digits * mpad * np, "X slices"));
            r.xs = as<double>(ctx->get(ctx->xscale, (size_t)mpad * 8, "X scale"));
            hip_check(mgw::launch_slice_x(as<double>(ctx->eri.p), r.dpk, as<int>(ctx->colidx.p),
                                          reinterpret_cast<int8_t*>(r.xh), r.xs, n, np, mpad, r.scheme,
                                          st),
                      "slice X");
        }
        // the error judge's FP64 X (mixed forms), before the batch takes the room
        if ((r.mixed || r.i8) && ctx->judge_n > 0) {
            r.xj = as<double>(ctx->get(ctx->hxp, (size_t)mpad * np * 8, "packed X (judge)"));
            hip_check(mgw::launch_pack_x(as<double>(ctx->eri.p), r.dpk, as<int>(ctx->colidx.p), r.xj,
                                         n, np, mpad, st),
                      "pack X (judge)");
        }
        // per-molecule outputs before the batch, which takes the room that is left
        r.sre = as<double>(ctx->get(ctx->hsig_re, (size_t)ns * nfreq * 8, "sigma re"));
        r.sim = as<double>(ctx->get(ctx->hsig_im, (size_t)ns * nfreq * 8, "sigma im"));
        r.sx = as<double>(ctx->get(ctx->hsx, (size_t)nmo * 8, "sigma_x"));
        r.qp = as<double>(ctx->get(ctx->hqp, (size_t)nmo * 8, "qp"));
        r.iters = as<int>(ctx->get(ctx->hiters, (size_t)nmo * 4, "qp iterations"));
        if (const size_t pb = mgw::pade_work_bytes(nfreq, r.npar, ns))
            r.pade = as<double>(ctx->get(ctx->hpade, pb, "pade tables"));
        r.gq = as<double2>(
            ctx->get(ctx->hsg, mgw::sigma_workspace_bytes(nmo, ns, nfreq), "quadrature factors"));
        r.info = as<int>(ctx->get(ctx->hinfo, (size_t)nfreq * 4, "info"));
        r.f0p = as<int>(ctx->get(ctx->f0buf, 2 * sizeof(int), "batch offset"));
        hip_check(hipMemsetAsync(r.info, 0, (size_t)nfreq * 4, st), "hipMemsetAsync");
        hip_check(hipEventRecord(r.ev[1], st), "hipEventRecord");
}

// One frequency batch's share of a molecule: its frequencies [f0, f0 + b) at
// matrix offset `off` of the batch and partial sums at qoff doubles of qpart.
struct Seg {
    MolRun* r;
    int f0, b, off;
    size_t qoff;
};

// Phase 2: the frequency batches of molecules that share the padded order,
// the form and eta = 0 (a molecule with eta != 0 or the decomposed form runs
// alone).  Everything is enqueued on lead's stream with lead's batch buffers;
// the molecules' own data must be complete (their streams synchronized).
void run_batches(mgw_hip_ctx* lead, const std::vector<MolRun*>& mols, bool allow_graphs) {
        MolRun& r0 = *mols[0];
        const mgw_g0w0_job* job = r0.job;
        const int np = r0.np, fp = r0.fp, precision = job->precision;
        const bool cplx = r0.cplx, cd = r0.cd, mixed = r0.mixed, i8 = r0.i8;
        const int scheme = r0.scheme;
        const size_t digits = r0.digits;
        need(!mols.empty(), "run_batches: no molecule");
        int total = 0, mpad_max = 0;
        for (MolRun* r : mols) {
            need(r->np == np && r->job->precision == precision && r->scheme == scheme &&
                     (mols.size() == 1 || (!r->cplx && !r->cd)),
                 "mgw_hip_g0w0_group: molecules of one group share the padded order and the form "
                 "(eta = 0, not decomposed)");
            total += r->nfreq;
            mpad_max = std::max(mpad_max, r->mpad);
        }
        hipStream_t st = lead->stream;

        // frequency batch size from the device budget
        size_t free_b = 0, total_b = 0;
        hip_check(device_free_bytes(lead->device, &free_b, &total_b), "hipMemGetInfo");
        // the batch buffers this context already holds are reused or released
        // before a larger allocation: they count as room
        size_t reusable = 0;
        for (const mgw_hip_ctx::Buf* b : {&lead->hA, &lead->hS, &lead->hdinv, &lead->hY, &lead->hfr,
                                          &lead->hfc, &lead->htbuf, &lead->hqpart, &lead->mhi,
                                          &lead->mlo, &lead->mscale, &lead->hinfob})
            reusable += b->bytes;
        const size_t avail = free_b + reusable;
        size_t room = avail > (size_t)(2ull << 30) ? avail - (size_t)(2ull << 30) : avail / 2;
        if (lead->budget) {
            const size_t fixed = lead->held - reusable;
            room = std::min(room, lead->budget > fixed ? lead->budget - fixed : 0);
        }
        const size_t per = cd ? batch_bytes(fp, mpad_max, MGW_PREC_FP64) + (size_t)np * 8
                              : batch_bytes(np, mpad_max, precision, cplx);
        int nb = (int)std::min<size_t>((size_t)total, std::max<size_t>(1, room / per));
        int uf = 0;
        if (mols.size() == 1) {
            nb = std::min(nb, cd ? std::max(256, lead->batch_cap) : lead->batch_cap);
            const Tuned tn = tuned(r0.nmo, r0.nocc, r0.ns, precision);
            const int pick = lead->forced_nb ? lead->forced_nb : tn.nb;
            if (pick > 0) nb = std::min(nb, pick);
            uf = tn.uf;
        } else {
            // stacked molecules: at most kMaxStack matrices per launch and the
            // frequency cap's worth of the largest molecule's bytes
            nb = std::min(nb, kMaxStack);
        }
        if (lead->max_nb > 0) nb = std::min(nb, lead->max_nb);
        // equal batches: the factorization's per-batch latency is paid once per
        // batch whatever its size, so no short last batch
        nb = (total + (total + nb - 1) / nb - 1) / ((total + nb - 1) / nb);
        // untuned unit shape: measured at nmo=200, 16 frequencies per unit for
        // the int8 contraction (1-1.5 % faster than 4), 4 for the FP16 split
        uf = lead->forced_uf ? lead->forced_uf : lead->split_uf ? lead->split_uf : uf ? uf : i8 ? 16 : 4;
        const size_t zf = cplx ? 2 : 1;    // complex entries are two doubles
        const int fnblk = fp / mgw::kNB;
        const int mpad = mpad_max;
        double* A = as<double>(lead->get(lead->hA, zf * nb * fp * fp * 8, "dielectric batch"));
        double* S = as<double>(lead->get(lead->hS, zf * nb * np * 8, "scale batch"));
        double* dinv = as<double>(lead->get(lead->hdinv, zf * nb * fnblk * 64 * 64 * 8, "dinv"));
        double* Y = cplx ? as<double>(lead->get(lead->hY, (size_t)nb * np * mpad * 16, "complex panel"))
                         : nullptr;
        char *fr = nullptr, *fc = nullptr;
        if ((mixed || i8) && lead->int8_factor) {       // the mixed forms' int8 block products
            fr = as<char>(lead->get(lead->hfr, (size_t)nb * mgw::fact_slice_bytes(np, scheme),
                                    "factor row slices"));
            fc = as<char>(lead->get(lead->hfc, (size_t)nb * mgw::fact_slice_bytes(np, scheme),
                                    "factor column slices"));
        }
        double* tbuf = as<double>(lead->get(lead->htbuf, (size_t)(fr ? 2 : 1) * nb * fp * 64 * 8,
                                            "trtri scratch"));
        double* qpart = as<double>(
            lead->get(lead->hqpart, (size_t)nb * (fp / mgw::kCT) * mpad * 8, "qpart"));
        uint16_t *mh = nullptr, *ml = nullptr;
        double* ms = nullptr;
        if (mixed) {
            mh = as<uint16_t>(lead->get(lead->mhi, (size_t)nb * np * np * 2, "M hi"));
            ml = as<uint16_t>(lead->get(lead->mlo, (size_t)nb * np * np * 2, "M lo"));
            ms = as<double>(lead->get(lead->mscale, (size_t)nb * np * 8, "M scale"));
        }
        if (i8) {
            mh = as<uint16_t>(lead->get(lead->mhi, (size_t)nb * digits * np * np, "M slices"));
            ms = as<double>(lead->get(lead->mscale, (size_t)nb * np * 8, "M scale"));
        }
        int* info = as<int>(lead->get(lead->hinfob, (size_t)nb * 4, "batch info"));
        hip_check(hipMemsetAsync(info, 0, (size_t)nb * 4, st), "hipMemsetAsync");
        const mgw::Split16 fmt = mgw::Split16::FP16;

        hipEvent_t ev[8];
        for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        struct EvGuard {
            hipEvent_t* e;
            ~EvGuard() {
                for (int i = 0; i < 8; ++i) (void)hipEventDestroy(e[i]);
            }
        } evg{ev};
        float t_diel = 0, t_fact = 0, t_contr = 0, t_aux = 0, t_graph = 0;
        int n_contr = 0;
        // One frequency batch: kernels only, so that it can also be captured.
        // Per molecule the frequency offset is read on the device (its f0p).
        auto enqueue_batch = [&](const std::vector<Seg>& segs, int B, bool timed) {
            if (cplx) {
                const Seg& g = segs[0];
                MolRun& r = *g.r;
                if (timed) hip_check(hipEventRecord(ev[2], st), "hipEventRecord");
                hip_check(mgw::launch_scale_grid_c(r.ddelta, r.dom, r.job->eta, S, r.n, np, r.f0p, B,
                                                   st),
                          "scale (complex)");
                hip_check(mgw::launch_dielectric_c(r.dV, S, A, np, B, st), "dielectric (complex)");
                if (timed) hip_check(hipEventRecord(ev[3], st), "hipEventRecord");
                hip_check(mgw::launch_potrf_c(A, dinv, info, np, B, st), "potrf (complex)");
                if (timed) hip_check(hipEventRecord(ev[4], st), "hipEventRecord");
                hip_check(mgw::launch_make_y(S, r.xp, Y, np, r.mpad, B, st), "S X (complex)");
                if (timed) hip_check(hipEventRecord(ev[6], st), "hipEventRecord");
                hip_check(mgw::launch_fsub_c(A, dinv, Y, np, r.mpad, B, st), "L^-1 S X (complex)");
                if (timed) hip_check(hipEventRecord(ev[7], st), "hipEventRecord");
                hip_check(mgw::launch_sumsq_c(Y, r.dcolmap, r.dq, r.dqi, np, r.m, r.mpad, r.f0p, B, st),
                          "q (complex)");
                if (timed) hip_check(hipEventRecord(ev[5], st), "hipEventRecord");
                return;
            }
            if (cd) {
                const Seg& g = segs[0];
                MolRun& r = *g.r;
                if (timed) hip_check(hipEventRecord(ev[2], st), "hipEventRecord");
                hip_check(mgw::launch_scale_from_grid(r.ddelta, r.dom, S, r.n, np, r.f0p, B, st),
                          "scale");
                hip_check(mgw::launch_aux_dielectric(r.cdBt, S, A, fp, np, B, st),
                          "pair-space dielectric");
                if (timed) hip_check(hipEventRecord(ev[3], st), "hipEventRecord");
                factor_batch(lead, A, dinv, info, tbuf, fp, B);
                if (timed) hip_check(hipEventRecord(ev[4], st), "hipEventRecord");
                if (timed) hip_check(hipEventRecord(ev[6], st), "hipEventRecord");
                hip_check(mgw::launch_contract_f64(A, nullptr, r.cdLx, qpart, fp, r.mpad, B, st),
                          "contract (pair space)");
                if (timed) hip_check(hipEventRecord(ev[7], st), "hipEventRecord");
                hip_check(mgw::launch_reduce_q_cd(qpart, r.cdnrm, r.dcolmap, r.dq, fp / mgw::kCT, r.m,
                                                  r.mpad, r.f0p, B, st),
                          "reduce q (pair space)");
                if (timed) hip_check(hipEventRecord(ev[5], st), "hipEventRecord");
                return;
            }
            if (timed) hip_check(hipEventRecord(ev[2], st), "hipEventRecord");
            for (const Seg& g : segs) {
                MolRun& r = *g.r;
                double* Sg = S + (size_t)g.off * np;
                hip_check(mgw::launch_scale_from_grid(r.ddelta, r.dom, Sg, r.n, np, r.f0p, g.b, st),
                          "scale");
                hip_check(mgw::launch_dielectric(r.dV, Sg, A + (size_t)g.off * np * np, np, g.b, st),
                          "dielectric");
            }
            if (timed) hip_check(hipEventRecord(ev[3], st), "hipEventRecord");
            factor_batch(lead, A, dinv, info, tbuf, np, B, fr, fc, scheme);
            if (timed) hip_check(hipEventRecord(ev[4], st), "hipEventRecord");
            if (mixed)
                hip_check(mgw::launch_split_m(fmt, A, S, mh, ml, ms, np, B, st), "split M");
            if (i8)
                hip_check(mgw::launch_slice_m(A, S, reinterpret_cast<int8_t*>(mh), ms, np, B, scheme,
                                              st),
                          "slice M");
            if (timed) hip_check(hipEventRecord(ev[6], st), "hipEventRecord");
            for (const Seg& g : segs) {
                MolRun& r = *g.r;
                double* qg = qpart + g.qoff;
                const size_t o = (size_t)g.off;
                if (mixed)
                    hip_check(mgw::launch_contract_split(fmt, mh + o * np * np, ml + o * np * np,
                                                         ms + o * np, r.xh, r.xl, qg, np, r.mpad,
                                                         g.b, uf, st),
                              "contract split");
                else if (i8)
                    hip_check(mgw::launch_contract_i8(reinterpret_cast<const int8_t*>(mh) +
                                                          o * digits * np * np,
                                                      ms + o * np, reinterpret_cast<const int8_t*>(r.xh),
                                                      qg, np, r.mpad, g.b, uf, scheme, st),
                              "contract int8");
                else
                    hip_check(mgw::launch_contract_f64(A + o * np * np, S + o * np, r.xp, qg, np,
                                                       r.mpad, g.b, st),
                              "contract");
            }
            if (timed) hip_check(hipEventRecord(ev[7], st), "hipEventRecord");
            for (const Seg& g : segs) {
                MolRun& r = *g.r;
                double* qg = qpart + g.qoff;
                if (mixed || i8)
                    hip_check(mgw::launch_reduce_q_scaled(qg, r.xs, r.dcolmap, r.dq,
                                                          mixed ? mgw::contract_split_row_tiles(np)
                                                                : mgw::contract_i8_row_tiles(np),
                                                          r.m, r.mpad, r.f0p, g.b, st),
                              "reduce q");
                else
                    hip_check(mgw::launch_reduce_q(qg, r.dcolmap, r.dq, np / mgw::kCT, r.m, r.mpad,
                                                   r.f0p, g.b, st),
                              "reduce q");
            }
            if (timed) hip_check(hipEventRecord(ev[5], st), "hipEventRecord");
        };
        // Factorization breakdowns: the kernels flag matrix w of a batch in
        // info[w]; after each batch the flags are read, recorded as the
        // molecule's frequency f0 + (w - off) and cleared for the next batch
        // (outside any captured graph).
        std::vector<int> hinfo(nb);
        auto collect_info = [&](const std::vector<Seg>& segs, int B) {
            lead->d2h(hinfo.data(), info, B);
            lead->sync();
            bool any = false;
            for (const Seg& g : segs)
                for (int w = 0; w < g.b; ++w)
                    if (hinfo[g.off + w] != 0) {
                        g.r->failed.push_back(g.f0 + w);
                        any = true;
                    }
            if (any) hip_check(hipMemsetAsync(info, 0, (size_t)B * 4, st), "hipMemsetAsync");
        };
        // batches over the molecules' frequencies in order
        size_t mi = 0;
        int fi = 0, done = 0;
        while (done < total) {
            // a molecule alone stops between batches when cancelled (a grouped
            // one is dropped after its group's batches)
            if (mols.size() == 1 && r0.ctx->cancel &&
                __atomic_load_n(r0.ctx->cancel, __ATOMIC_ACQUIRE) != 0) {
                lead->sync();
                fail(MGW_ERR_STATE, "mgw_hip_g0w0: cancelled after %d of %d frequencies", done,
                     r0.nfreq);
            }
            const int B = std::min(nb, total - done);
            std::vector<Seg> segs;
            size_t qoff = 0;
            for (int off = 0; off < B;) {
                MolRun* r = mols[mi];
                const int b = std::min(B - off, r->nfreq - fi);
                segs.push_back({r, fi, b, off, qoff});
                qoff += (size_t)b * (fp / mgw::kCT) * r->mpad;
                off += b;
                fi += b;
                if (fi == r->nfreq) {
                    ++mi;
                    fi = 0;
                }
            }
            for (const Seg& g : segs)
                hip_check(mgw::launch_set_int(g.r->f0p, g.f0, st), "set offset");
            // A graph captured across the factorization's side streams (fork
            // and join events) keeps the device memory of the buffers it
            // referenced after they are freed and the graph is destroyed
            // (measured: 16 GB per nph = 5402 molecule); batches that split
            // run eagerly (at those orders a batch takes hundreds of ms, so
            // replay saves nothing measurable).
            const bool splits = fp >= lead->split_order && std::min(B, lead->factor_streams) >= 2;
            if (allow_graphs && lead->use_graphs && !cplx && !splits) {
                // Replay a graph of this batch when one with the same shape and
                // buffers was captured before (the offsets are read on the device).
                GraphKey key;
                key.sig = {(uint64_t)np, (uint64_t)fp, (uint64_t)B, (uint64_t)precision, (uint64_t)uf,
                           (uint64_t)(uintptr_t)A, (uint64_t)(uintptr_t)S, (uint64_t)(uintptr_t)dinv,
                           (uint64_t)(uintptr_t)tbuf, (uint64_t)(uintptr_t)info,
                           (uint64_t)(uintptr_t)qpart, (uint64_t)(uintptr_t)mh,
                           (uint64_t)(uintptr_t)ml, (uint64_t)(uintptr_t)ms, (uint64_t)(uintptr_t)fr,
                           (uint64_t)(uintptr_t)fc};
                for (const Seg& g : segs) {
                    const MolRun& r = *g.r;
                    for (uint64_t v : {(uint64_t)r.n, (uint64_t)r.mpad, (uint64_t)r.m, (uint64_t)g.b,
                                       (uint64_t)g.off, (uint64_t)g.qoff,
                                       (uint64_t)(uintptr_t)r.dq, (uint64_t)(uintptr_t)r.xp,
                                       (uint64_t)(uintptr_t)r.xh, (uint64_t)(uintptr_t)r.xl,
                                       (uint64_t)(uintptr_t)r.xs, (uint64_t)(uintptr_t)r.dV,
                                       (uint64_t)(uintptr_t)r.ddelta, (uint64_t)(uintptr_t)r.dom,
                                       (uint64_t)(uintptr_t)r.dcolmap, (uint64_t)(uintptr_t)r.f0p,
                                       (uint64_t)(uintptr_t)r.cdLx, (uint64_t)(uintptr_t)r.cdnrm})
                        key.sig.push_back(v);
                }
                hipGraphExec_t exec = lead->find_graph(key);
                if (!exec) {
                    hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                              "hipStreamBeginCapture");
                    enqueue_batch(segs, B, false);
                    hipGraph_t graph = nullptr;
                    hip_check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
                    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    hip_check(e, "hipGraphInstantiate");
                    lead->graphs.push_back({key, exec});
                }
                hip_check(hipEventRecord(ev[2], st), "hipEventRecord");
                hip_check(hipGraphLaunch(exec, st), "hipGraphLaunch");
                hip_check(hipEventRecord(ev[5], st), "hipEventRecord");
                hip_check(hipEventSynchronize(ev[5]), "hipEventSynchronize");
                float a = 0;
                hip_check(hipEventElapsedTime(&a, ev[2], ev[5]), "hipEventElapsedTime");
                t_graph += a;
            } else {
                {
                    Range batch_range("build W_c + sigma_c (frequency batch)");
                    enqueue_batch(segs, B, true);
                    hip_check(hipEventSynchronize(ev[5]), "hipEventSynchronize");
                }
                float a = 0, bb = 0, c = 0, d = 0, e = 0;
                hip_check(hipEventElapsedTime(&a, ev[2], ev[3]), "hipEventElapsedTime");
                hip_check(hipEventElapsedTime(&bb, ev[3], ev[4]), "hipEventElapsedTime");
                hip_check(hipEventElapsedTime(&c, ev[6], ev[7]), "hipEventElapsedTime");
                hip_check(hipEventElapsedTime(&d, ev[4], ev[6]), "hipEventElapsedTime");
                hip_check(hipEventElapsedTime(&e, ev[7], ev[5]), "hipEventElapsedTime");
                t_diel += a;
                t_fact += bb;
                t_contr += c;
                t_aux += d + e;
            }
            ++n_contr;
            collect_info(segs, B);
            done += B;
        }
        for (MolRun* r : mols) {
            r->t_diel = t_diel;
            r->t_fact = t_fact;
            r->t_contr = t_contr;
            r->t_aux = t_aux;
            r->t_graph = t_graph;
            r->n_contr = n_contr;
            r->nb = nb;
        }
}

// Phase 3a (on the molecule's own context and stream): frequencies whose
// dielectric matrix was not positive definite re-screened by pivoted LU
// (plain FP64 form; the other forms report the breakdown).
void mol_lu(MolRun& r, mgw_hip_ctx* ctx) {
        const int n = r.n, np = r.np, m = r.m, mpad = r.mpad;
        hipStream_t st = ctx->stream;
        ctx->lu_freqs = 0;
        ctx->wc_ready = false;
        std::sort(r.failed.begin(), r.failed.end());
        if (r.failed.empty()) return;
        if (r.cplx || r.cd || r.mixed || r.i8)
            fail(MGW_ERR_NUMERIC, "potrf: symmetrised dielectric matrix %d is not positive definite",
                 r.failed[0]);
        // one frequency's buffers (the context's batch buffers when it led
        // the batches, else allocated here)
        auto one = [&](mgw_hip_ctx::Buf& b, size_t bytes, const char* name) {
            return as<double>(ctx->get(b, std::max(bytes, b.bytes), name));
        };
        double* A = one(ctx->hA, (size_t)np * np * 8, "dielectric batch");
        double* S = one(ctx->hS, (size_t)np * 8, "scale batch");
        double* qpart = one(ctx->hqpart, (size_t)(np / mgw::kCT) * mpad * 8, "qpart");
        double* W = as<double>(ctx->get(ctx->luz, (size_t)np * np * 8, "E^-1"));
        for (int f : r.failed) {
            hip_check(mgw::launch_set_int(r.f0p, f, st), "set offset");
            hip_check(mgw::launch_scale_from_grid(r.ddelta, r.dom, S, n, np, r.f0p, 1, st), "scale");
            lu_screen(ctx, r.dV, S, A, W, np, 1);
            hip_check(mgw::launch_contract_dense(W, S, r.xp, qpart, np, mpad, 1, st),
                      "contract (dense)");
            hip_check(mgw::launch_reduce_q(qpart, r.dcolmap, r.dq, np / mgw::kCT, m, mpad, r.f0p, 1, st),
                      "reduce q");
        }
        ctx->lu_freqs = (int)r.failed.size();
}

// Phase 3b: the error judge of the mixed forms.  A sample of the frequencies
// is recomputed with every step in FP64 (FP64 dielectric, factorization,
// contraction) and its q compared with the form's: the relative error is the
// largest |q_form - q_fp64| over the largest |q_fp64| of each sampled
// frequency.  The sampled frequencies of a molecule are a frequency batch of
// their own, run by run_batches through a view of the molecule: its grid the
// sampled omegas, its output the judge's q rows, its panel the FP64 X.
struct JudgeRun {
    MolRun* r = nullptr;
    MolRun view;
    mgw_g0w0_job job{};
    std::vector<int> jf;
};

bool judged(const MolRun& r) {
    return (r.mixed || r.i8) && r.ctx->judge_n > 0;
}

void judge_view(MolRun& r, JudgeRun& j) {
        mgw_hip_ctx* ctx = r.ctx;
        const int nj = std::min(ctx->judge_n, r.nfreq);
        j.r = &r;
        j.jf.resize(nj);
        std::vector<double> jom(nj);
        for (int i = 0; i < nj; ++i) {
            j.jf[i] = (int)(((int64_t)2 * i + 1) * r.nfreq / (2 * nj));
            jom[i] = r.omega[j.jf[i]];
        }
        j.job = *r.job;
        j.job.precision = MGW_PREC_FP64;
        j.job.nfreq = nj;
        MolRun& v = j.view;
        v.ctx = ctx;
        v.job = &j.job;
        v.nmo = r.nmo;
        v.nocc = r.nocc;
        v.nfreq = nj;
        v.npar = r.npar;
        v.ns = r.ns;
        v.n = r.n;
        v.np = r.np;
        v.fp = r.np;
        v.m = r.m;
        v.mu = r.mu;
        v.mpad = r.mpad;
        v.omega = jom;
        v.dom = as<double>(ctx->get(ctx->hjom, (size_t)nj * 8, "judge grid"));
        v.dq = as<double>(ctx->get(ctx->hqj, (size_t)nj * r.m * 8, "judge q"));
        v.ddelta = r.ddelta;
        v.dV = r.dV;
        v.dcolmap = r.dcolmap;
        v.xp = r.xj;
        v.f0p = r.f0p + 1;
        ctx->h2d(v.dom, v.omega.data(), nj);
        ctx->sync();
}

// The judge's verdict on its batch: true when the form's result stands.
bool judge_verdict(JudgeRun& j) {
        MolRun& r = *j.r;
        mgw_hip_ctx* ctx = r.ctx;
        const int nj = (int)j.jf.size(), m = r.m;
        if (!j.view.failed.empty())
            fail(MGW_ERR_NUMERIC,
                 "potrf (judge): symmetrised dielectric matrix %d is not positive definite",
                 j.jf[j.view.failed[0]]);
        const double tol = ctx->judge_tol > 0.0                 ? ctx->judge_tol
                           : r.job->precision == MGW_PREC_MIXED_INT8 ? 1e-7
                           : r.i8                                   ? 1e-6
                                                                    : 1e-5;
        std::vector<double> qa((size_t)nj * m), qb((size_t)nj * m);
        ctx->d2h(qa.data(), j.view.dq, (size_t)nj * m);
        for (int i = 0; i < nj; ++i) ctx->d2h(qb.data() + (size_t)i * m, r.dq + (size_t)j.jf[i] * m, m);
        ctx->sync();
        double worst = 0.0;
        for (int i = 0; i < nj; ++i) {
            double e = 0.0, sc = 0.0;
            for (int c = 0; c < m; ++c) {
                e = std::max(e, std::fabs(qa[(size_t)i * m + c] - qb[(size_t)i * m + c]));
                sc = std::max(sc, std::fabs(qa[(size_t)i * m + c]));
            }
            const double rel = sc > 0.0 ? e / sc : (e > 0.0 ? INFINITY : 0.0);
            worst = std::max(worst, std::isnan(rel) ? INFINITY : rel);
        }
        ctx->judge_err = worst;
        ctx->judge_used = nj;
        return worst <= tol || !ctx->judge_fallback;
}

// Phase 3c: Sigma_c, Sigma_x, Pade/QP and the downloads, enqueued on the
// molecule's stream (mol_tail_end waits for them and records the timings).
struct Tail {
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<double> nan;
    std::vector<int> zero;
    ~Tail() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void mol_tail_begin(MolRun& r, mgw_hip_ctx* ctx, mgw_g0w0_out* out, Tail& t) {
        const mgw_g0w0_job* job = r.job;
        const int nmo = r.nmo, nfreq = r.nfreq, ns = r.ns;
        hipStream_t st = ctx->stream;
        if (out->q) ctx->d2h(out->q, r.dq, (size_t)nfreq * r.m);
        for (auto& e : t.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventRecord(t.ev[0], st), "hipEventRecord");
        {
            Range rg("sigma_c wall");
            hip_check(mgw::launch_sigma_c(r.dq, r.dqi, r.dmo, r.dom, r.dwt, job->fermi_energy, nmo, ns,
                                          nfreq, r.gq, r.sre, r.sim, st),
                      "sigma_c");
        }
        hip_check(hipEventRecord(t.ev[1], st), "hipEventRecord");
        {
            Range rg("exchange");
            // (pi|pi): eri[p, i, p, i] of the tensor, or the streamed storage's diagonal
            const int64_t n2 = (int64_t)nmo * nmo;
            const bool res = ctx->eri_storage == MGW_ERI_RESIDENT;
            hip_check(mgw::launch_sigma_x(as<double>(res ? ctx->eri.p : ctx->exsd.p), r.sx, nmo,
                                          r.nocc, res ? n2 * nmo + nmo : r.nocc, res ? n2 + 1 : 1,
                                          st),
                      "sigma_x");
        }
        Range rg("pade/output pade");
        t.nan.assign(nmo, std::numeric_limits<double>::quiet_NaN());
        t.zero.assign(nmo, 0);
        ctx->h2d(r.qp, t.nan.data(), nmo);
        ctx->h2d(r.iters, t.zero.data(), nmo);
        hip_check(mgw::launch_pade_qp(r.sre, r.sim, r.dom, r.dmo, r.sx, r.dvxc, r.dstates,
                                      job->fermi_energy, ns, nfreq, r.npar, r.qp, r.iters, r.pade,
                                      st),
                  "pade/qp");
        hip_check(hipEventRecord(t.ev[2], st), "hipEventRecord");
        if (out->sigma_c_re) ctx->d2h(out->sigma_c_re, r.sre, (size_t)ns * nfreq);
        if (out->sigma_c_im) ctx->d2h(out->sigma_c_im, r.sim, (size_t)ns * nfreq);
        if (out->sigma_x_diag) ctx->d2h(out->sigma_x_diag, r.sx, nmo);
        ctx->d2h(out->qp, r.qp, nmo);
        if (out->qp_iterations) ctx->d2h(out->qp_iterations, r.iters, nmo);
}

void mol_tail_end(MolRun& r, mgw_hip_ctx* ctx, mgw_g0w0_out* out, Tail& t) {
        ctx->sync();
        if (out->omega) std::copy(r.omega.begin(), r.omega.end(), out->omega);
        if (out->weight) std::copy(r.weight.begin(), r.weight.end(), out->weight);
        float tv = 0, tsig = 0, tqp = 0, ttot = 0;
        hip_check(hipEventElapsedTime(&tv, r.ev[0], r.ev[1]), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&tsig, t.ev[0], t.ev[1]), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&tqp, t.ev[1], t.ev[2]), "hipEventElapsedTime");
        // total: from the start of the molecule's setup to its QP energies
        hip_check(hipEventElapsedTime(&ttot, r.ev[0], t.ev[2]), "hipEventElapsedTime");
        const double tm[7] = {tv, r.t_diel, r.t_fact, r.t_contr, tsig, tqp, ttot};
        std::copy(tm, tm + 7, ctx->timings);
        ctx->timings[7] = r.nb;
        ctx->timings[8] = r.n_contr;
        ctx->timings[9] = r.t_aux;
        ctx->timings[10] = r.mu;
        ctx->timings[11] = r.t_graph;
        ctx->timings[12] = r.cd_rank;
        ctx->timings[13] = r.cd_ms;
        ctx->timings[14] = ctx->lu_freqs;
}

// One molecule on its own context.  Returns false, having written nothing to
// `out`, when the error judge failed and the context asks for the FP64
// fallback.
bool g0w0_body(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, mgw_g0w0_out* out) {
    MolRun r;
    mol_prepare(r, ctx, job, out);
    run_batches(ctx, {&r}, true);
    mol_lu(r, ctx);
    if (judged(r)) {
        JudgeRun j;
        judge_view(r, j);
        run_batches(ctx, {&j.view}, false);
        if (!judge_verdict(j)) return false;
    }
    Tail t;
    mol_tail_begin(r, ctx, out, t);
    mol_tail_end(r, ctx, out, t);
    return true;
}

// A mixed form's fallback to FP64: 1 after a failed judge, 2 after a
// factorization breakdown under an active judge (whose error is kept).
bool falls_back(const mgw_g0w0_job* job, const mgw_hip_ctx* ctx, const MgwError& e) {
    return mixed_form(job->precision) && ctx->judge_n > 0 && ctx->judge_fallback &&
           e.code == MGW_ERR_NUMERIC;
}

void rerun_fp64(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, mgw_g0w0_out* out, int why) {
    ctx->fell_back = why;
    mgw_g0w0_job plain = *job;
    plain.precision = MGW_PREC_FP64;
    const double err = ctx->judge_err;
    const int used = ctx->judge_used;
    g0w0_body(ctx, &plain, out);
    ctx->judge_err = err;
    ctx->judge_used = used;
}

void reset_judge(mgw_hip_ctx* ctx) {
    ctx->fell_back = 0;
    ctx->judge_err = -1.0;
    ctx->judge_used = 0;
}
}  // namespace

int mgw_hip_g0w0(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, mgw_g0w0_out* out) {
    return guarded(ctx, [&] {
        need(ctx && job && out, "mgw_hip_g0w0: NULL argument");
        reset_judge(ctx);
        int why = 0;
        try {
            if (!g0w0_body(ctx, job, out)) why = 1;
        } catch (const MgwError& e) {
            if (!falls_back(job, ctx, e)) throw;
            why = 2;
        }
        if (why) rerun_fp64(ctx, job, out, why);
    });
}

int mgw_hip_g0w0_group(mgw_hip_ctx* const* ctxs, const mgw_g0w0_job* jobs, mgw_g0w0_out* outs,
                       int32_t count, int32_t* status) {
    mgw_hip_ctx* lead = ctxs && count > 0 ? ctxs[0] : nullptr;
    std::vector<int> st(count > 0 ? count : 0, MGW_OK);
    std::vector<std::string> msg(st.size());
    const int rc = guarded(lead, [&] {
        need(ctxs && jobs && outs && count > 0, "mgw_hip_g0w0_group: NULL argument or count < 1");
        for (int i = 0; i < count; ++i) {
            need(ctxs[i] != nullptr, "mgw_hip_g0w0_group: context %d is NULL", i);
            need(ctxs[i]->device == lead->device,
                 "mgw_hip_g0w0_group: the contexts must share one device");
            for (int k = 0; k < i; ++k)
                need(ctxs[k] != ctxs[i],
                     "mgw_hip_g0w0_group: context %d appears twice (a context holds one resident ERI)",
                     i);
        }
        auto job_error = [&](int i, const MgwError& e) {
            st[i] = e.code;
            msg[i] = e.what();
        };
        std::vector<std::unique_ptr<MolRun>> runs(count);
        for (int i = 0; i < count; ++i) {
            reset_judge(ctxs[i]);
            runs[i] = std::make_unique<MolRun>();
            try {
                mol_prepare(*runs[i], ctxs[i], &jobs[i], &outs[i]);
            } catch (const MgwError& e) {
                job_error(i, e);
            }
        }
        for (int i = 0; i < count; ++i)
            if (st[i] == MGW_OK) ctxs[i]->sync();
        // molecules of one padded order and form share batches; eta != 0 and
        // the decomposed form run alone, each on its own context
        auto batch_groups = [&](const std::vector<MolRun*>& mols, bool graphs) {
            std::vector<std::vector<MolRun*>> groups;
            std::map<std::pair<int, int>, size_t> slot;
            for (MolRun* r : mols) {
                if (r->cplx || r->cd) {
                    groups.push_back({r});
                    continue;
                }
                const auto key = std::make_pair(r->np, r->job->precision);
                auto it = slot.find(key);
                if (it == slot.end()) {
                    slot[key] = groups.size();
                    groups.push_back({r});
                } else {
                    groups[it->second].push_back(r);
                }
            }
            for (auto& g : groups) run_batches(g.size() == 1 ? g[0]->ctx : lead, g, graphs);
        };
        {
            std::vector<MolRun*> mols;
            for (int i = 0; i < count; ++i)
                if (st[i] == MGW_OK) mols.push_back(runs[i].get());
            batch_groups(mols, true);
        }
        // per job: 0 the form's result stands, 1 / 2 rerun in FP64 (judge, breakdown)
        std::vector<int> why(count, 0);
        auto job_failure = [&](int i, const MgwError& e) {
            if (falls_back(&jobs[i], ctxs[i], e))
                why[i] = 2;
            else
                job_error(i, e);
        };
        for (int i = 0; i < count; ++i) {
            if (st[i] != MGW_OK) continue;
            try {
                // a job cancelled while its group ran keeps no result
                if (ctxs[i]->cancel && __atomic_load_n(ctxs[i]->cancel, __ATOMIC_ACQUIRE) != 0)
                    fail(MGW_ERR_STATE, "mgw_hip_g0w0: cancelled during its group's frequency batches");
                mol_lu(*runs[i], ctxs[i]);
            } catch (const MgwError& e) {
                job_failure(i, e);
            }
        }
        // the judges' sampled frequencies, stacked like the molecules' own
        std::vector<std::unique_ptr<JudgeRun>> judges(count);
        {
            std::vector<MolRun*> views;
            for (int i = 0; i < count; ++i)
                if (st[i] == MGW_OK && !why[i] && judged(*runs[i])) {
                    judges[i] = std::make_unique<JudgeRun>();
                    judge_view(*runs[i], *judges[i]);
                    views.push_back(&judges[i]->view);
                }
            batch_groups(views, false);
        }
        for (int i = 0; i < count; ++i) {
            if (!judges[i]) continue;
            try {
                if (!judge_verdict(*judges[i])) why[i] = 1;
            } catch (const MgwError& e) {
                job_failure(i, e);
            }
        }
        // the tails of every molecule enqueued before any is waited for
        std::vector<Tail> tails(count);
        for (int i = 0; i < count; ++i)
            if (st[i] == MGW_OK && !why[i]) mol_tail_begin(*runs[i], ctxs[i], &outs[i], tails[i]);
        for (int i = 0; i < count; ++i)
            if (st[i] == MGW_OK && !why[i]) mol_tail_end(*runs[i], ctxs[i], &outs[i], tails[i]);
        for (int i = 0; i < count; ++i) {
            if (st[i] != MGW_OK || !why[i]) continue;
            try {
                rerun_fp64(ctxs[i], &jobs[i], &outs[i], why[i]);
            } catch (const MgwError& e) {
                job_error(i, e);
            }
        }
    });
    if (rc != MGW_OK) std::fill(st.begin(), st.end(), rc);
    int first = rc;
    for (size_t i = 0; i < st.size(); ++i) {
        if (rc == MGW_OK && st[i] != MGW_OK) {
            ctxs[i]->error = msg[i];
            if (first == MGW_OK) first = st[i];
        } else if (rc == MGW_OK && ctxs[i]) {
            ctxs[i]->error.clear();
        }
        if (status) status[i] = st[i];
    }
    return first;
}

int mgw_hip_pade_qp(mgw_hip_ctx* ctx, int32_t nmo, int32_t nstates, const int32_t* states,
                    int32_t nfreq, int32_t npar, const double* omega, const double* sigma_re,
                    const double* sigma_im, const double* sigma_x_diag, const double* mo_energy,
                    const double* vxc_diag, double fermi_energy, double* qp, int32_t* iterations) {
    return guarded(ctx, [&] {
        need(ctx && omega && sigma_re && sigma_im && sigma_x_diag && mo_energy && vxc_diag && qp,
             "mgw_hip_pade_qp: NULL argument");
        need(nmo > 0 && nfreq > 0, "mgw_hip_pade_qp: nmo and nfreq must be positive");
        need(npar > 0, "Pade: npar must be positive");
        std::vector<int> st;
        if (nstates <= 0 || !states) {
            st.resize(nmo);
            std::iota(st.begin(), st.end(), 0);
        } else {
            st.assign(states, states + nstates);
            for (size_t i = 0; i < st.size(); ++i) {
                need(st[i] >= 0 && st[i] < nmo, "Selected state is out of range");
                need(i == 0 || st[i] > st[i - 1], "states must be ascending and unique");
            }
        }
        const int ns = (int)st.size();
        hipStream_t s = ctx->stream;
        double* dom = as<double>(ctx->get(ctx->homega, (size_t)nfreq * 8, "omega"));
        double* sre = as<double>(ctx->get(ctx->hsig_re, (size_t)ns * nfreq * 8, "sigma re"));
        double* sim = as<double>(ctx->get(ctx->hsig_im, (size_t)ns * nfreq * 8, "sigma im"));
        double* sx = as<double>(ctx->get(ctx->hsx, (size_t)nmo * 8, "sigma_x"));
        double* dmo = as<double>(ctx->get(ctx->hmo, (size_t)nmo * 8, "mo energies"));
        double* dvxc = as<double>(ctx->get(ctx->hvxc, (size_t)nmo * 8, "vxc"));
        int* dst = as<int>(ctx->get(ctx->hstates, (size_t)ns * 4, "states"));
        double* dqp = as<double>(ctx->get(ctx->hqp, (size_t)nmo * 8, "qp"));
        int* dit = as<int>(ctx->get(ctx->hiters, (size_t)nmo * 4, "qp iterations"));
        const size_t pb = mgw::pade_work_bytes(nfreq, npar, ns);
        double* dpw = pb ? as<double>(ctx->get(ctx->hpade, pb, "pade tables")) : nullptr;
        ctx->h2d(dom, omega, nfreq);
        ctx->h2d(sre, sigma_re, (size_t)ns * nfreq);
        ctx->h2d(sim, sigma_im, (size_t)ns * nfreq);
        ctx->h2d(sx, sigma_x_diag, nmo);
        ctx->h2d(dmo, mo_energy, nmo);
        ctx->h2d(dvxc, vxc_diag, nmo);
        ctx->h2d(dst, st.data(), ns);
        std::vector<double> nan(nmo, std::numeric_limits<double>::quiet_NaN());
        std::vector<int> zero(nmo, 0);
        ctx->h2d(dqp, nan.data(), nmo);
        ctx->h2d(dit, zero.data(), nmo);
        hip_check(mgw::launch_pade_qp(sre, sim, dom, dmo, sx, dvxc, dst, fermi_energy, ns, nfreq,
                                      npar, dqp, dit, dpw, s),
                  "pade/qp");
        ctx->d2h(qp, dqp, nmo);
        if (iterations) ctx->d2h(iterations, dit, nmo);
        ctx->sync();
    });
}

int mgw_hip_pade_real_axis(mgw_hip_ctx* ctx, const double* sigma_re, const double* sigma_im,
                           const double* omega, int32_t nfreq, int32_t npar, double fermi_energy,
                           double x_start, double interval, int32_t npoints, double* out_re,
                           double* out_im) {
    return guarded(ctx, [&] {
        need(ctx && sigma_re && sigma_im && omega && out_re && out_im,
             "mgw_hip_pade_real_axis: NULL argument");
        need(nfreq > 0 && npoints >= 0, "mgw_hip_pade_real_axis: nfreq must be positive");
        need(npar > 0, "Pade: npar must be positive");
        if (npoints == 0) return;
        const size_t pb = mgw::pade_work_bytes(nfreq, npar, 1);
        const size_t wo = ((size_t)3 * nfreq + 2 * (size_t)npoints + 1) & ~(size_t)1;   // 16 B aligned
        double* d = as<double>(ctx->get(ctx->wsrc, wo * 8 + pb, "pade"));
        ctx->h2d(d, sigma_re, nfreq);
        ctx->h2d(d + nfreq, sigma_im, nfreq);
        ctx->h2d(d + 2 * nfreq, omega, nfreq);
        double* ore = d + 3 * nfreq;
        double* oim = ore + npoints;
        hip_check(mgw::launch_pade_real_axis(d, d + nfreq, d + 2 * nfreq, nfreq, npar, fermi_energy,
                                             x_start, interval, npoints, ore, oim,
                                             pb ? d + wo : nullptr, ctx->stream),
                  "pade real axis");
        ctx->d2h(out_re, ore, npoints);
        ctx->d2h(out_im, oim, npoints);
        ctx->sync();
    });
}

int mgw_hip_selftest_tile(mgw_hip_ctx* ctx, const double* a, const double* b, double* c) {
    return guarded(ctx, [&] {
        need(ctx && a && b && c, "mgw_hip_selftest_tile: NULL argument");
        double* d = as<double>(ctx->get(ctx->wsrc, 3 * 64 * 64 * 8, "selftest"));
        ctx->h2d(d, a, 64 * 64);
        ctx->h2d(d + 4096, b, 64 * 64);
        hip_check(mgw::launch_tile_selftest(d, d + 4096, d + 8192, ctx->stream), "tile selftest");
        ctx->d2h(c, d + 8192, 64 * 64);
        ctx->sync();
    });
}

int mgw_hip_probe_mfma(mgw_hip_ctx* ctx, int32_t kind, int32_t ntests, const void* a,
                       const void* b, void* d) {
    return guarded(ctx, [&] {
        need(ctx && a && b && d && ntests > 0 && ntests <= 65536, "mgw_hip_probe_mfma: bad argument");
        need(kind >= 0 && kind <= 5,
             "mgw_hip_probe_mfma: kind must be 0 (f64), 1/2 (bf16/f16 16x16), 3/4 (bf16/f16 32x32), "
             "5 (i8 32x32x32)");
        const size_t in_bytes = kind == 0 ? 64 * 8 : 64 * 16;
        const size_t out_bytes = kind == 0 ? 64 * 4 * 8 : kind <= 2 ? 64 * 4 * 4 : 64 * 16 * 4;
        char* buf = as<char>(ctx->get(ctx->wsrc, (size_t)ntests * (2 * in_bytes + out_bytes), "probe"));
        char* da = buf;
        char* db = buf + (size_t)ntests * in_bytes;
        char* dd = db + (size_t)ntests * in_bytes;
        ctx->h2d(da, static_cast<const char*>(a), (size_t)ntests * in_bytes);
        ctx->h2d(db, static_cast<const char*>(b), (size_t)ntests * in_bytes);
        if (kind == 0)
            hip_check(mgw::launch_probe_f64(as<double>(da), as<double>(db), as<double>(dd), ntests,
                                            ctx->stream), "probe f64");
        else if (kind <= 2)
            hip_check(mgw::launch_probe16(kind == 1 ? mgw::Split16::BF16 : mgw::Split16::FP16,
                                          as<uint16_t>(da), as<uint16_t>(db), as<float>(dd), ntests,
                                          ctx->stream), "probe 16");
        else if (kind <= 4)
            hip_check(mgw::launch_probe32(kind == 3 ? mgw::Split16::BF16 : mgw::Split16::FP16,
                                          as<uint16_t>(da), as<uint16_t>(db), as<float>(dd), ntests,
                                          ctx->stream), "probe 32");
        else
            hip_check(mgw::launch_probe_i8(as<int8_t>(da), as<int8_t>(db), as<int>(dd), ntests,
                                           ctx->stream), "probe i8");
        ctx->d2h(static_cast<char*>(d), dd, (size_t)ntests * out_bytes);
        ctx->sync();
    });
}

int mgw_hip_measure_peaks(mgw_hip_ctx* ctx, double* out) {
    return guarded(ctx, [&] {
        need(ctx && out, "mgw_hip_measure_peaks: NULL argument");
        int dev = 0, cus = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                  "hipDeviceGetAttribute");
        const size_t count = (size_t)1 << 28;                 // 2 GiB of doubles
        char* buf = as<char>(ctx->get(ctx->wsrc, 2 * count * 8 + 4096, "peak scratch"));
        float* sink = reinterpret_cast<float*>(buf + 2 * count * 8);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        const double flop[9] = {2048.0 * 16, 16384.0 * 8, 16384.0 * 8, 32768.0 * 4, 32768.0 * 4,
                                65536.0 * 4, 32768.0 * 8, 65536.0 * 8, 2048.0 * 16};
        const int iters[9] = {1000, 4000, 4000, 4000, 4000, 4000, 4000, 2000, 1000};
        const int slot[9] = {0, 1, 2, 4, 5, 6, 7, 8, 9};
        for (int kind = 0; kind < 9; ++kind) {
            const int blocks = cus * 8;
            hip_check(mgw::launch_mfma_peak(kind, blocks, 10, sink, ctx->stream), "mfma peak");
            hip_check(hipEventRecord(e0, ctx->stream), "hipEventRecord");
            hip_check(mgw::launch_mfma_peak(kind, blocks, iters[kind], sink, ctx->stream), "mfma peak");
            hip_check(hipEventRecord(e1, ctx->stream), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            const double total = (double)blocks * 4 * iters[kind] * flop[kind];
            out[slot[kind]] = total / (ms * 1e-3);
        }
        double* a = reinterpret_cast<double*>(buf);
        double* b = a + count;
        hip_check(hipMemsetAsync(a, 0, count * 8, ctx->stream), "hipMemsetAsync");
        hip_check(mgw::launch_stream_copy(a, b, count, ctx->stream), "stream copy");
        hip_check(hipEventRecord(e0, ctx->stream), "hipEventRecord");
        for (int r = 0; r < 5; ++r)
            hip_check(mgw::launch_stream_copy(a, b, count, ctx->stream), "stream copy");
        hip_check(hipEventRecord(e1, ctx->stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        out[3] = 5.0 * 2.0 * count * 8 / (ms * 1e-3);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->release(ctx->wsrc);
    });
}

int mgw_hip_autotune_result(const mgw_g0w0_job* job, int32_t* batch, int32_t* unit_freqs) {
    if (!job || !batch || !unit_freqs) return MGW_ERR_ARG;
    const int ns = job->nstates > 0 && job->states ? job->nstates : job->nmo;
    const Tuned tn = tuned(job->nmo, job->nocc, ns, job->precision);
    *batch = tn.nb;
    *unit_freqs = tn.uf;
    return MGW_OK;
}

int mgw_hip_autotune(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, int32_t* best_batch,
                     double* best_ms_per_freq) {
    return guarded(ctx, [&] {
        need(ctx && job && best_batch, "mgw_hip_autotune: NULL argument");
        // the decomposed form's pair-space batches are small: it takes up to 256
        const std::vector<int> cands = job->precision == MGW_PREC_FP64_CD
                                           ? std::vector<int>{16, 32, 64, 128, 256}
                                           : std::vector<int>{8, 16, 32, 64};
        const int nmo = job->nmo;
        const int ns = job->nstates > 0 && job->states ? job->nstates : nmo;
        std::vector<double> qp(nmo), sre((size_t)ns * 128), sim((size_t)ns * 128);
        int best = 0;
        double best_t = 0.0;
        const bool graphs = ctx->use_graphs;
        ctx->use_graphs = false;
        for (int c : cands) {
            if (c > job->nfreq && best > 0) break;
            mgw_g0w0_job j = *job;
            j.nfreq = std::min(job->nfreq, 2 * c);
            sre.resize((size_t)ns * j.nfreq);
            sim.resize((size_t)ns * j.nfreq);
            mgw_g0w0_out o{};
            o.qp = qp.data();
            o.sigma_c_re = sre.data();
            o.sigma_c_im = sim.data();
            ctx->forced_nb = c;
            const int rc = mgw_hip_g0w0(ctx, &j, &o);
            ctx->forced_nb = 0;
            if (rc != MGW_OK) {
                ctx->use_graphs = graphs;
                throw MgwError(rc, ctx->error);
            }
            if ((int)ctx->timings[7] < c) break;      // memory capped the batch: stop growing
            const double t = (ctx->timings[1] + ctx->timings[2] + ctx->timings[3] + ctx->timings[9]) / j.nfreq;
            if (best == 0 || t < best_t) {
                best = c;
                best_t = t;
            }
        }
        // mixed forms: the contraction's unit shape (frequencies per 32-tile
        // unit sharing an XCD's L2) at the chosen batch
        int best_uf = 0;
        if (mixed_form(job->precision) && best > 0) {
            double best_c = 0.0;
            for (int uf : {4, 2, 8, 16}) {
                mgw_g0w0_job j = *job;
                j.nfreq = std::min(job->nfreq, 2 * best);
                sre.resize((size_t)ns * j.nfreq);
                sim.resize((size_t)ns * j.nfreq);
                mgw_g0w0_out o{};
                o.qp = qp.data();
                o.sigma_c_re = sre.data();
                o.sigma_c_im = sim.data();
                ctx->forced_nb = best;
                ctx->forced_uf = uf;
                const int rc = mgw_hip_g0w0(ctx, &j, &o);
                ctx->forced_nb = 0;
                ctx->forced_uf = 0;
                if (rc != MGW_OK) {
                    ctx->use_graphs = graphs;
                    throw MgwError(rc, ctx->error);
                }
                const double tc = ctx->timings[3] / j.nfreq;
                if (best_uf == 0 || tc < best_c) {
                    best_uf = uf;
                    best_c = tc;
                }
            }
        }
        ctx->use_graphs = graphs;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[{nmo, job->nocc, ns, job->precision}] = Tuned{best, best_uf};
        }
        *best_batch = best;
        if (best_ms_per_freq) *best_ms_per_freq = best_t;
    });
}

int mgw_hip_last_timings(const mgw_hip_ctx* ctx, double* ms, int32_t n) {
    if (!ctx || !ms || n < 0) {
        g_thread_error = "mgw_hip_last_timings: bad argument";
        return MGW_ERR_ARG;
    }
    for (int i = 0; i < n && i < 15; ++i) ms[i] = ctx->timings[i];
    return MGW_OK;
}

size_t mgw_hip_estimate_bytes(int32_t nmo, int32_t nocc, int32_t nfreq, int32_t nstates,
                              int32_t nbatch, int32_t precision) {
    if (nmo <= 1 || nocc <= 0 || nocc >= nmo || nfreq <= 0) return 0;
    const PhLayout L = ph_layout(nmo, nocc);
    const int ns = nstates > 0 ? nstates : nmo;
    const int m = ns * nmo, mpad = mgw::round_up(m, mgw::kPad);
    const size_t eri = (size_t)nmo * nmo * nmo * nmo * 8;
    if (precision == MGW_PREC_FP64_CD) {
        // decomposed form, sized for a rank up to 4 nmo (the factor's first
        // capacity; a larger rank grows it, which the budget then checks)
        const int N = nmo * (nmo + 1) / 2, Npad = mgw::round_up(N, 64);
        const int cap = std::min(N, mgw::round_up(std::max(256, 4 * nmo), 64));
        const int rpad = mgw::round_up(std::min(N, 4 * nmo), mgw::kCT);
        const size_t fixed = eri + mgw::sigma_workspace_bytes(nmo, ns, nfreq) + (size_t)nfreq * m * 8 +
                             (size_t)cap * Npad * 8 +
                             (size_t)Npad * 24 + (size_t)rpad * L.np * 8 +
                             (size_t)mpad * (rpad + 1) * 8 + ((size_t)N + L.np + mpad) * 4;
        const int nb = std::max(1, std::min({nbatch > 0 ? nbatch : nfreq, nfreq, 256}));
        return fixed + (size_t)nb * (batch_bytes(rpad, mpad, MGW_PREC_FP64) + (size_t)L.np * 8);
    }
    const size_t v = (size_t)L.np * L.np * 8;
    const size_t sg = mgw::sigma_workspace_bytes(nmo, ns, nfreq);     // quadrature factors
    size_t fixed = eri + v + sg + (size_t)nfreq * m * 8 + (size_t)L.np * 12 + (size_t)mpad * 8;
    fixed += precision == MGW_PREC_MIXED        ? (size_t)mpad * L.np * 4 + (size_t)mpad * 8
             : int8_form(precision) ? (size_t)mpad * L.np * mgw::i8_digits(digit_scheme(precision)) +
                                          (size_t)mpad * 8
                                                : (size_t)mpad * L.np * 8;
    const int nb = std::max(1, std::min(nbatch > 0 ? nbatch : nfreq, nfreq));
    return fixed + (size_t)nb * batch_bytes(L.np, mpad, precision);
}

}  // extern "C"
