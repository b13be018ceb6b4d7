/*
 * mgw_hip.h -- C ABI of the MI355X (gfx950) screening + Sigma_c backend.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (MGW_OK = 0); on failure mgw_hip_last_error(ctx) gives the message, which
 * the C++ side of an integration turns into
 * std::runtime_error("mgw_hip_<fn> failed: <message>") like the reference's
 * check_cuda / check_cublas helpers
 * (src/workspace/device_screening_workspace.cu:25-43).
 *
 * Contexts are not thread-safe: one context per host thread, as the
 * reference's device workspaces are one per MPI rank
 * (BackendCapabilities.thread_safe = false, src/linalg/linalg_cublas.cu:208-214).
 * All device work of a context runs on its own HIP streams of default
 * priority: one main stream, plus side streams that the factorization forks
 * from it and joins back within a call; every call returns after its results
 * are on the host.
 *
 * Arrays are row-major (C order) like the reference's Matrix / Tensor4
 * (include/types.hpp:56-59,182-185).  Complex values are passed as separate
 * real and imaginary arrays.
 */
#ifndef MGW_HIP_H
#define MGW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgw_hip_ctx mgw_hip_ctx;

enum mgw_status {
    MGW_OK = 0,
    MGW_ERR_ARG = 1,         /* bad argument or inconsistent dimensions          */
    MGW_ERR_HIP = 2,         /* HIP runtime error                                */
    MGW_ERR_NUMERIC = 3,     /* factorization breakdown (non-SPD dielectric)      */
    MGW_ERR_UNSUPPORTED = 4, /* valid request this backend does not implement     */
    MGW_ERR_MEMORY = 5,      /* device allocation failed / over the HBM budget    */
    MGW_ERR_STATE = 6,       /* call order (e.g. quadratic forms before compute_wc) */
    MGW_ERR_IO = 7           /* file missing, unreadable or malformed (mgw_io.h)  */
};

enum mgw_precision {
    MGW_PREC_FP64 = 0,       /* everything in FP64                                */
    MGW_PREC_MIXED = 1,      /* FP16 hi/lo split contraction on MFMA (3 products, */
                             /* FP32 acc; ~5e-8 relative in Sigma_c); the         */
                             /* factorization's block products on int8 MFMA       */
                             /* (MGW_INT8_FACTOR=0: FP64); Sigma_c/Pade/QP FP64   */
    MGW_PREC_MIXED_INT8 = 2, /* contraction on int8 MFMA: 5 slices per operand,   */
                             /* exact int32 products, FP64 epilogue (~1e-10       */
                             /* relative in Sigma_c); factorization as MIXED      */
    MGW_PREC_FP64_CD = 3,    /* FP64 throughout, with the ERI decomposed on the   */
                             /* device (pivoted Cholesky over unique pairs, to    */
                             /* mgw_hip_set_cd_tolerance): the dielectric is      */
                             /* factored in the rank-r pair space (r x r instead  */
                             /* of nph x nph, Woodbury identity).  Needs an ERI   */
                             /* with (pq|rs) = (qp|rs) = (rs|pq) (else            */
                             /* UNSUPPORTED) and eta == 0; no counterpart in the  */
                             /* C++ reference (its PySCF companion's auxiliary-   */
                             /* basis form, pyscf_prep/gw_ac.py:169-219)          */
    MGW_PREC_MIXED_INT8_4 = 4 /* as MIXED_INT8 with four balanced 8-bit digits per */
                             /* operand (10 int8 products instead of 15) in the  */
                             /* contraction and the factorization's block        */
                             /* products (~2^-32 of a row's largest element,     */
                             /* ~1e-9 relative in Sigma_c); nph <= 32512         */
};

/* ---- library / context ------------------------------------------------ */

/* Version string of the library. */
const char* mgw_hip_version(void);

/* Message of the last failed call on ctx; with ctx == NULL, of the last
 * failed call on this host thread (e.g. a failed mgw_hip_create). */
const char* mgw_hip_last_error(const mgw_hip_ctx* ctx);

/* Number of visible HIP devices. */
int mgw_hip_device_count(int* count);

/* Device for an MPI-style local rank: (local_rank / tasks_per_gpu) % count.
 * Replaces set_cuda_device_for_local_rank (include/linalg/linalg_cublas.hpp:46-48,
 * src/linalg/linalg_cublas.cu:393-405). */
int mgw_hip_device_for_local_rank(size_t local_rank, size_t tasks_per_gpu, int* device);

/* Create / destroy a context bound to one device.  Replaces the construction
 * of DeviceScreeningWorkspace + DevicePqPhPanelView
 * (src/gw/sigma_c_driver.cpp:489-491).  Read at creation: MGW_INT8_FACTOR
 * (0: FP64 block products in the mixed forms' factorization),
 * MGW_SPLIT_UNIT_FREQS (unit shape of the mixed contractions),
 * MGW_FACTOR_STREAMS (1, 2 or 4 streams for the parts of a frequency batch
 * in the factorization; default 2), MGW_FACTOR_SPLIT_ORDER (smallest
 * matrix order that is split; default 4096), MGW_ERI_SLAB_BYTES (largest
 * slab of a streamed ERI upload; default 512 MiB, at least one row p) and,
 * read once per process, MGW_ASYNC_ALLOC (1: stream-ordered allocation
 * from the device's default memory pool instead of hipMalloc / hipFree;
 * default 0).  All but the first change scheduling or memory handling only,
 * never a result. */
int mgw_hip_create(int device, mgw_hip_ctx** out);
int mgw_hip_destroy(mgw_hip_ctx* ctx);

/* Device bytes currently held by the context
 * (DeviceScreeningWorkspace::estimated_device_bytes,
 * include/workspace/device_screening_workspace.hpp:39). */
int mgw_hip_device_bytes(const mgw_hip_ctx* ctx, size_t* bytes);

/* Cap on the device bytes a context may hold (0 = no cap).  Allocations past
 * the cap fail with MGW_ERR_MEMORY; the hoisted path sizes its frequency
 * batches to fit. */
int mgw_hip_set_memory_budget(mgw_hip_ctx* ctx, size_t bytes);

/* ---- DeviceScreeningWorkspace (include/workspace/device_screening_workspace.hpp:27-59) */

/* Upload V_ph (n x n, row-major, V[jb, ia] = (jb|ia)); replaces the
 * DeviceScreeningWorkspace(const MatrixReal& v_ph) constructor
 * (src/workspace/device_screening_workspace.cu:289-310). */
int mgw_hip_upload_vph(mgw_hip_ctx* ctx, const double* v_rowmajor, int32_t n);

/* W_c = (I - diag(pi0) V)^-1 diag(pi0), kept resident (as the factor of
 * I + S V S, S = sqrt(-pi0)); replaces DeviceScreeningWorkspace::compute_w_c
 * (src/workspace/device_screening_workspace.cu:337-374).  pi0_im may be NULL
 * (all zero, eta == 0: real FP64 MFMA path); a non-zero imaginary part
 * (eta != 0) takes the complex FP64 path.  Re(pi0) > 0 (a non-positive ph
 * gap) is MGW_ERR_UNSUPPORTED.  At eta == 0 a matrix I + S V S that is not
 * positive definite (V_ph not positive semidefinite) is factored by LU with
 * partial pivoting instead and kept as its dense inverse, as the reference's
 * getrf/getrs do; a singular one is MGW_ERR_NUMERIC. */
int mgw_hip_compute_wc(mgw_hip_ctx* ctx, const double* pi0_re, const double* pi0_im, int32_t n);

/* Copy the resident W_c (n x n, row-major) to the host.  wc_im may be NULL. */
int mgw_hip_download_wc(mgw_hip_ctx* ctx, double* wc_re, double* wc_im, int32_t n);

/* q_j = x_j^T W_c x_j for a host panel X (n x nvec, row-major); replaces
 * DeviceScreeningWorkspace::quadratic_forms_panel(const MatrixReal&)
 * (src/workspace/device_screening_workspace.cu:419-436).  q_im may be NULL. */
int mgw_hip_quadratic_forms_host(mgw_hip_ctx* ctx, const double* x_rowmajor, int32_t n,
                                 int32_t nvec, double* q_re, double* q_im);

/* ---- DevicePqPhPanelView (include/workspace/device_pq_ph_panel.hpp:67-99) */

/* How a context holds the ERI; mirrors DevicePqPhPanelStorageMode
 * (include/workspace/device_pq_ph_panel.hpp:37-41, chosen in
 * src/workspace/device_pq_ph_panel.cu:168-190).
 *   MGW_ERI_RESIDENT: the whole nmo^4 tensor in HBM (FullEriResident).
 *   MGW_ERI_STREAMED: the ERI passes through the device once, in slabs of
 *     whole rows p (MGW_ERI_SLAB_BYTES each), and only what the path reads stays
 *     resident: the particle-hole block (pk|ia) (nmo^2 x nocc nvir values) and
 *     the exchange diagonal (pi|pi).  Replaces StreamingPanel, which keeps the
 *     ERI on the host and copies every panel; every result is bit for bit the
 *     resident one.  The decomposed form (MGW_PREC_FP64_CD) and
 *     mgw_hip_download_eri need the whole tensor (MGW_ERR_UNSUPPORTED).
 *   MGW_ERI_AUTO: RESIDENT while the tensor takes less than a quarter of the
 *     free device memory (of the context's budget left, if smaller), else
 *     STREAMED, by the reference's rule. */
enum mgw_eri_storage {
    MGW_ERI_AUTO = 0,
    MGW_ERI_RESIDENT = 1,
    MGW_ERI_STREAMED = 2
};

/* Storage of the ERI for the uploads that follow on ctx (default MGW_ERI_AUTO). */
int mgw_hip_set_eri_storage(mgw_hip_ctx* ctx, int32_t storage);

/* Storage the uploaded ERI took (MGW_ERI_RESIDENT or MGW_ERI_STREAMED) and the
 * device bytes it holds; replaces storage_mode() / estimated_device_bytes()
 * (include/workspace/device_pq_ph_panel.hpp:85-88).  MGW_ERR_STATE before an
 * upload. */
int mgw_hip_eri_storage(const mgw_hip_ctx* ctx, int32_t* storage, size_t* device_bytes);

/* Device bytes an ERI of this size holds in a storage (RESIDENT or STREAMED,
 * the latter with its upload slab of MGW_ERI_SLAB_BYTES as now set); 0 for
 * invalid sizes. */
size_t mgw_hip_eri_bytes(int32_t nmo, int32_t nocc, int32_t storage);

/* Upload the MO ERI (nmo^4, C order) into the context's storage; replaces
 * DevicePqPhPanelView's constructor (src/workspace/device_pq_ph_panel.cu:204-252). */
int mgw_hip_upload_eri(mgw_hip_ctx* ctx, const double* eri, int32_t nmo, int32_t nocc);

/* Build the ERI on the device from its factor B (nmo^2 x kdim, row-major,
 * symmetric in (p,q)): eri[pq][rs] = sum_K B[pq][K] B[rs][K], formed with
 * FP64 MFMA and exactly symmetric in (pq, rs), into the context's storage (a
 * streamed one forms one slab of rows at a time, same values bit for bit).
 * Saves the host product and the nmo^4 upload for factor-form inputs
 * (density fitting / Cholesky vectors, the seeded synthetic molecules). */
int mgw_hip_upload_eri_factor(mgw_hip_ctx* ctx, const double* b, int32_t nmo, int32_t nocc,
                              int32_t kdim);

/* Copy the resident ERI (nmo^4, C order) back to the host (MGW_ERI_RESIDENT only). */
int mgw_hip_download_eri(mgw_hip_ctx* ctx, double* eri);

/* Free every device buffer the context holds (the context stays usable). */
int mgw_hip_release_buffers(mgw_hip_ctx* ctx);

/* Build the resident V_ph from the resident ERI on the device; replaces the
 * host gather calculate_v_ph_matrix (src/gw/sigma.cpp:56-71) + upload. */
int mgw_hip_build_vph_from_eri(mgw_hip_ctx* ctx);

/* q[kk] = x_{p,k0+kk}^T W_c x_{p,k0+kk}, kk < width, with x_{pk}[ia] = (pk|ia)
 * streamed from the resident ERI; replaces fill_panel + quadratic_forms_panel
 * (src/workspace/device_pq_ph_panel.cu:284-314, device_screening_workspace.cu:376-417). */
int mgw_hip_quadratic_forms(mgw_hip_ctx* ctx, int32_t p, int32_t k0, int32_t width,
                            double* q_re, double* q_im);

/* The two halves of mgw_hip_quadratic_forms for callers that keep the
 * reference's split between the panel source and the screening workspace
 * (two contexts on one device).  mgw_hip_fill_panel replaces
 * DevicePqPhPanelView::fill_panel (src/workspace/device_pq_ph_panel.cu:284-314):
 * it writes X(ia, kk) = (p, k0 + kk | i a) into a device buffer owned by ctx,
 * column-major with leading dimension *ld (nph rounded up to 256, rows past
 * nph zero), and returns it in *x_dev; the panel stays valid until the next
 * fill on ctx or its destruction, and is complete when the call returns.
 * The values are real (the ERI is real), not the reference's complex copy. */
int mgw_hip_fill_panel(mgw_hip_ctx* ctx, int32_t p, int32_t k0, int32_t width,
                       const double** x_dev, int32_t* ld);

/* q[c] = x_c^T W_c x_c for the nvec columns of a device panel x_dev (column c
 * at x_dev + c * ld, n = nph rows read, ld >= n) on ctx's device; replaces
 * DeviceScreeningWorkspace::quadratic_forms_panel(const DeviceComplexPanelView&)
 * (src/workspace/device_screening_workspace.cu:376-417).  A pointer that is
 * not device memory of ctx's device is MGW_ERR_ARG. */
int mgw_hip_quadratic_forms_device(mgw_hip_ctx* ctx, const double* x_dev, int32_t n, int32_t ld,
                                   int32_t nvec, double* q_re, double* q_im);

/* mgw_hip_quadratic_forms without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * panel and every intermediate live in a fixed device scratch array of the
 * library (about 16.2 MiB per device, held until the process ends), which
 * limits width to 256 and the padded nph (nph rounded up to 256) to 8192;
 * larger requests fail with MGW_ERR_ARG.  q_dev: width doubles in device
 * memory, written with the real part of q (eta == 0 only; a complex W_c is
 * MGW_ERR_UNSUPPORTED).  Needs the resident ERI (mgw_hip_upload_eri) and
 * W_c (mgw_hip_compute_wc), returns after the results are written.  Every
 * call on one device shares that device's scratch: calls that overlap on one
 * device (from two contexts or threads) race. */
int mgw_hip_quadratic_forms_noalloc(mgw_hip_ctx* ctx, int32_t p, int32_t k0, int32_t width,
                                    double* q_dev);

/* ---- hoisted G0W0 on the resident ERI (compute_sigma_c_with_backend's
 *      device branch, src/gw/sigma_c_driver.cpp:471-514, and run_g0w0,
 *      src/gw/gw.cpp:125-191) */

typedef struct mgw_g0w0_job {
    int32_t nmo, nocc;
    int32_t nfreq;              /* transformed Gauss-Legendre points          */
    int32_t npar;               /* Pade parameters (>= 1; min(npar, nfreq) used) */
    int32_t nstates;            /* 0 = all states                             */
    const int32_t* states;      /* nstates 0-based indices (ascending), or NULL */
    const double* mo_energy;    /* nmo                                        */
    const double* vxc_diag;     /* nmo: vxc(p, p)                             */
    double fermi_energy;
    double eta;                 /* broadening; != 0 runs the complex path,    */
                                /* MGW_PREC_FP64 only (else UNSUPPORTED)      */
    int32_t precision;          /* enum mgw_precision                         */
} mgw_g0w0_job;

typedef struct mgw_g0w0_out {
    double* omega;              /* nfreq, may be NULL                         */
    double* weight;             /* nfreq, may be NULL                         */
    double* sigma_x_diag;       /* nmo, may be NULL                           */
    double* sigma_c_re;         /* nstates x nfreq (row-major), may be NULL   */
    double* sigma_c_im;         /* nstates x nfreq, may be NULL               */
    double* qp;                 /* nmo, NaN for states not computed           */
    int32_t* qp_iterations;     /* nmo, may be NULL                           */
    double* q;                  /* nfreq x (nstates * nmo), may be NULL: the  */
                                /* quadratic forms q[f][s * nmo + k] =        */
                                /* x_{p_s k}^T W_c(i w_f) x_{p_s k} (real     */
                                /* part); Sigma_c is their quadrature         */
} mgw_g0w0_out;

/* One molecule, all frequencies, on the ERI made resident by mgw_hip_upload_eri. */
int mgw_hip_g0w0(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, mgw_g0w0_out* out);

/* Several molecules at once, each job on the ERI resident in its own context
 * (jobs[i] on ctxs[i], all contexts on one device, none twice); no
 * counterpart in the reference, whose driver runs one molecule per process
 * (src/gw/sigma_c_driver.cpp:225-260 loops over its frequencies).  Molecules
 * with the same padded particle-hole order and form (eta == 0, not
 * MGW_PREC_FP64_CD) share every launch of their frequency batches: their
 * frequencies are stacked in one batch of dielectric matrices (up to 2048 at
 * once, within ctxs[0]'s memory budget and batch limit), factored, inverted
 * and cut into digits by the same launches on ctxs[0]'s stream, with
 * ctxs[0]'s batch buffers; the contraction and reduction are one launch per
 * molecule on its own panel.  The rest (setup, LU re-screening, error judge
 * and FP64 fallback, Sigma_c, Pade/QP) runs per molecule on its context, as
 * mgw_hip_g0w0 does; every result equals bit for bit the molecule's
 * mgw_hip_g0w0 result.  Other molecules (eta != 0, decomposed form, an order
 * of their own) run their batches alone on their own context.  status[i]
 * (may be NULL) receives each job's code, whose message is
 * mgw_hip_last_error(ctxs[i]); a job cancelled through its context's flag
 * while its group runs ends with MGW_ERR_STATE, the others complete.  Returns
 * MGW_OK when every job succeeded, else the first failing code. */
int mgw_hip_g0w0_group(mgw_hip_ctx* const* ctxs, const mgw_g0w0_job* jobs, mgw_g0w0_out* outs,
                       int32_t count, int32_t* status);

/* Error judge and FP64 fallback of the mixed forms (no counterpart in the
 * reference, whose only failure check is the factorization's devInfo,
 * src/workspace/device_screening_workspace.cu:324-334).  With nsample > 0,
 * every MGW_PREC_MIXED / MGW_PREC_MIXED_INT8 / MGW_PREC_MIXED_INT8_4 run of
 * mgw_hip_g0w0 on ctx recomputes nsample of its frequencies (evenly spaced)
 * with every step in FP64 and compares their quadratic forms with the
 * form's: the error is the largest |q_form - q_fp64| over the largest
 * |q_fp64| of a frequency.  When it exceeds tol (tol = 0: 1e-7 for
 * MIXED_INT8, 1e-6 for MIXED_INT8_4, 1e-5 for MIXED) or the
 * form's factorization breaks down, and fallback != 0, the call reruns the
 * job in MGW_PREC_FP64 and returns that result.  nsample = 0 turns it off. */
int mgw_hip_set_judge(mgw_hip_ctx* ctx, int32_t nsample, double tol, int32_t fallback);

/* Outcome of the last mgw_hip_g0w0 on ctx: the judge's error (-1 if not
 * judged), the frequencies it sampled, and fell_back = 0 (the form's own
 * result), 1 (rerun in FP64 after a failed judge) or 2 (rerun in FP64 after
 * a factorization breakdown). */
int mgw_hip_last_judge(const mgw_hip_ctx* ctx, double* max_rel_err, int32_t* samples,
                       int32_t* fell_back);

/* Pade fit and QP fixed point on the device for a given Sigma_c(i w): the
 * reference's solve_qp_energies (src/gw/qp_solver.cpp:10-47) with
 * get_pade_coefficients_continued_fraction / evaluate_pade_continued_fraction
 * (src/gw/pade.cpp:10-101), as run at the end of mgw_hip_g0w0.
 * sigma_re / sigma_im: nstates x nfreq row-major for `states` (ascending,
 * unique; NULL = all nmo states); omega: the nfreq imaginary frequencies.
 * Writes qp[nmo] (NaN for states not listed) and, if not NULL,
 * iterations[nmo] (0 for states not listed). */
int mgw_hip_pade_qp(mgw_hip_ctx* ctx, int32_t nmo, int32_t nstates, const int32_t* states,
                    int32_t nfreq, int32_t npar, const double* omega, const double* sigma_re,
                    const double* sigma_im, const double* sigma_x_diag, const double* mo_energy,
                    const double* vxc_diag, double fermi_energy, double* qp, int32_t* iterations);

/* Real-axis continuation of one state's Sigma_c(i w) (sigma_re / sigma_im:
 * nfreq values on the imaginary frequencies omega): the Thiele Pade of
 * mgw_hip_pade_qp evaluated at e_i = x_start + i * interval - fermi_energy,
 * i < npoints, into out_re / out_im; the values of E_c.out
 * (output_self_energy_after_pade, src/io.cpp:94-108). */
int mgw_hip_pade_real_axis(mgw_hip_ctx* ctx, const double* sigma_re, const double* sigma_im,
                           const double* omega, int32_t nfreq, int32_t npar, double fermi_energy,
                           double x_start, double interval, int32_t npoints, double* out_re,
                           double* out_im);

/* ---- request loop (continuous batching; no counterpart in the reference,
 *      which runs one molecule per process) */

typedef struct mgw_hip_server mgw_hip_server;

typedef struct mgw_serve_job {
    mgw_g0w0_job job;           /* sizes, energies, states, precision         */
    const double* eri;          /* host ERI (nmo^4), or NULL                  */
    const double* eri_factor;   /* host factor (nmo^2 x kdim), or NULL;       */
    int32_t kdim;               /* exactly one of eri / eri_factor is given   */
} mgw_serve_job;

/* Start `workers` threads on `device`, each with its own context and stream
 * (graphs on; MGW_SERVE_GRAPHS=0 in the environment at creation runs the
 * frequency batches eagerly, results unchanged).  A worker's lead context
 * keeps its device buffers from one job to the next while they hold under a
 * quarter of hbm_budget / workers (MGW_SERVE_KEEP_BUFFERS=0: freed after
 * every job; results unchanged).  Jobs are admitted in arrival order while the bytes reserved
 * by running jobs plus the next job's estimate fit hbm_budget; batch_freqs
 * caps the frequencies factored at once per job (mgw_hip_set_batch_limit).
 * Mixed-precision jobs run under the error judge with FP64 fallback
 * (mgw_hip_set_judge: 2 sampled frequencies, the form's tolerance) unless
 * mgw_hip_server_set_judge changes it. */
int mgw_hip_server_create(int device, int32_t workers, size_t hbm_budget, int32_t batch_freqs,
                          mgw_hip_server** out);

/* Cancel jobs not yet started, wait for running ones, free everything. */
int mgw_hip_server_destroy(mgw_hip_server* server);

/* Error judge of the jobs submitted from now on (mgw_hip_set_judge's
 * arguments; nsample = 0 turns it off). */
int mgw_hip_server_set_judge(mgw_hip_server* server, int32_t nsample, double tol,
                             int32_t fallback);

/* Most jobs a worker runs together (1..64, default 16; 1 turns grouping off):
 * admitted jobs that share a padded particle-hole order and form (eta == 0,
 * not MGW_PREC_FP64_CD) run through mgw_hip_g0w0_group, their frequencies
 * stacked in shared batches; each result is bit for bit the job's result
 * alone.  A grouped job's cancellation takes effect when its group's batches
 * end.  No counterpart in the reference. */
int mgw_hip_server_set_grouping(mgw_hip_server* server, int32_t max_jobs);

/* Fault injection for tests of failure handling: each job started from now
 * on gets scale (0 < scale <= 1) times its estimate as its context's device
 * budget, so that a scale below 1 makes the job's allocations fail with
 * MGW_ERR_MEMORY as an HBM shortfall would.  Default 1. */
int mgw_hip_server_set_budget_scale(mgw_hip_server* server, double scale);

/* ERI storage of the jobs submitted from now on (enum mgw_eri_storage):
 * MGW_ERI_RESIDENT (default) or MGW_ERI_STREAMED for every job; with
 * MGW_ERI_AUTO a job starts resident when its resident estimate fits what is
 * left of the budget, else streamed when its streamed estimate fits, and
 * fails with MGW_ERR_MEMORY only if even that exceeds the whole budget.  No
 * counterpart in the reference (its panel view chooses per process). */
int mgw_hip_server_set_eri_storage(mgw_hip_server* server, int32_t storage);

/* Storage a started job's ERI took (MGW_ERI_RESIDENT or MGW_ERI_STREAMED;
 * MGW_ERI_AUTO while the job has not started). */
int mgw_hip_server_job_eri_storage(mgw_hip_server* server, int64_t ticket, int32_t* storage);

/* Queue a job.  mo_energy / vxc_diag / states are copied; the ERI or factor
 * and the output buffers must stay valid until mgw_hip_server_wait reports the
 * job finished. */
int mgw_hip_server_submit(mgw_hip_server* server, const mgw_serve_job* job,
                          const mgw_g0w0_out* out, int64_t* ticket);

/* Cancel a job: a pending job finishes at once, a running one at its next
 * frequency batch, both with MGW_ERR_STATE; a finished job is not changed. */
int mgw_hip_server_cancel(mgw_hip_server* server, int64_t ticket);

/* Wait for a job (timeout_s < 0: no limit).  *status = -1 if still pending or
 * running at the timeout, else the job's mgw_status; latency_ms = submit to
 * finish, run_ms = start to finish. */
int mgw_hip_server_wait(mgw_hip_server* server, int64_t ticket, double timeout_s, int32_t* status,
                        double* latency_ms, double* run_ms);

/* Error message of a failed job ("" if none). */
const char* mgw_hip_server_job_error(mgw_hip_server* server, int64_t ticket);

/* Judge outcome of a finished job: its error (-1 if not judged) and whether
 * it was rerun in FP64 (mgw_hip_last_judge's fell_back). */
int mgw_hip_server_job_judge(mgw_hip_server* server, int64_t ticket, double* max_rel_err,
                             int32_t* fell_back);

/* Forget a finished job: its record (status, error, judge outcome) is kept
 * until this call or the server's destruction. */
int mgw_hip_server_release(mgw_hip_server* server, int64_t ticket);

/* Jobs finished so far, jobs running now, most jobs running at once, bytes
 * reserved by running jobs. */
int mgw_hip_server_stats(mgw_hip_server* server, int64_t* completed, int32_t* running,
                         int32_t* peak_running, size_t* reserved);

/* Phase times of the last mgw_hip_g0w0 call in milliseconds (HIP events on the
 * context's stream): [0] V build (+ X split; with MGW_PREC_FP64_CD the ERI
 * decomposition), [1] dielectric build, [2] factorization (potrf + trtri),
 * [3] contraction kernel, [4] Sigma_c quadrature, [5] Pade + QP, [6] total,
 * [7] frequencies per batch, [8] contraction launches, [9] M split + q
 * reduction, [10] contraction columns (unordered (p,k) pairs when the ERI is
 * exactly pk-symmetric), [11] time in replayed batch graphs (phases 1-3 and 9
 * are then 0), [12] rank of the ERI decomposition (MGW_PREC_FP64_CD, else 0),
 * [13] its host wall time, [14] frequencies whose symmetrised dielectric
 * matrix was not positive definite and was screened by pivoted LU instead
 * (MGW_PREC_FP64).  n is the number of entries the caller provides (<= 15). */
int mgw_hip_last_timings(const mgw_hip_ctx* ctx, double* ms, int32_t n);

/* Autotuner (no counterpart in the reference): times mgw_hip_g0w0 on the
 * resident ERI with 8, 16, 32 and 64 frequencies factored per batch (16 to
 * 256 for MGW_PREC_FP64_CD; two batches each) and, for the mixed forms, the
 * contraction's tile unit shape
 * (2, 4, 8 or 16 frequencies per 32-tile unit sharing an XCD's L2) at that batch;
 * keeps the fastest per (nmo, nocc, nstates, precision) for every later
 * mgw_hip_g0w0 call of the process and returns the batch with its time per
 * frequency.  MGW_SPLIT_UNIT_FREQS, when set, overrides the unit shape. */
int mgw_hip_autotune(mgw_hip_ctx* ctx, const mgw_g0w0_job* job, int32_t* best_batch,
                     double* best_ms_per_freq);

/* The choice the autotuner keeps for this job's shape: batch and unit
 * frequencies (0 = not tuned / default). */
int mgw_hip_autotune_result(const mgw_g0w0_job* job, int32_t* batch, int32_t* unit_freqs);

/* Stop flag of ctx (no counterpart in the reference): mgw_hip_g0w0 reads
 * *flag (atomically) before each frequency batch and, when it is non-zero,
 * waits for the work in flight and fails with MGW_ERR_STATE ("cancelled"),
 * writing no results.  NULL removes the flag; the caller keeps it valid while
 * it is set. */
int mgw_hip_set_cancel_flag(mgw_hip_ctx* ctx, const int32_t* flag);

/* Cap on the frequencies mgw_hip_g0w0 factors at once on ctx (0 = none: the
 * largest batch that fits the device budget, at most 64, or the autotuned
 * one).  Scheduling only; results do not depend on it. */
int mgw_hip_set_batch_limit(mgw_hip_ctx* ctx, int32_t max_freqs);

/* Run each frequency batch of mgw_hip_g0w0 as a captured hipGraph, replayed
 * for later batches and molecules of the same shape on the same buffers
 * (enable = 1), or launch kernels directly (0, the default, which also gives
 * the per-phase timings). */
int mgw_hip_set_graphs(mgw_hip_ctx* ctx, int32_t enable);

/* Stopping tolerance of the MGW_PREC_FP64_CD decomposition: pivots stop once
 * the largest residual diagonal of the pair matrix is <= tol times the first
 * (largest) pivot; every residual entry is then bounded by that value.
 * 0 <= tol < 1, default 1e-12.  No counterpart in the reference. */
int mgw_hip_set_cd_tolerance(mgw_hip_ctx* ctx, double tol);

/* Transformed Gauss-Legendre grid (src/gw/frequency_grids.cpp:62-93). */
int mgw_grid_transformed_legendre(int32_t nfreq, double* omega, double* weight);

/* HBM bytes mgw_hip_g0w0 needs for one molecule with nbatch frequencies
 * resident at once (ERI + V + dielectric batch + contraction scratch);
 * the admission estimate of the request loop.  For MGW_PREC_FP64_CD it
 * assumes a decomposition rank of at most 4 nmo (ERI + factor + pair-space
 * batch). */
size_t mgw_hip_estimate_bytes(int32_t nmo, int32_t nocc, int32_t nfreq, int32_t nstates,
                              int32_t nbatch, int32_t precision);

#ifdef __cplusplus
}
#endif

#endif /* MGW_HIP_H */
