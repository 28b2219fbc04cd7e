/*
 * richmol_amd: MI355X (gfx950) device path for the field-driven rotational
 * TDSE step of richmol (richmol.tdse.TDSE.update).
 *
 * C ABI: plain pointers and sizes, no C++ or torch types.  Complex numbers are
 * interleaved (re, im) doubles.  "device" pointers are HIP device addresses
 * (e.g. torch.Tensor.data_ptr() of a CUDA/HIP tensor); "host" pointers are
 * ordinary process memory.  Every entry point returns RM_OK (0) or a status
 * code and never throws; rm_last_error() returns the message of the last
 * failure on the calling thread.  Kernels are enqueued asynchronously on the
 * given hipStream_t (passed as void*; NULL = default stream).  Calls on one
 * plan must be serialised by the caller.
 *
 * The operator of a plan is the reference's field-contracted block operator
 * flattened to a sparse matrix in the reference's full-form state order
 * (richmol/field.py:659-692):
 *
 *   H(F) = sum_e A_e |row_e><col_e|
 *   A_e  = sum_{p in pairs(e)} K_p * MF_{term_p}                 (K:  field.py:1235)
 *   MF_t = thr_{leaf_t}( sum_{q in t} fprod_{field_q} * M_q )     (MF: field.py:1115-1142)
 *
 * where fprod are the field products (one slot per Cartesian component of
 * each tensor in a sum) and thr zeroes |MF_t| < thresh (thresh < 0: none).
 */
#ifndef RICHMOL_AMD_H
#define RICHMOL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RM_OK              0
#define RM_NOT_CONVERGED   1    /* a column hit maxorder (tdse.py:480-484) */
#define RM_DEVICE_TIMEOUT  2    /* status bit: a cooperative launch's grid barrier timed out */
#define RM_STALE_DOMAIN    4    /* status bit: reused column domains no longer held; column untouched */
#define RM_ERR_ARG        -1
#define RM_ERR_HIP        -2
#define RM_ERR_ALLOC      -3

#define RM_MAX_FIELD      32    /* field-product slots per plan */
#define RM_MAX_LEAF        8    /* tensors summed into one operator */
#define RM_MAX_KRYLOV    128    /* Krylov dimension bound (reference default maxorder=100) */

typedef struct rm_plan rm_plan;

/* Message of the last failed call on this thread ("" if none). */
const char* rm_last_error(void);

/* Library ABI version (RM_ABI_VERSION of the header it was built from). */
#define RM_ABI_VERSION 5
int rm_version(void);

/*
 * Build a device plan for one operator structure.  Replaces the per-call
 * block loop of CarTens.vec (richmol/field.py:1212-1243) and the per-step
 * MF contraction of CarTens.field (field.py:1115-1142): the structure is
 * uploaded once, afterwards only field products cross the ABI.
 *
 *   n                      basis dimension (rows = columns of H)
 *   nnz, row_ptr[n+1], col_idx[nnz]              CSR pattern (host)
 *   pair_ptr[nnz+1], pair_term[npair], pair_kval[2*npair]   K factors (host)
 *   nterm, term_ptr[nterm+1], term_field[nq], term_mval[2*nq], term_leaf[nterm]
 *                                                M factors per MF element (host)
 *   nfield <= RM_MAX_FIELD, nleaf <= RM_MAX_LEAF
 */
int rm_plan_create(int64_t n,
                   int64_t nnz, const int64_t* row_ptr, const int32_t* col_idx,
                   const int64_t* pair_ptr, const int32_t* pair_term,
                   const double* pair_kval,
                   int64_t nterm, const int64_t* term_ptr, const int32_t* term_field,
                   const double* term_mval, const int32_t* term_leaf,
                   int32_t nfield, int32_t nleaf,
                   rm_plan** out);

int rm_plan_destroy(rm_plan* plan);

/* Basis dimension and number of stored matrix entries of a plan. */
int rm_plan_info(const rm_plan* plan, int64_t* n, int64_t* nnz);

/*
 * Contract the operator with the field of one time step (CarTens.field,
 * field.py:1073-1142) on the device.  fprod: host, 2*nfield doubles;
 * thresh: host, nleaf doubles (< 0 = no threshold).  Also records whether
 * any MF element is non-zero; rm_step skips the Krylov propagation when none
 * is (the `len(H.mfmat) > 0` test of tdse.py:377).
 */
int rm_plan_set_field(rm_plan* plan, const double* fprod, const double* thresh,
                      void* stream);

/* y[:, b] = H x[:, b] for b < ncols (CarTens.vec, field.py:1145-1245).
 * x, y: device, complex, column b at offset b*ld (ld >= n). */
int rm_matvec(rm_plan* plan, const void* x, void* y, int64_t ld, int64_t ncols,
              void* stream);

/* phase[i] = exp(f/2 * diag[i])  (tdse.py:373).  diag: device, complex n;
 * phase: device, complex n. */
int rm_phase_init(int64_t n, const void* diag, double f_re, double f_im,
                  void* phase, void* stream);

/*
 * One propagation step for ncols wavepackets (TDSE.update, tdse.py:365-414;
 * Lanczos expv tdse.py:417-486), FP64:
 *   phi = P*psi;  phi = expv(f*H, phi) if the field is non-zero;  out = P*phi
 * (P = phase, or identity when phase == NULL).
 *   psi_in/psi_out: device, complex, column b at offset b*ld (may alias).
 *   krylov_fixed > 0: fixed Krylov dimension (tol/maxorder ignored);
 *   krylov_fixed == 0: the reference's adaptive rule with tol and maxorder.
 *   kused (device int32[ncols], may be NULL): Krylov dimension used per column
 *   (0 when the step skipped the expv).
 *   status (device int32[ncols], may be NULL): 0 ok, RM_NOT_CONVERGED.
 */
int rm_step(rm_plan* plan, const void* psi_in, void* psi_out, int64_t ld,
            int64_t ncols, double f_re, double f_im, const void* phase,
            int32_t krylov_fixed, int32_t maxorder, double tol,
            int32_t* kused, int32_t* status, void* stream);
/* rm_step runs the full-vector kernels, which compute in FP64 only: a plan
 * set to RM_PREC_FP32 (rm_plan_set_precision) is refused with RM_ERR_ARG. */

/*
 * nsteps consecutive steps of the same wavepackets (the user time loop of
 * examples/ocs_alignment.py:89-100 around TDSE.update, without the host in
 * between): step s contracts the operator with row s of fprod_dev (device,
 * nsteps x nfield complex) and advances every column as rm_step does.
 *   psi: device, in/out; work: device scratch of the same shape (ping-pong).
 *   kused: device int32[nsteps*ncols] or NULL (Krylov dimension per step).
 *   status: device int32[ncols] or NULL (zeroed at the start), per column
 *   the bits RM_NOT_CONVERGED (some step hit maxorder), RM_DEVICE_TIMEOUT (a
 *   cooperative launch's bounded wait gave up: the column is invalid) and
 *   RM_STALE_DOMAIN (domain reuse only, below: the column was left as it was).
 *   With status == NULL the call reports a cooperative timeout by returning
 *   RM_DEVICE_TIMEOUT (it then waits for the stream).
 * thresh: host, nleaf doubles (< 0 = none), the same for every step.
 * The full-vector kernels (taken when thresholds are given) compute in FP64
 * only: a plan set to RM_PREC_FP32 gets RM_ERR_ARG there, as from rm_step.
 */
int rm_propagate(rm_plan* plan, void* psi, void* work, int64_t ld, int64_t ncols,
                 int64_t nsteps, const double* fprod_dev, const double* thresh,
                 double f_re, double f_im, const void* phase, int32_t krylov_fixed,
                 int32_t maxorder, double tol, int32_t* kused, int32_t* status,
                 void* stream);

/*
 * rm_propagate with an observable read after every step: the loop of
 * examples/ocs_alignment.py:89-103 (TDSE.update, then <psi|O|psi> of every
 * vector and their sum) with the host out of it.  obs: a plan of the same n
 * holding the observable's operator -- the matrix of its last
 * rm_plan_set_field (e.g. cos^2 theta, tomat(form="full", cart="0")); not
 * `plan` itself.  After step s (phases and expv applied, as TDSE.update
 * returns it):
 *   expect: device complex [nsteps][ncols] (or NULL),
 *     expect[s*ncols + b] = <psi_b|O|psi_b>;
 *   expect_sum: device complex [nsteps] (or NULL), the sum over the columns
 *     (the ensemble value of ocs_alignment.py:99-100, without its +1/3).
 * At least one of them non-NULL.  The observable is evaluated in FP64 on the
 * FP64 wavepackets whatever the plan's precision.  A column whose status has
 * RM_STALE_DOMAIN gets no values (the call is to be repeated); a zero column
 * gets zeros.  obs's matrix is read on `stream`: set its field on that stream
 * (or synchronise) before the call.
 */
int rm_propagate_expect(rm_plan* plan, void* psi, void* work, int64_t ld, int64_t ncols,
                        int64_t nsteps, const double* fprod_dev, const double* thresh,
                        double f_re, double f_im, const void* phase, int32_t krylov_fixed,
                        int32_t maxorder, double tol, int32_t* kused, int32_t* status,
                        rm_plan* obs, void* expect, void* expect_sum, void* stream);

/*
 * Propagator of rm_propagate / rm_propagate_expect (TDSE.update's `propag`,
 * tdse.py:377-412): RM_PROPAG_LANCZOS (default, 'internal': the Lanczos
 * expv of tdse.py:417-486) or RM_PROPAG_EXPOKIT ('external': Expokit's
 * ZHEXPV, expokit/expokit.f:2783-3170, as tdse.py:383-391 calls it -- the
 * operator iH(F), t = Im(f), the whole matrix's 1-norm for the first step
 * size).  With Expokit, krylov_fixed is ZHEXPV's Krylov size m (the
 * reference passes 12; 1 <= m <= RM_EXPOKIT_MAX_M, m < n), tol its
 * tolerance (maxorder unused), kused receives the matvecs of each step, and
 * a column that needed more than 500 sub-steps gets RM_NOT_CONVERGED (the
 * reference raises).  FP64 only; no MF thresholds.
 */
#define RM_PROPAG_LANCZOS 0
#define RM_PROPAG_EXPOKIT 1
#define RM_EXPOKIT_MAX_M 16
int rm_plan_set_propagator(rm_plan* plan, int32_t propag);

/* Measurement hooks (no reference counterpart).  With timing enabled, every
 * Krylov-step launch of rm_step / rm_propagate is bracketed by HIP events on
 * its stream; rm_plan_kernel_time waits for the last one, returns the summed
 * launch durations and the number of launches, and resets the record. */
int rm_plan_timing(rm_plan* plan, int32_t enable);
int rm_plan_kernel_time(rm_plan* plan, double* total_ms, int64_t* launches);

/* Kernel used by rm_step / rm_propagate (no reference counterpart: a device
 * tuning knob).  All variants compute the same step.
 *
 * RM_KERNEL_AUTO: rm_propagate takes the invariant-subspace kernel when it
 * applies (below); otherwise, and in rm_step, the ELL kernel (persistent
 * workgroups, operator staged in LDS when it fits) for n <= 4096, else the
 * global-memory kernel.
 *
 * RM_KERNEL_SUBSPACE (rm_propagate only): the rows are split into the
 * connected components of the operator's sparsity graph (invariant subspaces
 * of every H(F) the plan can form).  A wavepacket's Krylov vectors stay in
 * the union of the subspaces its initial support touches ("domain"), so each
 * wavepacket is advanced by one wave on its domain only, through all nsteps
 * steps in one launch, with the operator A(F) = sum_s F_s B_s assembled per
 * step from the field table.  Applies when no MF threshold is given; results
 * equal the full-vector kernels' to rounding (entries outside the domain stay
 * exactly zero).  Domains of more than 64 rows (RM_TUNE_WAVE_MAX_ROWS; at most
 * RM_SUBSPACE_MAX_ROWS) are advanced by a cooperative launch (many
 * workgroups on one wavepacket, grid barriers between the Lanczos
 * reductions) when there are at most 8 of them and each holds at least
 * their count x 6400 stored entries (RM_TUNE_COOP_MAX,
 * RM_TUNE_COOP_MIN_ENTRIES), otherwise by the block kernel (one workgroup
 * per wavepacket, all in one launch).
 *
 * RM_KERNEL_COOP, RM_KERNEL_BLOCK (rm_propagate only): every wavepacket
 * through the cooperative, or the block, kernel whatever its domain (no MF
 * thresholds). */
#define RM_KERNEL_AUTO      0
#define RM_KERNEL_GLOBAL    1
#define RM_KERNEL_RESIDENT  2
#define RM_KERNEL_ELL       3
#define RM_KERNEL_SUBSPACE  4
#define RM_KERNEL_COOP      5
#define RM_KERNEL_BLOCK     6
#define RM_SUBSPACE_MAX_ROWS 1024
#define RM_BLOCK_MAX_ROWS   65536   /* longest domain of an rm_batch_propagate job */
/* RM_KERNEL_ELL_CFG0 + i: ELL kernel with explicit launch configuration i
 * (rm_ell_config() describes it); the autotuner's search space. */
#define RM_KERNEL_ELL_CFG0  16
int rm_plan_set_kernel(rm_plan* plan, int32_t variant);

/* Launch choices of rm_propagate's invariant-subspace path, for an autotuner
 * over the production kernels (no counterpart in richmol, whose step has one
 * code path; callers that do not tune leave the defaults):
 *  RM_TUNE_PAIR      1 (default): two wavepackets per wave on domains of
 *                    <= 32 rows under a fixed Krylov dimension <= 16;
 *                    0: one wave each;
 *  RM_TUNE_COOP_MAX  long-domain wavepackets per call up to which the
 *                    cooperative form runs (default 8; more go one workgroup
 *                    each);
 *  RM_TUNE_COOP_GRID workgroups of a cooperative launch at most (default 0:
 *                    one per 64-row group of the domain, up to one per CU);
 *  RM_TUNE_FORK      1 (default): the launch classes of a call side by side
 *                    on auxiliary streams; 0: one after another;
 *  RM_TUNE_WAVE_MAX_ROWS  domains longer than this (default 64, at most
 *                    RM_SUBSPACE_MAX_ROWS) take the long-domain forms;
 *  RM_TUNE_COOP_MIN_ENTRIES  the cooperative form only while every long
 *                    domain holds at least (their count) x this many stored
 *                    entries (default 6400: it runs them one at a time).
 * Results do not depend on these choices beyond summation order.  A change
 * drops the launch lists kept for domain reuse. */
#define RM_TUNE_PAIR       0
#define RM_TUNE_COOP_MAX   1
#define RM_TUNE_COOP_GRID  2
#define RM_TUNE_FORK       3
#define RM_TUNE_WAVE_MAX_ROWS    4
#define RM_TUNE_COOP_MIN_ENTRIES 5
int rm_plan_set_tuning(rm_plan* plan, int32_t knob, int32_t value);

/* Launch configuration i of the ELL kernel: threads per workgroup, rows per
 * thread, wavepackets per workgroup, operator staged in LDS (1/0).
 * Returns RM_ERR_ARG when i is out of range. */
int rm_ell_config(int32_t i, int32_t* threads, int32_t* rows, int32_t* cols, int32_t* op_lds);

/* Invariant subspaces of the plan's operator (connected components of its
 * sparsity graph, built on first call): count and largest size. */
int rm_plan_subspaces(rm_plan* plan, int64_t* count, int64_t* max_rows);

/* Arithmetic of the bulk matvec in the invariant-subspace kernel (no
 * reference counterpart).  RM_PREC_FP64: complex128 throughout.
 * RM_PREC_FP32: the per-step operator values and the gathered Krylov vector
 * are rounded to complex64 and the matvec accumulates in FP32; Krylov
 * recurrence, norms and dots, exp(fT), the split-operator phases and the
 * wavepacket itself stay FP64. */
#define RM_PREC_FP64 0
#define RM_PREC_FP32 1
int rm_plan_set_precision(rm_plan* plan, int32_t precision);

/*
 * Domain of each column for the subspace kernel: lohi[2b], lohi[2b+1] =
 * [lo, hi) in the plan's permuted row order, the union of the invariant
 * subspaces the column's non-zero entries touch (lo >= hi: zero column; a
 * non-finite entry gives the whole basis).  psi: device; lohi: host, 2*ncols.
 * Synchronises the stream.
 */
int rm_plan_domains(rm_plan* plan, const void* psi, int64_t ld, int64_t ncols,
                    int32_t* lohi, void* stream);

/*
 * Domain reuse (no reference counterpart: removes the per-call device-to-host
 * read of the column domains).  With on != 0, rm_propagate keeps the domains
 * and launch lists it computed and reuses them on its next call with the same
 * psi, ld and ncols -- propagation keeps every column inside its domain.  A
 * reused call first checks on the device that every non-zero entry of psi
 * still lies in its column's domain; where it does not (psi was rewritten
 * between the calls), no column is propagated, psi is left as it was and the
 * columns get RM_STALE_DOMAIN in status.  Every call of this function, with
 * either value, forgets the kept domains.  Default: off.
 */
int rm_plan_set_domain_reuse(rm_plan* plan, int32_t on);

/* Test hook (no reference counterpart): with preset != 0 every cooperative
 * launch of the plan starts with its abort flag raised, so its bounded waits
 * give up at once, the launch runs to its end and its columns get
 * RM_DEVICE_TIMEOUT -- the timeout path without a timeout.  Default 0. */
int rm_plan_set_coop_abort(rm_plan* plan, int32_t preset);

/* sflag[s] = 1 if step s (row s of fprod_dev, nsteps x nfield complex,
 * device) makes any MF element non-zero, else 0: the `len(H.mfmat) > 0`
 * test of tdse.py:377 without thresholds.  sflag: device int32[nsteps]. */
int rm_plan_step_flags(rm_plan* plan, int64_t nsteps, const double* fprod_dev,
                       int32_t* sflag, void* stream);

/*
 * Batches of independent propagation jobs (no reference counterpart: the
 * serving loop's device call).  One rm_batch_propagate call advances every
 * job by its own number of steps with the subspace kernels: one wave per job
 * on a domain of up to RM_SUBSPACE_MAX_ROWS rows, one workgroup per job on a
 * longer one, all jobs in as many launches as there are launch classes
 * (domain size and operator width; a job's result does not depend on the
 * other jobs of the call, bit for bit).  Jobs may share plans; each has its
 * own wavepacket, phases and field table.  Calls on one batch object must
 * use one stream (descriptor staging is double-buffered and reused in call
 * order).
 */
typedef struct rm_batch rm_batch;
typedef struct {
  rm_plan* plan;             /* operator (only its subspace arrays are read) */
  void* psi;                 /* device: the wavepacket is column `col` at psi + col*ld */
  int64_t ld;
  int32_t col;
  int32_t lo, len;           /* its domain, from rm_plan_domains (len <= RM_BLOCK_MAX_ROWS) */
  const double* fprod_dev;   /* device: field products of this call's steps, nsteps x nfield complex */
  const int32_t* sflag_dev;  /* device: rm_plan_step_flags of those rows */
  int32_t nsteps;
  double f_re, f_im;         /* exponent factor of the step (TDSE.update, tdse.py:336-337) */
  const void* phase;         /* device: n complex split-operator phases (rm_phase_init) or NULL */
  int32_t* kused;            /* device int32[nsteps] or NULL: Krylov dimension per step */
  int32_t* status;           /* device int32[1] or NULL: set to RM_NOT_CONVERGED on failure */
} rm_job;
int rm_batch_create(rm_batch** out);
int rm_batch_destroy(rm_batch* batch);
/* Jobs whose domain has more than `rows` rows run on the block kernel (one
 * workgroup per job) instead of one wave; 0 <= rows <= RM_SUBSPACE_MAX_ROWS,
 * default RM_SUBSPACE_MAX_ROWS.  The class of a job still depends only on the
 * job, so results stay independent of the mix. */
int rm_batch_set_block_rows(rm_batch* batch, int32_t rows);
/* job_bytes: sizeof(rm_job) as the caller laid the array out; any other
 * value is rejected (RM_ERR_ARG) before a job is read. */
int rm_batch_propagate(rm_batch* batch, const rm_job* jobs, int32_t njobs, int32_t job_bytes,
                       int32_t krylov_fixed, int32_t maxorder, double tol, int32_t precision,
                       void* stream);

/*
 * rm_batch_propagate without allocating (no counterpart in the reference: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * jobs, kernels and results as rm_batch_propagate, with the job descriptors
 * and the Krylov workspace in a fixed device array of the library
 * (RM_STATIC_SCRATCH_BYTES per GPU, held while the process runs) instead of
 * memory the call allocates: no hipMalloc/hipFree, no pinned memory, no
 * streams or events.  Limits, each refused with RM_ERR_ARG and a message:
 * at most RM_STATIC_MAX_JOBS jobs; every job's domain on the one-wave kernels
 * (<= RM_SUBSPACE_MAX_ROWS rows, no block-form jobs); every plan's subspace
 * arrays built beforehand (rm_plan_subspaces, or any earlier propagate call on
 * the plan).  The launch classes run one after another on `stream`, with as
 * many waves as the workspace holds (fewer than rm_batch_propagate may use
 * for Krylov bases outside LDS or the adaptive rule's Gram matrices; the
 * results do not depend on the wave count).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller.
 */
#define RM_STATIC_SCRATCH_BYTES (32 << 20)
#define RM_STATIC_MAX_JOBS 4096
int rm_batch_propagate_static(const rm_job* jobs, int32_t njobs, int32_t job_bytes, int32_t krylov_fixed,
                              int32_t maxorder, double tol, int32_t precision, void* stream);

/* psi_out[:, b] = (psi_in[:, b] * phase) * phase: the step when the operator
 * carries no field at all (tdse.py:375, 394 with the Lanczos skipped). */
int rm_phase_apply(int64_t n, const void* psi_in, void* psi_out, int64_t ld, int64_t ncols,
                   const void* phase, void* stream);

/* out[b] = <psi_b| H |psi_b> (complex, device 2*ncols doubles), the
 * quadratic form of the observable loops in examples/ocs_alignment.py:99-100. */
int rm_expect(rm_plan* plan, const void* psi, int64_t ld, int64_t ncols,
              void* out, void* stream);

/*
 * Kronecker-block operators for large rotor bases (the factored form of
 * CarTens.vec, field.py:1145-1245: Y[J1] += MF_w X[J2] K_w^T per block and
 * irrep w).  The basis has every J of `Js` (ascending) with all m; a
 * wavepacket is stored per J as the (2J+1) x (2J+1) matrix X[J][m][k]
 * (m-major, k over all species), J blocks in the order of `Js` -- the
 * "internal order" below.  Block q couples J1[q] <- J2[q]:
 *   K:     complex (2J1+1) x (2J2+1) row-major at K + 2*k_off[q] (doubles);
 *   sigma: its m-diagonals sig[sig_off[q] .. sig_off[q+1]) (m2 = m1 - sigma,
 *          |sigma| <= 2);
 *   mdiag: complex [nsig][ncart[q]][2J1+1] at mdiag + 2*d_off[q]: the values
 *          of M_w,cart on those diagonals (indexed by m1);
 *   slot0: the field products of its Cartesian components are slots
 *          slot0[q] .. slot0[q]+ncart[q]-1 of a step's nslot products.
 * Per step MF_q = sum_c fprod[slot0+c] M_q,c (CarTens.field,
 * field.py:1073-1142); per output J1 the stacked MF_q X[J2] of every
 * wavepacket go through one dense product with the stacked K_q (rocBLAS).
 */
typedef struct rm_kplan rm_kplan;
int rm_kplan_create(int32_t nJ, const int32_t* Js, int32_t nslot, int32_t nblk,
                    const int32_t* J1, const int32_t* J2, const int32_t* slot0,
                    const int32_t* ncart, const int64_t* k_off, const double* K,
                    const int64_t* sig_off, const int32_t* sig, const int64_t* d_off,
                    const double* mdiag, rm_kplan** out);
int rm_kplan_destroy(rm_kplan* plan);
/* n: states per wavepacket; gemm_flops_per_col: real flops of one matvec's
 * dense products per wavepacket. */
int rm_kplan_info(const rm_kplan* plan, int64_t* n, double* gemm_flops_per_col);
/* Matvec form (no reference counterpart).  RM_KKERNEL_GEMM (default): the
 * m-mixed input (MF_w X) stored, then one rocBLAS product per output J, the
 * build of the next J chunk overlapping the products of the last.
 * RM_KKERNEL_FUSED: one launch per matvec on the matrix cores, the m-mixed
 * input built in LDS chunk by chunk inside the product (no stored copy).
 * RM_KKERNEL_MFMA: the stored form with the library's own matrix-core
 * product in place of rocBLAS (one launch per J chunk; FP64 64 x 128 output
 * tiles, FP32 128 x 128). */
#define RM_KKERNEL_FUSED 0
#define RM_KKERNEL_GEMM  1
#define RM_KKERNEL_MFMA  2
int rm_kplan_set_kernel(rm_kplan* plan, int32_t kernel);
/* y = H(F) x for ncols wavepackets (device, complex, row b at b*ld, internal
 * order); fprod_dev: device, nslot complex.  precision RM_PREC_FP32: the
 * products in FP32 (complex64 operands, FP32 accumulation). */
int rm_kplan_matvec(rm_kplan* plan, const void* x, void* y, int64_t ld, int32_t ncols,
                    const double* fprod_dev, int32_t precision, void* stream);
/*
 * nsteps steps of TDSE.update (tdse.py:365-414) for ncols wavepackets (rows
 * of psi, in place).  krylov_fixed > 0: that fixed Krylov dimension;
 * krylov_fixed == 0: the reference's adaptive rule (tdse.py:417-486) with
 * maxorder (<= RM_MAX_KRYLOV) and tol, every wavepacket stopping at its own
 * dimension (the batch goes on until all have stopped; one device-to-host
 * read of a counter per Krylov iteration).  An exact breakdown continues with
 * a zero vector whose coefficient in exp(fT)e0 is zero, as the reference's
 * restart vector's is (and whose Gram column is the restart vector's).
 * fprod: HOST, nsteps x nslot complex; a step whose products are all zero
 * applies the phases only (the reference skips the Lanczos when the field
 * leaves no MF block).  phase: device, n complex, internal order (from
 * rm_phase_init) or NULL.  kused: device int32[nsteps x ncols] or NULL
 * (dimension per step and wavepacket, 0 on phase-only steps); status: device
 * int32[ncols] or NULL, zeroed, RM_NOT_CONVERGED where a step reached
 * maxorder.  RM_PREC_FP32: the dense products in FP32; the recurrence,
 * norms, exp(fT)e0, phases and the wavepackets stay FP64.
 */
int rm_kplan_propagate(rm_kplan* plan, void* psi, int64_t ld, int32_t ncols, int64_t nsteps,
                       const double* fprod, double f_re, double f_im, const void* phase,
                       int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                       int32_t* status, int32_t precision, void* stream);

/*
 * rm_kplan_propagate with an observable read after every step (the loop of
 * examples/ocs_alignment.py:89-103 on the Kronecker-block form): obs, a
 * Kronecker-block plan over the same J list, with obs_fprod (HOST, obs's
 * nslot complex products, fixed for the run, e.g. those of the field
 * [0, 0, 1] for cos^2 theta).  After step s, Y = O psi (FP64 products,
 * whatever `precision`) and
 *   expect: device complex [nsteps][ncols] (or NULL), <psi_b|O|psi_b>;
 *   expect_sum: device complex [nsteps] (or NULL), their sum over b;
 * at least one of them non-NULL.  obs holds a workspace of its own for the
 * products (as large as a matvec's of this many wavepackets).
 */
int rm_kplan_propagate_expect(rm_kplan* plan, void* psi, int64_t ld, int32_t ncols, int64_t nsteps,
                              const double* fprod, double f_re, double f_im, const void* phase,
                              int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                              int32_t* status, int32_t precision, rm_kplan* obs, const double* obs_fprod,
                              void* expect, void* expect_sum, void* stream);

/*
 * Rotational densities of wavepackets on a product grid of Euler angles (the
 * at-scale use of richmol/rotdens.py:10-283): with a^J_mk the wavepacket's
 * coefficients over the primitive symmetric-top functions (including
 * sqrt((2J+1)/(8 pi^2))),
 *   psi(alpha, beta, gamma) = sum_J sum_mk a^J_mk e^{i m alpha} d^J_mk(beta) e^{i k gamma},
 * rho = |psi|^2 sin(beta).  m, k run over -jmax..jmax (M = 2 jmax + 1):
 *   coef:   device complex [nwp][jmax+1][M][M] (zero where |m| or |k| > J);
 *   ealpha: device complex [na][M], e^{i m alpha_a};
 *   dtab:   device double [jmax+1][nb][M][M], d^J_mk(beta_b);
 *   sinb:   device double [nb];  egamma: device complex [ng][M], e^{i k gamma_g};
 *   rho:    device double, [na][nb][ng] summed over the wavepackets (the
 *           ensemble's density; per_wavepacket 0) or [nwp][na][nb][ng].
 */
int rm_rotdens(int32_t jmax, int32_t nwp, const void* coef, int32_t na, const void* ealpha, int32_t nb,
               const double* dtab, const double* sinb, int32_t ng, const void* egamma,
               int32_t per_wavepacket, void* rho, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RICHMOL_AMD_H */
