// richmol_amd: gfx950 kernels and C ABI for the rotational TDSE step.
//
// Operator (per plan, see include/richmol_amd.h):
//   MF_t = thr( sum_q fprod[field_q] * M_q )        k_field_contract
//   A_e  = sum_p K_p * MF[term_p]                   k_assemble
//   y    = A x  (CSR, complex128)                   spmv_row / k_spmv
// Step (per wavepacket = one workgroup):            k_lanczos_step
//   phi = P*psi, Lanczos expv of phi (tdse.py:417-486), out = P*phi.
//
// Numerics: FP64 everywhere.  Reductions are fixed-order (wave butterfly then
// waves in index order) so a column's result does not depend on how many
// other columns share the launch.

#include <string>

#include <atomic>

#include "rm_common.h"

// ---------------------------------------------------------------------------
// errors

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* rm_last_error(void) { return g_err; }
extern "C" int rm_version(void) { return RM_ABI_VERSION; }

// ---------------------------------------------------------------------------
// plan

static void plan_free(rm_plan* p) {
  void* bufs[] = {p->row_ptr, p->col, p->pair_ptr, p->pair_term, p->pair_kval, p->term_ptr,
                  p->term_field, p->term_mval, p->term_leaf, p->mf, p->aval, p->flag, p->ws,
                  p->ell_col, p->ell_val, p->ell_pos, p->sub_perm, p->sub_label, p->sub_lo,
                  p->sub_hi, p->sub_col, p->sub_val, p->sub_rp, p->sub_ccol, p->sub_cval, p->sub_cval32, p->dom,
                  p->sflag, p->sflag_part, p->sflag_arrive, p->dom_keep, p->stale, p->int_status,
                  p->sub_iperm, p->obs_col, p->obs_val, p->ex_buf, p->expo_ws, p->expo_anorm,
                  p->expo_items};
  sub_scratch_release(p->sub_scratch);
  if (p->h_dom_pin) (void)hipHostFree(p->h_dom_pin);
  if (p->h_status_pin) (void)hipHostFree(p->h_status_pin);
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : p->ev0) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->ev1) (void)hipEventDestroy(e);
  delete p;
}

extern "C" int rm_plan_create(int64_t n, int64_t nnz, const int64_t* row_ptr, const int32_t* col_idx,
                              const int64_t* pair_ptr, const int32_t* pair_term,
                              const double* pair_kval, int64_t nterm, const int64_t* term_ptr,
                              const int32_t* term_field, const double* term_mval,
                              const int32_t* term_leaf, int32_t nfield, int32_t nleaf,
                              rm_plan** out) {
  if (!out) return fail(RM_ERR_ARG, "rm_plan_create: out is NULL");
  *out = nullptr;
  if (n < 0 || nnz < 0 || nterm < 0) return fail(RM_ERR_ARG, "rm_plan_create: negative size");
  if (n > INT32_MAX) return fail(RM_ERR_ARG, "rm_plan_create: n=%lld exceeds int32 column index", (long long)n);
  if (nfield < 0 || nfield > RM_MAX_FIELD)
    return fail(RM_ERR_ARG, "rm_plan_create: nfield=%d outside [0, %d]", nfield, RM_MAX_FIELD);
  if (nleaf < 0 || nleaf > RM_MAX_LEAF)
    return fail(RM_ERR_ARG, "rm_plan_create: nleaf=%d outside [0, %d]", nleaf, RM_MAX_LEAF);
  if (!row_ptr || (nnz > 0 && (!col_idx || !pair_ptr)) || (nterm > 0 && !term_ptr))
    return fail(RM_ERR_ARG, "rm_plan_create: NULL array");
  // host-side validation of every index the kernels will follow
  if (row_ptr[0] != 0 || row_ptr[n] != nnz) return fail(RM_ERR_ARG, "rm_plan_create: bad row_ptr ends");
  for (int64_t i = 0; i < n; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) return fail(RM_ERR_ARG, "rm_plan_create: row_ptr not monotone");
  for (int64_t e = 0; e < nnz; ++e)
    if (col_idx[e] < 0 || col_idx[e] >= n) return fail(RM_ERR_ARG, "rm_plan_create: col_idx out of range");
  int64_t npair = nnz > 0 ? pair_ptr[nnz] : 0;
  if (nnz > 0 && pair_ptr[0] != 0) return fail(RM_ERR_ARG, "rm_plan_create: bad pair_ptr");
  for (int64_t e = 0; e < nnz; ++e)
    if (pair_ptr[e + 1] < pair_ptr[e]) return fail(RM_ERR_ARG, "rm_plan_create: pair_ptr not monotone");
  for (int64_t p = 0; p < npair; ++p)
    if (pair_term[p] < 0 || pair_term[p] >= nterm) return fail(RM_ERR_ARG, "rm_plan_create: pair_term out of range");
  int64_t nq = nterm > 0 ? term_ptr[nterm] : 0;
  if (nterm > 0 && term_ptr[0] != 0) return fail(RM_ERR_ARG, "rm_plan_create: bad term_ptr");
  for (int64_t t = 0; t < nterm; ++t) {
    if (term_ptr[t + 1] < term_ptr[t]) return fail(RM_ERR_ARG, "rm_plan_create: term_ptr not monotone");
    if (term_leaf[t] < 0 || term_leaf[t] >= nleaf) return fail(RM_ERR_ARG, "rm_plan_create: term_leaf out of range");
  }
  for (int64_t q = 0; q < nq; ++q)
    if (term_field[q] < 0 || term_field[q] >= nfield) return fail(RM_ERR_ARG, "rm_plan_create: term_field out of range");

  rm_plan* p = new (std::nothrow) rm_plan();
  if (!p) return fail(RM_ERR_ALLOC, "rm_plan_create: host allocation");
  p->n = n; p->nnz = nnz; p->npair = npair; p->nterm = nterm; p->nq = nq;
  p->nfield = nfield; p->nleaf = nleaf;
  p->h_row_ptr.assign(row_ptr, row_ptr + n + 1);
  p->h_col.assign(col_idx, col_idx + nnz);
  p->h_pair_ptr.assign(pair_ptr, pair_ptr + (nnz > 0 ? nnz + 1 : 0));
  p->h_pair_term.assign(pair_term, pair_term + npair);
  p->h_pair_kval.assign((const double2*)pair_kval, (const double2*)pair_kval + npair);
  p->h_term_ptr.assign(term_ptr, term_ptr + (nterm > 0 ? nterm + 1 : 0));
  p->h_term_field.assign(term_field, term_field + nq);
  p->h_term_mval.assign((const double2*)term_mval, (const double2*)term_mval + nq);
  if (hipGetDevice(&p->device) != hipSuccess) { plan_free(p); return fail(RM_ERR_HIP, "hipGetDevice failed"); }
  int rc = RM_OK;
  if ((rc = upload(&p->row_ptr, row_ptr, n + 1)) ||
      (rc = upload(&p->col, col_idx, nnz)) ||
      (rc = upload(&p->pair_ptr, pair_ptr, nnz + 1)) ||
      (rc = upload(&p->pair_term, pair_term, npair)) ||
      (rc = upload(&p->pair_kval, (const double2*)pair_kval, npair)) ||
      (rc = upload(&p->term_ptr, term_ptr, nterm + 1)) ||
      (rc = upload(&p->term_field, term_field, nq)) ||
      (rc = upload(&p->term_mval, (const double2*)term_mval, nq)) ||
      (rc = upload(&p->term_leaf, term_leaf, nterm)) ||
      (rc = upload<double2>(&p->mf, nullptr, nterm)) ||
      (rc = upload<double2>(&p->aval, nullptr, nnz)) ||
      (rc = upload<int32_t>(&p->flag, nullptr, 1))) {
    plan_free(p);
    return rc;
  }
  if (hipMemset(p->aval, 0, sizeof(double2) * (size_t)(nnz > 0 ? nnz : 1)) != hipSuccess ||
      hipMemset(p->flag, 0xff, sizeof(int32_t)) != hipSuccess) {
    plan_free(p);
    return fail(RM_ERR_HIP, "rm_plan_create: hipMemset failed");
  }
  // ELL layout
  int64_t w = 0;
  for (int64_t i = 0; i < n; ++i) w = row_ptr[i + 1] - row_ptr[i] > w ? row_ptr[i + 1] - row_ptr[i] : w;
  if (w > INT32_MAX / 2 || (n > 0 && w * n > INT32_MAX)) {
    plan_free(p);
    return fail(RM_ERR_ARG, "rm_plan_create: ELL layout (width %lld) too large", (long long)w);
  }
  p->ell_w = (int32_t)w;
  {
    const int64_t ne = w * n;
    int32_t* hcol = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ne > 0 ? ne : 1));
    int32_t* hpos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    if (!hcol || !hpos) {
      free(hcol);
      free(hpos);
      plan_free(p);
      return fail(RM_ERR_ALLOC, "rm_plan_create: host allocation");
    }
    for (int64_t i = 0; i < n; ++i) {
      const int64_t len = row_ptr[i + 1] - row_ptr[i];
      for (int64_t j = 0; j < w; ++j) {
        if (j < len) {
          hcol[j * n + i] = col_idx[row_ptr[i] + j];
          hpos[row_ptr[i] + j] = (int32_t)(j * n + i);
        } else {
          hcol[j * n + i] = (int32_t)i;
        }
      }
    }
    rc = upload(&p->ell_col, hcol, ne);
    if (!rc) rc = upload(&p->ell_pos, hpos, nnz);
    if (!rc) rc = upload<double2>(&p->ell_val, nullptr, ne);
    free(hcol);
    free(hpos);
    if (rc) {
      plan_free(p);
      return rc;
    }
    if (hipMemset(p->ell_val, 0, sizeof(double2) * (size_t)(ne > 0 ? ne : 1)) != hipSuccess) {
      plan_free(p);
      return fail(RM_ERR_HIP, "rm_plan_create: hipMemset failed");
    }
  }
  if (hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
      p->n_cu <= 0)
    p->n_cu = 256;
  *out = p;
  return RM_OK;
}

extern "C" int rm_plan_destroy(rm_plan* plan) {
  if (!plan) return RM_OK;
  (void)hipDeviceSynchronize();
  plan_free(plan);
  return RM_OK;
}

extern "C" int rm_plan_info(const rm_plan* plan, int64_t* n, int64_t* nnz) {
  if (!plan) return fail(RM_ERR_ARG, "rm_plan_info: NULL plan");
  if (n) *n = plan->n;
  if (nnz) *nnz = plan->nnz;
  return RM_OK;
}

// ---------------------------------------------------------------------------
// field contraction (CarTens.field) and assembly

struct FieldArgs {
  double2 fprod[RM_MAX_FIELD];
  double thresh[RM_MAX_LEAF];
};

// fdev: field products of this step in device memory (a row of a table), or
// NULL to use the kernel-argument copy in fa.
__global__ void k_field_contract(int64_t nterm, const int64_t* __restrict__ term_ptr,
                                 const int32_t* __restrict__ term_field,
                                 const double2* __restrict__ term_mval,
                                 const int32_t* __restrict__ term_leaf, FieldArgs fa,
                                 const double2* __restrict__ fdev,
                                 double2* __restrict__ mf, int32_t* __restrict__ flag, int32_t id) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nterm) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t q = term_ptr[t]; q < term_ptr[t + 1]; ++q)
    acc = cfma(fdev ? fdev[term_field[q]] : fa.fprod[term_field[q]], term_mval[q], acc);
  double thr = fa.thresh[term_leaf[t]];
  if (thr > 0.0 && hypot(acc.x, acc.y) < thr) acc = make_double2(0.0, 0.0);
  mf[t] = acc;
  if (acc.x != 0.0 || acc.y != 0.0 || acc.x != acc.x || acc.y != acc.y) *flag = id;
}

__global__ void k_assemble(int64_t nnz, const int64_t* __restrict__ pair_ptr,
                           const int32_t* __restrict__ pair_term, const double2* __restrict__ pair_kval,
                           const double2* __restrict__ mf, double2* __restrict__ aval,
                           const int32_t* __restrict__ ell_pos, double2* __restrict__ ell_val) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t p = pair_ptr[e]; p < pair_ptr[e + 1]; ++p)
    acc = cfma(pair_kval[p], mf[pair_term[p]], acc);
  aval[e] = acc;
  ell_val[ell_pos[e]] = acc;
}

// Contraction and assembly in one pass: each entry evaluates the MF terms its
// pairs use (same arithmetic as k_field_contract + k_assemble, term by term).
// Chosen when terms are used about once (linear rotors: dk = 1), where the
// recomputation costs nothing and one launch per step is saved.
__global__ void k_field_assemble(int64_t nnz, const int64_t* __restrict__ pair_ptr,
                                 const int32_t* __restrict__ pair_term,
                                 const double2* __restrict__ pair_kval,
                                 const int64_t* __restrict__ term_ptr,
                                 const int32_t* __restrict__ term_field,
                                 const double2* __restrict__ term_mval,
                                 const int32_t* __restrict__ term_leaf, FieldArgs fa,
                                 const double2* __restrict__ fdev, double2* __restrict__ aval,
                                 const int32_t* __restrict__ ell_pos, double2* __restrict__ ell_val,
                                 int32_t* __restrict__ flag, int32_t id) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double2 acc = make_double2(0.0, 0.0);
  bool nz = false;
  for (int64_t p = pair_ptr[e]; p < pair_ptr[e + 1]; ++p) {
    const int32_t t = pair_term[p];
    double2 mf = make_double2(0.0, 0.0);
    for (int64_t q = term_ptr[t]; q < term_ptr[t + 1]; ++q)
      mf = cfma(fdev ? fdev[term_field[q]] : fa.fprod[term_field[q]], term_mval[q], mf);
    const double thr = fa.thresh[term_leaf[t]];
    if (thr > 0.0 && hypot(mf.x, mf.y) < thr) mf = make_double2(0.0, 0.0);
    nz |= (mf.x != 0.0 || mf.y != 0.0 || mf.x != mf.x || mf.y != mf.y);
    acc = cfma(pair_kval[p], mf, acc);
  }
  aval[e] = acc;
  ell_val[ell_pos[e]] = acc;
  if (nz) *flag = id;
}

// The field contraction of one step (fdev: device row of field products or
// NULL for fa's copy): fused when every term is used about once.
static void launch_field(rm_plan* p, const FieldArgs& fa, const double2* fdev, hipStream_t s) {
  if (p->nnz > 0 && p->npair <= 2 * p->nterm) {
    hipLaunchKernelGGL(k_field_assemble, dim3(grid_for(p->nnz, 256)), dim3(256), 0, s, p->nnz,
                       p->pair_ptr, p->pair_term, p->pair_kval, p->term_ptr, p->term_field,
                       p->term_mval, p->term_leaf, fa, fdev, p->aval, p->ell_pos, p->ell_val,
                       p->flag, p->step_id);
    return;
  }
  if (p->nterm > 0)
    hipLaunchKernelGGL(k_field_contract, dim3(grid_for(p->nterm, 256)), dim3(256), 0, s, p->nterm,
                       p->term_ptr, p->term_field, p->term_mval, p->term_leaf, fa, fdev, p->mf,
                       p->flag, p->step_id);
  if (p->nnz > 0)
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(p->nnz, 256)), dim3(256), 0, s, p->nnz,
                       p->pair_ptr, p->pair_term, p->pair_kval, p->mf, p->aval, p->ell_pos,
                       p->ell_val);
}

extern "C" int rm_plan_set_field(rm_plan* p, const double* fprod, const double* thresh, void* stream) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_field: NULL plan");
  if (p->nfield > 0 && !fprod) return fail(RM_ERR_ARG, "rm_plan_set_field: NULL fprod");
  FieldArgs fa;
  memset(&fa, 0, sizeof(fa));
  for (int i = 0; i < p->nfield; ++i) fa.fprod[i] = make_double2(fprod[2 * i], fprod[2 * i + 1]);
  for (int i = 0; i < RM_MAX_LEAF; ++i) fa.thresh[i] = -1.0;
  if (thresh)
    for (int i = 0; i < p->nleaf; ++i) fa.thresh[i] = thresh[i];
  hipStream_t s = (hipStream_t)stream;
  p->step_id = (p->step_id + 1) & 0x3fffffff;
  static std::atomic<int64_t> gen{0};
  p->field_gen = ++gen;
  launch_field(p, fa, nullptr, s);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

// ---------------------------------------------------------------------------
// sparse matvec

__device__ __forceinline__ double2 spmv_row(int64_t i, const int64_t* __restrict__ rp,
                                            const int32_t* __restrict__ col,
                                            const double2* __restrict__ a,
                                            const double2* __restrict__ x) {
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t e = rp[i]; e < rp[i + 1]; ++e) acc = cfma(a[e], x[col[e]], acc);
  return acc;
}

__global__ void k_spmv(int64_t n, int64_t ld, const int64_t* __restrict__ rp,
                       const int32_t* __restrict__ col, const double2* __restrict__ a,
                       const double2* __restrict__ x, double2* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t b = blockIdx.y;
  if (i >= n) return;
  y[b * ld + i] = spmv_row(i, rp, col, a, x + b * ld);
}

extern "C" int rm_matvec(rm_plan* p, const void* x, void* y, int64_t ld, int64_t ncols, void* stream) {
  if (!p) return fail(RM_ERR_ARG, "rm_matvec: NULL plan");
  if (ncols < 0 || ld < p->n) return fail(RM_ERR_ARG, "rm_matvec: bad ld/ncols");
  if (ncols > 65535) return fail(RM_ERR_ARG, "rm_matvec: ncols > 65535");
  if (ncols == 0 || p->n == 0) return RM_OK;
  if (!x || !y) return fail(RM_ERR_ARG, "rm_matvec: NULL vector");
  if (x == y) return fail(RM_ERR_ARG, "rm_matvec: x and y must not alias");
  hipLaunchKernelGGL(k_spmv, dim3(grid_for(p->n, 256), (unsigned)ncols), dim3(256), 0,
                     (hipStream_t)stream, p->n, ld, p->row_ptr, p->col, p->aval,
                     (const double2*)x, (double2*)y);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

// ---------------------------------------------------------------------------
// split-operator phases

__global__ void k_phase_init(int64_t n, const double2* __restrict__ d, double2 half_f,
                             double2* __restrict__ ph) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 z = cmul(half_f, d[i]);
  double m = exp(z.x);
  double sn, cs;
  sincos(z.y, &sn, &cs);
  ph[i] = make_double2(m * cs, m * sn);
}

extern "C" int rm_phase_init(int64_t n, const void* diag, double f_re, double f_im, void* phase,
                             void* stream) {
  if (n < 0) return fail(RM_ERR_ARG, "rm_phase_init: n < 0");
  if (n == 0) return RM_OK;
  if (!diag || !phase) return fail(RM_ERR_ARG, "rm_phase_init: NULL pointer");
  hipLaunchKernelGGL(k_phase_init, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                     (const double2*)diag, make_double2(0.5 * f_re, 0.5 * f_im), (double2*)phase);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

__global__ void k_phase_apply(int64_t n, int64_t ld, const double2* __restrict__ in,
                              double2* __restrict__ out, const double2* __restrict__ ph) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t b = blockIdx.y;
  if (i >= n) return;
  double2 p = ph[i];
  out[b * ld + i] = cmul(cmul(in[b * ld + i], p), p);
}

extern "C" int rm_phase_apply(int64_t n, const void* psi_in, void* psi_out, int64_t ld,
                              int64_t ncols, const void* phase, void* stream) {
  if (n < 0 || ncols < 0 || ld < n) return fail(RM_ERR_ARG, "rm_phase_apply: bad sizes");
  if (ncols > 65535) return fail(RM_ERR_ARG, "rm_phase_apply: ncols > 65535");
  if (n == 0 || ncols == 0) return RM_OK;
  if (!psi_in || !psi_out || !phase) return fail(RM_ERR_ARG, "rm_phase_apply: NULL pointer");
  hipLaunchKernelGGL(k_phase_apply, dim3(grid_for(n, 256), (unsigned)ncols), dim3(256), 0,
                     (hipStream_t)stream, n, ld, (const double2*)psi_in, (double2*)psi_out,
                     (const double2*)phase);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

// ---------------------------------------------------------------------------
// Lanczos step: one workgroup per wavepacket

namespace {

constexpr int NT = 256;
constexpr int NWAVE = NT / 64;

struct StepArgs {
  int64_t n, ld;
  const double2* psi_in;
  double2* psi_out;
  const double2* phase;
  const int64_t* row_ptr;
  const int32_t* col;
  const double2* aval;
  const int32_t* flag;
  int32_t flag_expect;
  double2 f;
  int32_t fixed_m, maxorder;
  double tol;
  double2* basis;   // per column kcap*n
  double2* wbuf;    // per column n
  double2* gram;    // per column kcap*kcap
  int32_t kcap;
  int32_t* kused;
  int32_t* status;
  // ELL kernel
  const int32_t* ell_col;
  const double2* ell_val;
  int32_t ell_w;
  int32_t ncols;
  double2* cscr;    // per column 2*kcap: exp(fT)e0 and its previous value
  int32_t status_sticky;   // 1: only record failures (multi-step runs)
};

// Per-column status of a step: RM_NOT_CONVERGED when the Krylov loop hit
// maxorder; in sticky mode a success does not overwrite an earlier failure.
__device__ __forceinline__ void put_status(const StepArgs& a, int b, bool bad) {
  if (a.status && (bad || !a.status_sticky)) a.status[b] = bad ? RM_NOT_CONVERGED : 0;
}

// Workgroup sum of NV doubles with DPP wave sums (red: [2][NW][16]).
template <int NV, int NW>
__device__ __forceinline__ void block_sum_dpp(double* v, double (*red)[NW][16], int& buf) {
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_allsum(v[i]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[buf][w][i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NW; ++j) s += red[buf][j][i];
    v[i] = s;
  }
  buf ^= 1;
}

// Sum NV doubles over the workgroup (NW waves); every thread receives the same
// value.  red is double-buffered: one barrier per call suffices because every
// thread reads red[buf] before it can reach the barrier of the next call.
template <int NV, int NW>
__device__ __forceinline__ void block_sum(double* v, double (*red)[NW][4], int& buf) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += __shfl_xor(v[i], off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[buf][w][i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NW; ++j) s += red[buf][j][i];
    v[i] = s;
  }
  buf ^= 1;
}

// Resident variant: the current Krylov vector lives in LDS (gathered by the
// sparse matvec), the recurrence vectors V_{k-1}, V_k, W in registers (R rows
// per thread, rows tid + r*NTR), the basis is written behind to global memory
// and read once for the final combination (and by the breakdown restart).
// One barrier separates each LDS overwrite from the gathers of the previous
// iteration: the alpha reduction (block_sum) that follows every gather.
template <int NTR, int R>
__global__ __launch_bounds__(NTR) void k_lanczos_resident(StepArgs a) {
  constexpr int NW = NTR / 64;
  extern __shared__ double2 s_x[];
  __shared__ double red[2][NW][4];
  __shared__ double2 s_alpha[KMAX];
  __shared__ double s_beta[KMAX + 1];
  __shared__ double2 s_c[KMAX];
  __shared__ double2 s_cprev[KMAX];
  __shared__ double s_conv;
  int buf = 0;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int n = (int)a.n;
  const double2* psi = a.psi_in + (int64_t)b * a.ld;
  double2* out = a.psi_out + (int64_t)b * a.ld;
  double2* V = a.basis + (int64_t)b * a.kcap * a.n;
  double2* G = a.gram + (int64_t)b * a.kcap * a.kcap;
  const int kc = a.kcap;
  const bool fixed = a.fixed_m > 0;
  const bool skip = a.flag != nullptr && (*a.flag != a.flag_expect);
  const double2 zero = make_double2(0.0, 0.0);

  double2 vk[R], vp[R], W[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * NTR;
    vk[r] = zero;
    vp[r] = zero;
    W[r] = zero;
    if (i < n) {
      double2 v = psi[i];
      if (a.phase) v = cmul(v, a.phase[i]);
      if (skip) {
        out[i] = a.phase ? cmul(v, a.phase[i]) : v;
      } else {
        vk[r] = v;
        V[i] = v;
        s_x[i] = v;
      }
    }
  }
  if (skip) {
    if (tid == 0) {
      if (a.kused) a.kused[b] = 0;
      put_status(a, b, false);
    }
    return;
  }
  __syncthreads();

  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * NTR;
    if (i < n) {
      double2 w = spmv_row(i, a.row_ptr, a.col, a.aval, s_x);
      W[r] = w;
      acc[0] += w.x * vk[r].x + w.y * vk[r].y;
      acc[1] += w.x * vk[r].y - w.y * vk[r].x;
      acc[2] += vk[r].x * vk[r].x + vk[r].y * vk[r].y;
    }
  }
  block_sum<3>(acc, red, buf);
  double2 alpha = make_double2(acc[0], acc[1]);
#pragma unroll
  for (int r = 0; r < R; ++r) W[r] = csub(W[r], cmul(alpha, vk[r]));
  if (tid == 0) {
    s_alpha[0] = alpha;
    s_cprev[0] = make_double2(1.0, 0.0);
    s_c[0] = make_double2(1.0, 0.0);
    if (!fixed) G[0] = make_double2(acc[2], 0.0);
  }

  int k = 1;
  double conv = 1.0;
  for (;;) {
    if (fixed) {
      if (k >= a.fixed_m) break;
    } else if (!(k < a.maxorder && conv > a.tol)) {
      break;
    }
    acc[0] = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) acc[0] += W[r].x * W[r].x + W[r].y * W[r].y;
    block_sum<1>(acc, red, buf);
    const double beta = sqrt(acc[0]);
    const bool breakdown = !(beta != 0.0);
#pragma unroll
    for (int r = 0; r < R; ++r) vp[r] = vk[r];
    if (!breakdown) {
#pragma unroll
      for (int r = 0; r < R; ++r) vk[r] = make_double2(W[r].x / beta, W[r].y / beta);
    } else {
      // exact breakdown: Gram-Schmidt of a vector of ones against the stored
      // basis (tdse.py:459-465); each thread owns its rows of every V_j
#pragma unroll
      for (int r = 0; r < R; ++r) W[r] = (tid + r * NTR < n) ? make_double2(1.0, 0.0) : zero;
      for (int j = 0; j < k; ++j) {
        const double2* Vj = V + (int64_t)j * n;
        acc[0] = acc[1] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * NTR;
          if (i < n) {
            double2 x = Vj[i];
            acc[0] += x.x * W[r].x + x.y * W[r].y;
            acc[1] += x.x * W[r].y - x.y * W[r].x;
          }
        }
        block_sum<2>(acc, red, buf);
        const double2 proj = make_double2(acc[0], acc[1]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * NTR;
          if (i < n) W[r] = csub(W[r], cmul(proj, Vj[i]));
        }
      }
      acc[0] = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) acc[0] += W[r].x * W[r].x + W[r].y * W[r].y;
      block_sum<1>(acc, red, buf);
      const double nv = sqrt(acc[0]);
#pragma unroll
      for (int r = 0; r < R; ++r) vk[r] = make_double2(W[r].x / nv, W[r].y / nv);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * NTR;
      if (i < n) {
        V[(int64_t)k * n + i] = vk[r];
        s_x[i] = vk[r];
      }
    }
    if (tid == 0) s_beta[k] = beta;
    __syncthreads();

    if (breakdown && !fixed) {
      for (int j = 0; j < k; ++j) {
        const double2* Vj = V + (int64_t)j * n;
        acc[0] = acc[1] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * NTR;
          if (i < n) {
            double2 x = Vj[i];
            acc[0] += x.x * vk[r].x + x.y * vk[r].y;
            acc[1] += x.x * vk[r].y - x.y * vk[r].x;
          }
        }
        block_sum<2>(acc, red, buf);
        if (tid == 0) {
          G[j * kc + k] = make_double2(acc[0], acc[1]);
          G[k * kc + j] = make_double2(acc[0], -acc[1]);
        }
      }
    }

    acc[0] = acc[1] = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * NTR;
      if (i < n) {
        double2 w = spmv_row(i, a.row_ptr, a.col, a.aval, s_x);
        W[r] = w;
        acc[0] += w.x * vk[r].x + w.y * vk[r].y;
        acc[1] += w.x * vk[r].y - w.y * vk[r].x;
      }
    }
    block_sum<2>(acc, red, buf);
    alpha = make_double2(acc[0], acc[1]);
#pragma unroll
    for (int r = 0; r < R; ++r) W[r] = csub(csub(W[r], cmul(alpha, vk[r])), cscale(beta, vp[r]));
    if (tid == 0) s_alpha[k] = alpha;

    if (!fixed) {
      __syncthreads();
      if (tid < 64) {
        double cv = adaptive_metric(k, breakdown, G, kc, s_alpha, s_beta, a.f, s_c, s_cprev);
        if (tid == 0) s_conv = cv;
      }
      __syncthreads();
      conv = s_conv;
    }
    ++k;
  }

  __syncthreads();
  if (fixed && tid < 64 && k > 1) expm_e0(k, s_alpha, s_beta, a.f, s_c);
  __syncthreads();

#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = tid + r * NTR;
    if (i < n) {
      double2 u = zero;
      for (int j = 0; j < k - 1; ++j) u = cfma(s_c[j], V[(int64_t)j * n + i], u);
      u = cfma(s_c[k - 1], vk[r], u);
      if (a.phase) u = cmul(u, a.phase[i]);
      out[i] = u;
    }
  }
  if (tid == 0) {
    if (a.kused) a.kused[b] = k;
    put_status(a, b, !fixed && k >= a.maxorder);
  }
}

// W[c][r] = (H x_c)[row tid + r*NTR] for the ELL operator, branch-free: rows
// past n read row n-1 (a valid address) and are zeroed at the end, so all R
// loads of one ELL slot j issue together.  x_c = s_x + c*n (LDS).
template <int NTR, int R, int C, bool OPL>
__device__ __forceinline__ void ell_apply(double2 (&W)[C][R], int n, int Wd, const double2* s_val,
                                          const uint16_t* s_col, const int32_t* __restrict__ g_col,
                                          const double2* __restrict__ g_val, const double2* s_x) {
  const int tid = threadIdx.x;
  const double2 zero = make_double2(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int r = 0; r < R; ++r) W[c][r] = zero;
  for (int j = 0; j < Wd; ++j) {
    int col[R];
    double2 val[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = min(tid + r * NTR, n - 1);
      const int q = j * n + i;
      col[r] = OPL ? (int)s_col[q] : g_col[q];
      val[r] = OPL ? s_val[q] : g_val[q];
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < C; ++c) W[c][r] = cfma(val[r], s_x[c * n + col[r]], W[c][r]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (tid + r * NTR >= n)
#pragma unroll
      for (int c = 0; c < C; ++c) W[c][r] = zero;
}

// ELL variant: persistent workgroups, C wavepackets per workgroup advanced in
// lockstep (shared barriers, batched reductions, one operator read serving C
// columns), R rows per thread in registers, the C current Krylov vectors in
// LDS, and (OPL) the ELL operator staged into LDS once per workgroup.
// Columns of a group that converge early idle until the group is done.
// FIXED (compile time) selects the fixed Krylov dimension; the adaptive
// instantiation carries the per-iteration stopping metric.
template <int NTR, int R, int C, bool OPL, bool FIXED>
__global__ __launch_bounds__(NTR) void k_lanczos_ell(StepArgs a) {
  constexpr int NW = NTR / 64;
  static_assert(C <= NW, "one wave per column for the adaptive metric");
  extern __shared__ double2 s_dyn[];
  __shared__ double red[2][NW][16];
  __shared__ double2 s_alpha[C][KMAX];
  __shared__ double s_beta[C][KMAX + 1];
  __shared__ double2 s_cf[C][KMAX];
  __shared__ double s_conv[C];
  int buf = 0;
  const int tid = threadIdx.x;
  const int n = (int)a.n;
  const int Wd = a.ell_w;
  const int kc = a.kcap;
  double2* s_x = s_dyn;
  double2* s_val = s_dyn + (size_t)C * n;
  uint16_t* s_col = (uint16_t*)(s_val + (size_t)Wd * n);
  constexpr bool fixed = FIXED;
  const bool skip = a.flag != nullptr && (*a.flag != a.flag_expect);
  const double2 zero = make_double2(0.0, 0.0);

  if (OPL && !skip) {
    for (int q = tid; q < Wd * n; q += NTR) {
      s_val[q] = a.ell_val[q];
      s_col[q] = (uint16_t)a.ell_col[q];
    }
  }

  for (int g = blockIdx.x; g * C < a.ncols; g += gridDim.x) {
    bool act[C];
    double2* V[C];
    double2* G[C];
    double2* cv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int b = g * C + c;
      act[c] = b < a.ncols;
      const int bb = act[c] ? b : 0;
      V[c] = a.basis + (int64_t)bb * kc * n;
      G[c] = a.gram + (int64_t)bb * kc * kc;
      cv[c] = a.cscr + (int64_t)bb * 2 * kc;
    }
    double2 vk[C][R], vp[C][R], W[C][R];
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = tid + r * NTR;
        vk[c][r] = zero;
        vp[c][r] = zero;
        W[c][r] = zero;
        if (act[c] && i < n) {
          const int64_t off = (int64_t)(g * C + c) * a.ld + i;
          double2 v = a.psi_in[off];
          if (a.phase) v = cmul(v, a.phase[i]);
          if (skip) {
            a.psi_out[off] = a.phase ? cmul(v, a.phase[i]) : v;
          } else {
            vk[c][r] = v;
            V[c][i] = v;
            s_x[c * n + i] = v;
          }
        }
      }
    if (skip) {
      if (tid == 0)
        for (int c = 0; c < C; ++c)
          if (act[c]) {
            if (a.kused) a.kused[g * C + c] = 0;
            put_status(a, g * C + c, false);
          }
      continue;
    }
    __syncthreads();

    // w = H V_0 for every column of the group
    double acc[3 * C];
#pragma unroll
    for (int q = 0; q < 3 * C; ++q) acc[q] = 0.0;
    ell_apply<NTR, R, C, OPL>(W, n, Wd, s_val, s_col, a.ell_col, a.ell_val, s_x);
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 w = W[c][r], v = vk[c][r];
        acc[3 * c] += w.x * v.x + w.y * v.y;
        acc[3 * c + 1] += w.x * v.y - w.y * v.x;
        acc[3 * c + 2] += v.x * v.x + v.y * v.y;
      }
    block_sum_dpp<3 * C>(acc, red, buf);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const double2 al = make_double2(acc[3 * c], acc[3 * c + 1]);
#pragma unroll
      for (int r = 0; r < R; ++r) W[c][r] = csub(W[c][r], cmul(al, vk[c][r]));
      if (tid == 0) {
        s_alpha[c][0] = al;
        if (!fixed && act[c]) {
          G[c][0] = make_double2(acc[3 * c + 2], 0.0);
          cv[c][kc] = make_double2(1.0, 0.0);   // c_prev = [1]
        }
      }
    }

    bool live[C];
    double conv[C];
    int kk[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      conv[c] = 1.0;
      kk[c] = 1;
      live[c] = act[c];
    }
    for (int it = 1;; ++it) {
      bool any = false;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (live[c]) {
          if (fixed) live[c] = it < a.fixed_m;
          else live[c] = it < a.maxorder && conv[c] > a.tol;
        }
        any |= live[c];
      }
      if (!any) break;
      // beta_it = |W| (tdse.py:453)
      double bs[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        bs[c] = 0.0;
        if (live[c])
#pragma unroll
          for (int r = 0; r < R; ++r) bs[c] += W[c][r].x * W[c][r].x + W[c][r].y * W[c][r].y;
      }
      block_sum_dpp<C>(bs, red, buf);
      bool brk[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const double beta = sqrt(bs[c]);
        bs[c] = beta;
        brk[c] = live[c] && !(beta != 0.0);
        if (live[c] && !brk[c]) {
          const double ib = 1.0 / beta;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            vp[c][r] = vk[c][r];
            vk[c][r] = cscale(ib, W[c][r]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (!brk[c]) continue;
        // exact breakdown: Gram-Schmidt of ones against the stored basis (tdse.py:459-465)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          vp[c][r] = vk[c][r];
          W[c][r] = (tid + r * NTR < n) ? make_double2(1.0, 0.0) : zero;
        }
        for (int j = 0; j < it; ++j) {
          const double2* Vj = V[c] + (int64_t)j * n;
          double d[2] = {0.0, 0.0};
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = tid + r * NTR;
            if (i < n) {
              const double2 x = Vj[i];
              d[0] += x.x * W[c][r].x + x.y * W[c][r].y;
              d[1] += x.x * W[c][r].y - x.y * W[c][r].x;
            }
          }
          block_sum_dpp<2>(d, red, buf);
          const double2 proj = make_double2(d[0], d[1]);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = tid + r * NTR;
            if (i < n) W[c][r] = csub(W[c][r], cmul(proj, Vj[i]));
          }
        }
        double nv = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) nv += W[c][r].x * W[c][r].x + W[c][r].y * W[c][r].y;
        block_sum_dpp<1>(&nv, red, buf);
        nv = sqrt(nv);
#pragma unroll
        for (int r = 0; r < R; ++r) vk[c][r] = make_double2(W[c][r].x / nv, W[c][r].y / nv);
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (!live[c]) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int i = tid + r * NTR;
          if (i < n) {
            V[c][(int64_t)it * n + i] = vk[c][r];
            s_x[c * n + i] = vk[c][r];
          }
        }
        if (tid == 0) s_beta[c][it] = bs[c];
      }
      __syncthreads();

#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (!(brk[c] && !fixed)) continue;
        for (int j = 0; j < it; ++j) {
          const double2* Vj = V[c] + (int64_t)j * n;
          double d[2] = {0.0, 0.0};
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = tid + r * NTR;
            if (i < n) {
              const double2 x = Vj[i];
              d[0] += x.x * vk[c][r].x + x.y * vk[c][r].y;
              d[1] += x.x * vk[c][r].y - x.y * vk[c][r].x;
            }
          }
          block_sum_dpp<2>(d, red, buf);
          if (tid == 0) {
            G[c][j * kc + it] = make_double2(d[0], d[1]);
            G[c][it * kc + j] = make_double2(d[0], -d[1]);
          }
        }
      }

      // w = H V_it; alpha = vdot(w, V_it); W = w - alpha V_it - beta V_{it-1}
      // (finished columns are multiplied too; their W is never read again)
      double as[2 * C];
#pragma unroll
      for (int c = 0; c < C; ++c) as[2 * c] = as[2 * c + 1] = 0.0;
      ell_apply<NTR, R, C, OPL>(W, n, Wd, s_val, s_col, a.ell_col, a.ell_val, s_x);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (!live[c]) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double2 w = W[c][r], v = vk[c][r];
          as[2 * c] += w.x * v.x + w.y * v.y;
          as[2 * c + 1] += w.x * v.y - w.y * v.x;
        }
      }
      block_sum_dpp<2 * C>(as, red, buf);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (!live[c]) continue;
        const double2 al = make_double2(as[2 * c], as[2 * c + 1]);
#pragma unroll
        for (int r = 0; r < R; ++r)
          W[c][r] = csub(csub(W[c][r], cmul(al, vk[c][r])), cscale(bs[c], vp[c][r]));
        if (tid == 0) s_alpha[c][it] = al;
        kk[c] = it + 1;
      }

      if (!fixed) {
        __syncthreads();
        const int w = tid >> 6;
        if (w < C && live[w]) {
          const double cvw = adaptive_metric(it, brk[w], G[w], kc, s_alpha[w], s_beta[w], a.f,
                                             cv[w], cv[w] + kc);
          if ((tid & 63) == 0) s_conv[w] = cvw;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < C; ++c)
          if (live[c]) conv[c] = s_conv[c];
      }
    }

    // coefficients of the final combination, into LDS for every wave
    __syncthreads();
    {
      const int w = tid >> 6, l = tid & 63;
      if (w < C && act[w]) {
        if (fixed) {
          if (kk[w] > 1) expm_e0(kk[w], s_alpha[w], s_beta[w], a.f, s_cf[w]);
          else if (l == 0) s_cf[w][0] = make_double2(1.0, 0.0);
        } else if (kk[w] == 1) {
          if (l == 0) s_cf[w][0] = make_double2(1.0, 0.0);
        } else {
          for (int j = l; j < kk[w]; j += 64) s_cf[w][j] = cv[w][j];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (!act[c]) continue;
      const int k = kk[c];
      double2 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = zero;
      for (int j = 0; j < k - 1; ++j) {
        const double2 cj = s_cf[c][j];
        const double2* Vj = V[c] + (int64_t)j * n;
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = cfma(cj, Vj[min(tid + r * NTR, n - 1)], u[r]);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = tid + r * NTR;
        if (i < n) {
          double2 uu = cfma(s_cf[c][k - 1], vk[c][r], u[r]);
          if (a.phase) uu = cmul(uu, a.phase[i]);
          a.psi_out[(int64_t)(g * C + c) * a.ld + i] = uu;
        }
      }
      if (tid == 0) {
        if (a.kused) a.kused[g * C + c] = k;
        put_status(a, g * C + c, !fixed && k >= a.maxorder);
      }
    }
    // the next group rewrites s_x only after the barrier that follows its loads
    __syncthreads();
  }
}

__global__ __launch_bounds__(NT) void k_lanczos_step(StepArgs a) {
  __shared__ double red[2][NWAVE][4];
  __shared__ double2 s_alpha[KMAX];
  __shared__ double s_beta[KMAX + 1];
  __shared__ double2 s_c[KMAX];
  __shared__ double2 s_cprev[KMAX];
  __shared__ double s_conv;
  int buf = 0;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t n = a.n;
  const double2* psi = a.psi_in + (int64_t)b * a.ld;
  double2* out = a.psi_out + (int64_t)b * a.ld;
  double2* V = a.basis + (int64_t)b * a.kcap * n;
  double2* W = a.wbuf + (int64_t)b * n;
  double2* G = a.gram + (int64_t)b * a.kcap * a.kcap;
  const int kc = a.kcap;
  const bool fixed = a.fixed_m > 0;
  const bool skip = a.flag != nullptr && (*a.flag != a.flag_expect);

  // phi = P * psi  (tdse.py:375)
  for (int64_t i = tid; i < n; i += NT) {
    double2 v = psi[i];
    if (a.phase) v = cmul(v, a.phase[i]);
    if (skip)
      out[i] = a.phase ? cmul(v, a.phase[i]) : v;   // tdse.py:377 skipped, 394 applied
    else
      V[i] = v;
  }
  if (skip) {
    if (tid == 0) {
      if (a.kused) a.kused[b] = 0;
      put_status(a, b, false);
    }
    return;
  }
  __syncthreads();

  // first Krylov vector (tdse.py:443-446): w = H V0, T00 = vdot(w, V0), W0 = w - T00 V0
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = tid; i < n; i += NT) {
    double2 w = spmv_row(i, a.row_ptr, a.col, a.aval, V);
    double2 v = V[i];
    W[i] = w;
    acc[0] += w.x * v.x + w.y * v.y;
    acc[1] += w.x * v.y - w.y * v.x;
    acc[2] += v.x * v.x + v.y * v.y;
  }
  block_sum<3>(acc, red, buf);
  double2 alpha = make_double2(acc[0], acc[1]);
  for (int64_t i = tid; i < n; i += NT) W[i] = csub(W[i], cmul(alpha, V[i]));
  if (tid == 0) {
    s_alpha[0] = alpha;
    s_cprev[0] = make_double2(1.0, 0.0);
    s_c[0] = make_double2(1.0, 0.0);
    if (!fixed) G[0] = make_double2(acc[2], 0.0);
  }

  int k = 1;
  double conv = 1.0;
  for (;;) {
    if (fixed) {
      if (k >= a.fixed_m) break;
    } else if (!(k < a.maxorder && conv > a.tol)) {
      break;
    }
    double2* Vk = V + (int64_t)k * n;
    const double2* Vkm1 = V + (int64_t)(k - 1) * n;
    // beta_k = |W_{k-1}|  (tdse.py:453)
    acc[0] = 0.0;
    for (int64_t i = tid; i < n; i += NT) acc[0] += W[i].x * W[i].x + W[i].y * W[i].y;
    block_sum<1>(acc, red, buf);
    const double beta = sqrt(acc[0]);
    const bool breakdown = !(beta != 0.0);
    if (!breakdown) {
      for (int64_t i = tid; i < n; i += NT) Vk[i] = make_double2(W[i].x / beta, W[i].y / beta);
    } else {
      // exact breakdown: Gram-Schmidt of a vector of ones (tdse.py:459-465)
      for (int64_t i = tid; i < n; i += NT) W[i] = make_double2(1.0, 0.0);
      for (int j = 0; j < k; ++j) {
        const double2* Vj = V + (int64_t)j * n;
        acc[0] = acc[1] = 0.0;
        for (int64_t i = tid; i < n; i += NT) {
          double2 x = Vj[i], y = W[i];
          acc[0] += x.x * y.x + x.y * y.y;
          acc[1] += x.x * y.y - x.y * y.x;
        }
        block_sum<2>(acc, red, buf);
        double2 proj = make_double2(acc[0], acc[1]);
        for (int64_t i = tid; i < n; i += NT) W[i] = csub(W[i], cmul(proj, Vj[i]));
      }
      acc[0] = 0.0;
      for (int64_t i = tid; i < n; i += NT) acc[0] += W[i].x * W[i].x + W[i].y * W[i].y;
      block_sum<1>(acc, red, buf);
      const double nv = sqrt(acc[0]);
      for (int64_t i = tid; i < n; i += NT) Vk[i] = make_double2(W[i].x / nv, W[i].y / nv);
    }
    if (tid == 0) s_beta[k] = beta;
    __syncthreads();   // V_k complete

    if (breakdown && !fixed) {
      // Gram column of the re-started vector, explicit dots
      for (int j = 0; j < k; ++j) {
        const double2* Vj = V + (int64_t)j * n;
        acc[0] = acc[1] = 0.0;
        for (int64_t i = tid; i < n; i += NT) {
          double2 x = Vj[i], y = Vk[i];
          acc[0] += x.x * y.x + x.y * y.y;
          acc[1] += x.x * y.y - x.y * y.x;
        }
        block_sum<2>(acc, red, buf);
        if (tid == 0) {
          G[j * kc + k] = make_double2(acc[0], acc[1]);
          G[k * kc + j] = make_double2(acc[0], -acc[1]);
        }
      }
    }

    // w = H V_k, T_kk = vdot(w, V_k), W_k = w - T_kk V_k - beta V_{k-1}  (tdse.py:467-470)
    acc[0] = acc[1] = 0.0;
    for (int64_t i = tid; i < n; i += NT) {
      double2 w = spmv_row(i, a.row_ptr, a.col, a.aval, Vk);
      double2 v = Vk[i];
      W[i] = w;
      acc[0] += w.x * v.x + w.y * v.y;
      acc[1] += w.x * v.y - w.y * v.x;
    }
    block_sum<2>(acc, red, buf);
    alpha = make_double2(acc[0], acc[1]);
    for (int64_t i = tid; i < n; i += NT)
      W[i] = csub(csub(W[i], cmul(alpha, Vk[i])), cscale(beta, Vkm1[i]));
    if (tid == 0) s_alpha[k] = alpha;

    if (!fixed) {
      __syncthreads();
      if (tid < 64) {
        double cv = adaptive_metric(k, breakdown, G, kc, s_alpha, s_beta, a.f, s_c, s_cprev);
        if (tid == 0) s_conv = cv;
      }
      __syncthreads();
      conv = s_conv;
    }
    ++k;
  }

  __syncthreads();
  if (fixed && tid < 64 && k > 1) expm_e0(k, s_alpha, s_beta, a.f, s_c);
  __syncthreads();

  // u = sum_i c_i V_i ; out = u * P  (tdse.py:475, 394)
  for (int64_t i = tid; i < n; i += NT) {
    double2 u = make_double2(0.0, 0.0);
    for (int j = 0; j < k; ++j) u = cfma(s_c[j], V[(int64_t)j * n + i], u);
    if (a.phase) u = cmul(u, a.phase[i]);
    out[i] = u;
  }
  if (tid == 0) {
    if (a.kused) a.kused[b] = k;
    put_status(a, b, !fixed && k >= a.maxorder);
  }
}

template <int NTR, int R>
hipError_t launch_resident(const StepArgs& a, unsigned ncols, hipStream_t s) {
  const size_t lds = sizeof(double2) * (size_t)a.n;
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void*)k_lanczos_resident<NTR, R>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 8192);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_lanczos_resident<NTR, R>), dim3(ncols), dim3(NTR), lds, s, a);
  return hipGetLastError();
}

// Largest basis handled by the resident kernel: its Krylov vector fills the
// dynamic LDS (16 B per state) next to ~6 KB of static LDS.
constexpr int64_t RESIDENT_MAX_N = 8192;
constexpr int64_t ELL_MAX_N = 4096;
constexpr size_t LDS_BYTES = 160 * 1024;

// Launch the ELL kernel if its LDS need fits.  *launched tells whether it
// did; the return value is the HIP status of the attempt.
template <int NTR, int R, int C, bool OPL, bool FIXED>
hipError_t try_launch_ell_k(const StepArgs& a, int ncols, hipStream_t s, int n_cu, bool* launched) {
  *launched = false;
  auto kern = k_lanczos_ell<NTR, R, C, OPL, FIXED>;
  static size_t static_lds = (size_t)-1;
  if (static_lds == (size_t)-1) {
    hipFuncAttributes fa;
    hipError_t e = hipFuncGetAttributes(&fa, (const void*)kern);
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(LDS_BYTES - fa.sharedSizeBytes));
    if (e != hipSuccess) return e;
    static_lds = fa.sharedSizeBytes;
  }
  const size_t lds = (size_t)C * a.n * sizeof(double2) +
                     (OPL ? (size_t)a.ell_w * a.n * (sizeof(double2) + sizeof(uint16_t)) : 0);
  if (lds + static_lds > LDS_BYTES) return hipSuccess;
  // The occupancy query does not count LDS beyond 64 KB; with more, the LDS
  // bound (one workgroup per 160 KB CU) is the binding one.
  int occ = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)kern, NTR, lds);
  if (e != hipSuccess) return e;
  const int occ_lds = (int)(LDS_BYTES / (lds + static_lds));
  if (occ < 1) occ = 1;
  if (occ > occ_lds) occ = occ_lds;
  const int groups = (ncols + C - 1) / C;
  const int grid = groups < n_cu * occ ? groups : n_cu * occ;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NTR), lds, s, a);
  *launched = true;
  return hipGetLastError();
}

template <int NTR, int R, int C, bool OPL>
hipError_t try_launch_ell(const StepArgs& a, int ncols, hipStream_t s, int n_cu, bool* launched) {
  if (a.fixed_m > 0) return try_launch_ell_k<NTR, R, C, OPL, true>(a, ncols, s, n_cu, launched);
  return try_launch_ell_k<NTR, R, C, OPL, false>(a, ncols, s, n_cu, launched);
}

template <int NTR, int R>
hipError_t launch_ell(const StepArgs& a, int ncols, hipStream_t s, int n_cu) {
  bool done = false;
  hipError_t e;
  if (NTR <= 512 && R <= 4 && ncols >= 2) {
    e = try_launch_ell<NTR, R, 2, true>(a, ncols, s, n_cu, &done);
    if (done || e != hipSuccess) return e;
  }
  e = try_launch_ell<NTR, R, 1, true>(a, ncols, s, n_cu, &done);
  if (done || e != hipSuccess) return e;
  if (NTR <= 512 && R <= 4 && ncols >= 2) {
    e = try_launch_ell<NTR, R, 2, false>(a, ncols, s, n_cu, &done);
    if (done || e != hipSuccess) return e;
  }
  e = try_launch_ell<NTR, R, 1, false>(a, ncols, s, n_cu, &done);
  if (!done && e == hipSuccess) e = hipErrorInvalidConfiguration;
  return e;
}

// Explicit ELL configurations (threads, rows per thread, columns per
// workgroup, operator in LDS), selected by rm_plan_set_kernel(RM_KERNEL_ELL_CFG0 + i).
struct EllCfg {
  int ntr, r, c, opl;
};
constexpr EllCfg ELL_CFGS[] = {
    {256, 4, 1, 1}, {256, 4, 2, 1}, {256, 8, 1, 1}, {512, 4, 1, 1}, {512, 4, 2, 1},
    {512, 8, 1, 1}, {256, 4, 1, 0}, {256, 4, 2, 0}, {256, 8, 1, 0}, {512, 4, 1, 0},
    {512, 4, 2, 0}, {512, 8, 1, 0}};
constexpr int N_ELL_CFGS = (int)(sizeof(ELL_CFGS) / sizeof(ELL_CFGS[0]));

hipError_t launch_ell_cfg(int i, const StepArgs& a, int ncols, hipStream_t s, int n_cu, bool* done) {
  switch (i) {
    case 0: return try_launch_ell<256, 4, 1, true>(a, ncols, s, n_cu, done);
    case 1: return try_launch_ell<256, 4, 2, true>(a, ncols, s, n_cu, done);
    case 2: return try_launch_ell<256, 8, 1, true>(a, ncols, s, n_cu, done);
    case 3: return try_launch_ell<512, 4, 1, true>(a, ncols, s, n_cu, done);
    case 4: return try_launch_ell<512, 4, 2, true>(a, ncols, s, n_cu, done);
    case 5: return try_launch_ell<512, 8, 1, true>(a, ncols, s, n_cu, done);
    case 6: return try_launch_ell<256, 4, 1, false>(a, ncols, s, n_cu, done);
    case 7: return try_launch_ell<256, 4, 2, false>(a, ncols, s, n_cu, done);
    case 8: return try_launch_ell<256, 8, 1, false>(a, ncols, s, n_cu, done);
    case 9: return try_launch_ell<512, 4, 1, false>(a, ncols, s, n_cu, done);
    case 10: return try_launch_ell<512, 4, 2, false>(a, ncols, s, n_cu, done);
    default: return try_launch_ell<512, 8, 1, false>(a, ncols, s, n_cu, done);
  }
}

bool ell_cfg_valid(int i, int64_t n, int64_t ncols) {
  if (i < 0 || i >= N_ELL_CFGS) return false;
  const EllCfg& c = ELL_CFGS[i];
  return n <= (int64_t)c.ntr * c.r && (c.c == 1 || ncols >= 2);
}

hipError_t launch_step(const StepArgs& a, unsigned ncols, hipStream_t s, int variant, int n_cu) {
  const int64_t n = a.n;
  if (variant >= RM_KERNEL_ELL_CFG0) {
    const int i = variant - RM_KERNEL_ELL_CFG0;
    if (!ell_cfg_valid(i, n, ncols)) return hipErrorInvalidConfiguration;
    bool done = false;
    hipError_t e = launch_ell_cfg(i, a, (int)ncols, s, n_cu, &done);
    if (!done && e == hipSuccess) e = hipErrorInvalidConfiguration;
    return e;
  }
  if ((variant == RM_KERNEL_AUTO || variant == RM_KERNEL_ELL) && n <= ELL_MAX_N) {
    if (n <= 1024) return launch_ell<256, 4>(a, (int)ncols, s, n_cu);
    if (n <= 2048) return launch_ell<512, 4>(a, (int)ncols, s, n_cu);
    return launch_ell<512, 8>(a, (int)ncols, s, n_cu);
  }
  if (variant == RM_KERNEL_RESIDENT && n <= RESIDENT_MAX_N) {
    if (n <= 256) return launch_resident<256, 1>(a, ncols, s);
    if (n <= 512) return launch_resident<256, 2>(a, ncols, s);
    if (n <= 1024) return launch_resident<256, 4>(a, ncols, s);
    if (n <= 2048) return launch_resident<256, 8>(a, ncols, s);
    if (n <= 4096) return launch_resident<512, 8>(a, ncols, s);
    return launch_resident<1024, 8>(a, ncols, s);
  }
  hipLaunchKernelGGL(k_lanczos_step, dim3(ncols), dim3(NT), 0, s, a);
  return hipGetLastError();
}

}  // namespace

extern "C" int rm_ell_config(int32_t i, int32_t* threads, int32_t* rows, int32_t* cols,
                             int32_t* op_lds) {
  if (i < 0 || i >= N_ELL_CFGS) return fail(RM_ERR_ARG, "rm_ell_config: no configuration %d", i);
  if (threads) *threads = ELL_CFGS[i].ntr;
  if (rows) *rows = ELL_CFGS[i].r;
  if (cols) *cols = ELL_CFGS[i].c;
  if (op_lds) *op_lds = ELL_CFGS[i].opl;
  return RM_OK;
}

extern "C" int rm_plan_set_kernel(rm_plan* p, int32_t variant) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_kernel: NULL plan");
  const bool cfg = variant >= RM_KERNEL_ELL_CFG0 && variant < RM_KERNEL_ELL_CFG0 + N_ELL_CFGS;
  if (!cfg && (variant < RM_KERNEL_AUTO || variant > RM_KERNEL_BLOCK))
    return fail(RM_ERR_ARG, "rm_plan_set_kernel: unknown variant %d", variant);
  if (cfg && p->n > (int64_t)ELL_CFGS[variant - RM_KERNEL_ELL_CFG0].ntr * ELL_CFGS[variant - RM_KERNEL_ELL_CFG0].r)
    return fail(RM_ERR_ARG, "rm_plan_set_kernel: n=%lld exceeds ELL configuration %d",
                (long long)p->n, variant - RM_KERNEL_ELL_CFG0);
  if (variant == RM_KERNEL_RESIDENT && p->n > RESIDENT_MAX_N)
    return fail(RM_ERR_ARG, "rm_plan_set_kernel: n=%lld exceeds the resident kernel's %lld states",
                (long long)p->n, (long long)RESIDENT_MAX_N);
  if (variant == RM_KERNEL_ELL && p->n > ELL_MAX_N)
    return fail(RM_ERR_ARG, "rm_plan_set_kernel: n=%lld exceeds the ELL kernel's %lld states",
                (long long)p->n, (long long)ELL_MAX_N);
  p->variant = variant;
  return RM_OK;
}

int ensure_ws(rm_plan* p, size_t bytes) {
  if (p->ws_bytes >= bytes) return RM_OK;
  if (p->ws) {
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(p->ws));
    p->ws = nullptr;
    p->ws_bytes = 0;
  }
  HIPCHK(hipMalloc(&p->ws, bytes));
  p->ws_bytes = bytes;
  return RM_OK;
}

// Event pair around one timed launch (no-ops unless the plan's timing is on).
hipError_t timing_begin(rm_plan* p, hipStream_t s) {
  if (!p->timing) return hipSuccess;
  if (p->nrec >= (int)p->ev0.size()) {
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
      (void)hipEventDestroy(e0);
      return e;
    }
    p->ev0.push_back(e0);
    p->ev1.push_back(e1);
  }
  return hipEventRecord(p->ev0[p->nrec], s);
}

hipError_t timing_end(rm_plan* p, hipStream_t s) {
  if (!p->timing) return hipSuccess;
  hipError_t e = hipEventRecord(p->ev1[p->nrec], s);
  p->nrec++;
  return e;
}

// launch_step, bracketed by HIP events on the same stream when timing is on.
static hipError_t timed_launch_step(rm_plan* p, const StepArgs& a, unsigned ncols, hipStream_t s) {
  hipError_t e = timing_begin(p, s);
  if (e != hipSuccess) return e;
  const bool prop_only = p->variant == RM_KERNEL_SUBSPACE || p->variant == RM_KERNEL_COOP ||
                         p->variant == RM_KERNEL_BLOCK;
  e = launch_step(a, ncols, s, prop_only ? RM_KERNEL_AUTO : p->variant, p->n_cu);
  if (e != hipSuccess) return e;
  return timing_end(p, s);
}

extern "C" int rm_plan_timing(rm_plan* p, int32_t enable) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_timing: NULL plan");
  p->timing = enable ? 1 : 0;
  p->nrec = 0;
  return RM_OK;
}

extern "C" int rm_plan_kernel_time(rm_plan* p, double* total_ms, int64_t* launches) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_kernel_time: NULL plan");
  double tot = 0.0;
  if (p->nrec > 0) HIPCHK(hipEventSynchronize(p->ev1[p->nrec - 1]));
  for (int i = 0; i < p->nrec; ++i) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, p->ev0[i], p->ev1[i]));
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = p->nrec;
  p->nrec = 0;
  return RM_OK;
}

// Validate the step parameters and fill StepArgs (workspace sized for ncols).
static int step_setup(rm_plan* p, const char* who, int64_t ld, int64_t ncols, double f_re,
                      double f_im, const void* phase, int32_t krylov_fixed, int32_t maxorder,
                      double tol, StepArgs* a, bool workspace = true) {
  if (!p) return fail(RM_ERR_ARG, "%s: NULL plan", who);
  if (ncols < 0 || ld < p->n)
    return fail(RM_ERR_ARG, "%s: bad ld=%lld or ncols=%lld", who, (long long)ld, (long long)ncols);
  if (ncols > INT32_MAX) return fail(RM_ERR_ARG, "%s: ncols too large", who);
  if (krylov_fixed < 0 || krylov_fixed > RM_MAX_KRYLOV)
    return fail(RM_ERR_ARG, "%s: krylov_fixed=%d outside [0, %d]", who, krylov_fixed, RM_MAX_KRYLOV);
  if (krylov_fixed == 0 && (maxorder < 1 || maxorder > RM_MAX_KRYLOV))
    return fail(RM_ERR_ARG, "%s: maxorder=%d outside [1, %d]", who, maxorder, RM_MAX_KRYLOV);
  if (krylov_fixed == 0 && !(tol > 0.0 && tol <= 1.0)) return fail(RM_ERR_ARG, "%s: tol must be in (0, 1]", who);
  const int kcap = krylov_fixed > 0 ? krylov_fixed : maxorder;
  const size_t per_col = (size_t)kcap * p->n + (size_t)p->n + (size_t)kcap * kcap + 2 * (size_t)kcap;
  if (workspace && ncols > 0 && p->n > 0) {
    int rc = ensure_ws(p, sizeof(double2) * per_col * (size_t)ncols);
    if (rc) return rc;
  }
  double2* ws = (double2*)p->ws;
  memset(a, 0, sizeof(*a));
  a->n = p->n;
  a->ld = ld;
  a->phase = (const double2*)phase;
  a->row_ptr = p->row_ptr;
  a->col = p->col;
  a->aval = p->aval;
  a->flag = p->flag;
  a->flag_expect = p->step_id;
  a->f = make_double2(f_re, f_im);
  a->fixed_m = krylov_fixed;
  a->maxorder = maxorder;
  a->tol = tol;
  a->kcap = kcap;
  a->basis = ws;
  a->wbuf = ws + (size_t)ncols * kcap * p->n;
  a->gram = a->wbuf + (size_t)ncols * p->n;
  a->cscr = a->gram + (size_t)ncols * kcap * kcap;
  a->ell_col = p->ell_col;
  a->ell_val = p->ell_val;
  a->ell_w = p->ell_w;
  a->ncols = (int32_t)ncols;
  return RM_OK;
}

extern "C" int rm_step(rm_plan* p, const void* psi_in, void* psi_out, int64_t ld, int64_t ncols,
                       double f_re, double f_im, const void* phase, int32_t krylov_fixed,
                       int32_t maxorder, double tol, int32_t* kused, int32_t* status, void* stream) {
  if (p && p->precision != RM_PREC_FP64)
    return fail(RM_ERR_ARG, "rm_step: the full-vector kernels compute in FP64 only (plan set to FP32)");
  StepArgs a;
  int rc = step_setup(p, "rm_step", ld, ncols, f_re, f_im, phase, krylov_fixed, maxorder, tol, &a);
  if (rc) return rc;
  if (ncols == 0 || p->n == 0) return RM_OK;
  if (!psi_in || !psi_out) return fail(RM_ERR_ARG, "rm_step: NULL psi");
  a.psi_in = (const double2*)psi_in;
  a.psi_out = (double2*)psi_out;
  a.kused = kused;
  a.status = status;
  HIPCHK(timed_launch_step(p, a, (unsigned)ncols, (hipStream_t)stream));
  return RM_OK;
}

__global__ void k_fill_i32(int32_t* x, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_copy_cols(int64_t n, int64_t ld, const double2* __restrict__ src,
                            double2* __restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t b = blockIdx.y;
  if (i < n) dst[b * ld + i] = src[b * ld + i];
}

// invariant-subspace path (rm_subspace.hip)
int sub_propagate(rm_plan* p, void* psi, int64_t ld, int64_t ncols, int64_t nsteps,
                  const double* fprod_dev, double f_re, double f_im, const void* phase,
                  int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                  int32_t* status, int force, hipStream_t s, bool* done, const rm_plan* obs,
                  double2* expect);

__global__ __launch_bounds__(NT) void k_expect(int64_t n, int64_t ld, const int64_t* __restrict__ rp,
                                               const int32_t* __restrict__ col,
                                               const double2* __restrict__ a,
                                               const double2* __restrict__ psi,
                                               double2* __restrict__ out);

// sum[s] = sum over b of ex[s * ncols + b], one wave per step (fixed order)
__global__ __launch_bounds__(64) void k_expect_sum(int64_t nsteps, int64_t ncols, const double2* __restrict__ ex,
                                                   double2* __restrict__ sum) {
  for (int64_t st = blockIdx.x; st < nsteps; st += gridDim.x) {
    double re = 0.0, im = 0.0;
    for (int64_t b = threadIdx.x; b < ncols; b += 64) {
      const double2 v = ex[st * ncols + b];
      re += v.x;
      im += v.y;
    }
    re = wave_allsum(re);
    im = wave_allsum(im);
    if (threadIdx.x == 0) sum[st] = make_double2(re, im);
  }
}

int launch_expect_sum(int64_t nsteps, int64_t ncols, const double2* ex, double2* sum, hipStream_t s) {
  hipLaunchKernelGGL(k_expect_sum, dim3((unsigned)std::min<int64_t>(nsteps, 65535)), dim3(64), 0, s, nsteps, ncols,
                     ex, sum);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

static int propagate_impl(rm_plan* p, void* psi, void* work, int64_t ld, int64_t ncols, int64_t nsteps,
                          const double* fprod_dev, const double* thresh, double f_re, double f_im,
                          const void* phase, int32_t krylov_fixed, int32_t maxorder, double tol,
                          int32_t* kused, int32_t* status, rm_plan* obs, double2* expect,
                          double2* expect_sum, hipStream_t s);

extern "C" int rm_propagate(rm_plan* p, void* psi, void* work, int64_t ld, int64_t ncols,
                            int64_t nsteps, const double* fprod_dev, const double* thresh,
                            double f_re, double f_im, const void* phase, int32_t krylov_fixed,
                            int32_t maxorder, double tol, int32_t* kused, int32_t* status,
                            void* stream) {
  return propagate_impl(p, psi, work, ld, ncols, nsteps, fprod_dev, thresh, f_re, f_im, phase, krylov_fixed,
                        maxorder, tol, kused, status, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int rm_propagate_expect(rm_plan* p, void* psi, void* work, int64_t ld, int64_t ncols,
                                   int64_t nsteps, const double* fprod_dev, const double* thresh,
                                   double f_re, double f_im, const void* phase, int32_t krylov_fixed,
                                   int32_t maxorder, double tol, int32_t* kused, int32_t* status,
                                   rm_plan* obs, void* expect, void* expect_sum, void* stream) {
  if (!p || !obs) return fail(RM_ERR_ARG, "rm_propagate_expect: NULL plan or observable");
  if (obs == p) return fail(RM_ERR_ARG, "rm_propagate_expect: the observable must be a plan of its own");
  if (obs->n != p->n)
    return fail(RM_ERR_ARG, "rm_propagate_expect: observable of %lld states for a plan of %lld",
                (long long)obs->n, (long long)p->n);
  if (obs->field_gen == 0)
    return fail(RM_ERR_ARG, "rm_propagate_expect: the observable plan has no matrix yet (rm_plan_set_field)");
  if (!expect && !expect_sum) return fail(RM_ERR_ARG, "rm_propagate_expect: NULL expect and expect_sum");
  return propagate_impl(p, psi, work, ld, ncols, nsteps, fprod_dev, thresh, f_re, f_im, phase, krylov_fixed,
                        maxorder, tol, kused, status, obs, (double2*)expect, (double2*)expect_sum,
                        (hipStream_t)stream);
}

static int propagate_impl(rm_plan* p, void* psi, void* work, int64_t ld, int64_t ncols, int64_t nsteps,
                          const double* fprod_dev, const double* thresh, double f_re, double f_im,
                          const void* phase, int32_t krylov_fixed, int32_t maxorder, double tol,
                          int32_t* kused, int32_t* status, rm_plan* obs, double2* expect,
                          double2* expect_sum, hipStream_t s) {
  StepArgs a;
  // the full-vector workspace is allocated below, only when those kernels run
  int rc = step_setup(p, "rm_propagate", ld, ncols, f_re, f_im, phase, krylov_fixed, maxorder, tol, &a, false);
  if (rc) return rc;
  if (nsteps < 0) return fail(RM_ERR_ARG, "rm_propagate: nsteps < 0");
  if (ncols > 65535) return fail(RM_ERR_ARG, "rm_propagate: ncols > 65535");
  if (ncols == 0 || p->n == 0 || nsteps == 0) return RM_OK;
  // per-column observable values: the caller's array, or the plan's own when
  // the caller wants only the sums
  double2* ex = expect;
  if (obs && !ex) {
    if ((rc = grow(&p->ex_buf, &p->ex_cap, nsteps * ncols))) return rc;
    ex = p->ex_buf;
  }
  auto finish = [&]() -> int {
    return obs && expect_sum ? launch_expect_sum(nsteps, ncols, ex, expect_sum, s) : RM_OK;
  };
  if (!psi || !work) return fail(RM_ERR_ARG, "rm_propagate: NULL psi or work");
  if (psi == work) return fail(RM_ERR_ARG, "rm_propagate: psi and work must differ");
  if (p->nfield > 0 && !fprod_dev) return fail(RM_ERR_ARG, "rm_propagate: NULL field table");
  FieldArgs fa;
  memset(&fa, 0, sizeof(fa));
  bool any_thresh = false;
  for (int i = 0; i < RM_MAX_LEAF; ++i) fa.thresh[i] = -1.0;
  if (thresh)
    for (int i = 0; i < p->nleaf; ++i) {
      fa.thresh[i] = thresh[i];
      any_thresh |= thresh[i] > 0.0;
    }
  if (p->propag == RM_PROPAG_EXPOKIT) {
    // Expokit's ZHEXPV on the invariant-subspace domains (rm_expokit.hip)
    if (krylov_fixed < 1 || krylov_fixed > RM_EXPOKIT_MAX_M || krylov_fixed >= p->n)
      return fail(RM_ERR_ARG, "rm_propagate: Expokit Krylov size m=%d must be in [1, %d] and below n=%lld",
                  krylov_fixed, RM_EXPOKIT_MAX_M, (long long)p->n);
    if (any_thresh) return fail(RM_ERR_ARG, "rm_propagate: the Expokit propagator takes no MF thresholds");
    if (p->precision != RM_PREC_FP64)
      return fail(RM_ERR_ARG, "rm_propagate: the Expokit propagator computes in FP64 only (plan set to FP32)");
    if (nsteps > INT32_MAX)
      return fail(RM_ERR_ARG, "rm_propagate: more than 2^31 steps");
    bool done = false;
    rc = sub_propagate(p, psi, ld, ncols, nsteps, fprod_dev, f_re, f_im, phase, krylov_fixed, maxorder, tol,
                       kused, status, 0, s, &done, obs, ex);
    if (rc) return rc;
    if (!done) return fail(RM_ERR_ARG, "rm_propagate: the Expokit propagator does not apply");
    return finish();
  }
  const bool sub_kernel = p->variant == RM_KERNEL_SUBSPACE || p->variant == RM_KERNEL_COOP ||
                          p->variant == RM_KERNEL_BLOCK;
  if ((p->variant == RM_KERNEL_AUTO || sub_kernel) && !any_thresh) {
    bool done = false;
    const int force = p->variant == RM_KERNEL_COOP || p->variant == RM_KERNEL_BLOCK ? p->variant : 0;
    rc = sub_propagate(p, psi, ld, ncols, nsteps, fprod_dev, f_re, f_im, phase, krylov_fixed,
                       maxorder, tol, kused, status, force, s, &done, obs, ex);
    if (rc) return rc;
    if (done) return finish();
  }
  if (sub_kernel)
    return fail(RM_ERR_ARG, "rm_propagate: the subspace kernels do not apply (%s)",
                any_thresh ? "MF thresholds given" : "more than 2^31 steps");
  if (p->precision != RM_PREC_FP64)
    return fail(RM_ERR_ARG, "rm_propagate: %s -- the full-vector kernels compute in FP64 only (plan set to FP32)",
                any_thresh ? "MF thresholds send this call to the full-vector kernels" : "the subspace path does not apply");
  if ((rc = step_setup(p, "rm_propagate", ld, ncols, f_re, f_im, phase, krylov_fixed, maxorder, tol, &a))) return rc;
  if (status) hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(ncols, 256)), dim3(256), 0, s, status, ncols, 0);
  a.status = status;
  a.status_sticky = 1;
  double2* cur = (double2*)psi;
  double2* nxt = (double2*)work;
  for (int64_t st = 0; st < nsteps; ++st) {
    p->step_id = (p->step_id + 1) & 0x3fffffff;
    launch_field(p, fa, (const double2*)fprod_dev + st * p->nfield, s);
    a.flag_expect = p->step_id;
    a.psi_in = cur;
    a.psi_out = nxt;
    a.kused = kused ? kused + st * ncols : nullptr;
    HIPCHK(timed_launch_step(p, a, (unsigned)ncols, s));
    if (obs)
      hipLaunchKernelGGL(k_expect, dim3((unsigned)ncols), dim3(NT), 0, s, obs->n, ld, obs->row_ptr, obs->col,
                         obs->aval, (const double2*)nxt, ex + st * ncols);
    double2* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != (double2*)psi)
    hipLaunchKernelGGL(k_copy_cols, dim3(grid_for(p->n, 256), (unsigned)ncols), dim3(256), 0, s,
                       p->n, ld, (const double2*)cur, (double2*)psi);
  HIPCHK(hipGetLastError());
  return finish();
}

// ---------------------------------------------------------------------------
// observables

__global__ __launch_bounds__(NT) void k_expect(int64_t n, int64_t ld, const int64_t* __restrict__ rp,
                                               const int32_t* __restrict__ col,
                                               const double2* __restrict__ a,
                                               const double2* __restrict__ psi,
                                               double2* __restrict__ out) {
  __shared__ double red[2][NWAVE][4];
  int buf = 0;
  const double2* x = psi + (int64_t)blockIdx.x * ld;
  double acc[2] = {0.0, 0.0};
  for (int64_t i = threadIdx.x; i < n; i += NT) {
    double2 y = spmv_row(i, rp, col, a, x);
    double2 v = x[i];
    acc[0] += v.x * y.x + v.y * y.y;
    acc[1] += v.x * y.y - v.y * y.x;
  }
  block_sum<2>(acc, red, buf);
  if (threadIdx.x == 0) out[blockIdx.x] = make_double2(acc[0], acc[1]);
}

extern "C" int rm_expect(rm_plan* p, const void* psi, int64_t ld, int64_t ncols, void* out,
                         void* stream) {
  if (!p) return fail(RM_ERR_ARG, "rm_expect: NULL plan");
  if (ncols < 0 || ld < p->n) return fail(RM_ERR_ARG, "rm_expect: bad ld/ncols");
  if (ncols == 0) return RM_OK;
  if (!psi || !out) return fail(RM_ERR_ARG, "rm_expect: NULL pointer");
  hipLaunchKernelGGL(k_expect, dim3((unsigned)ncols), dim3(NT), 0, (hipStream_t)stream, p->n, ld,
                     p->row_ptr, p->col, p->aval, (const double2*)psi, (double2*)out);
  HIPCHK(hipGetLastError());
  return RM_OK;
}
