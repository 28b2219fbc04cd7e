// richmol_amd: declarations shared by the library's translation units -- error
// reporting, complex arithmetic, the plan object, wave-level reductions and the
// small dense Krylov pieces (exp(fT) e0, adaptive stopping metric).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/richmol_amd.h"

// ---------------------------------------------------------------------------
// errors (message per calling thread, returned by rm_last_error)

int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(RM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

static inline unsigned grid_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// complex helpers (double2 = re, im)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) {
  return make_double2(s * a.x, s * a.y);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// *p when c holds, else 0, as a value.  (`c ? p[k] : zero` with a named zero
// is an lvalue: the compiler selects between the two addresses, keeps the zero
// in scratch memory and reloads it wherever it is used.)
__device__ __forceinline__ double2 load_or_zero(bool c, const double2* p) {
  double2 v = make_double2(0.0, 0.0);
  if (c) v = *p;
  return v;
}
// acc + a*b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
  return make_double2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
}
__device__ __forceinline__ float2 cfmaf(float2 a, float2 b, float2 acc) {
  return make_float2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
}

// matvec operand type of the precision modes: FP64, or FP32 (complex64
// operator and gathered vector, FP32 accumulation)
template <bool F32>
struct XType { using T = double2; };
template <>
struct XType<true> { using T = float2; };

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <bool F32>
__device__ __forceinline__ typename XType<F32>::T to_x(double2 v);
template <>
__device__ __forceinline__ double2 to_x<false>(double2 v) { return v; }
template <>
__device__ __forceinline__ float2 to_x<true>(double2 v) {
  return make_float2(__double2float_rn(v.x), __double2float_rn(v.y));
}
__device__ __forceinline__ double2 xfma(double2 a, double2 b, double2 acc) { return cfma(a, b, acc); }
__device__ __forceinline__ float2 xfma(float2 a, float2 b, float2 acc) { return cfmaf(a, b, acc); }
__device__ __forceinline__ double2 xmul(double2 a, double2 b) { return cmul(a, b); }
__device__ __forceinline__ float2 xmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}


// ---------------------------------------------------------------------------
// plan

// launch choices of rm_propagate's invariant-subspace path
// (rm_plan_set_tuning); the defaults are the hand-set ones
struct RmTuning {
  int32_t pair = 1;        // two wavepackets per wave on domains <= 32 rows (fixed Krylov dimension <= 16)
  int32_t coop_max = 8;    // long-domain wavepackets per call up to which the cooperative form runs
  int32_t coop_grid = 0;   // workgroups of a cooperative launch at most (0: the domain's row groups, one per CU)
  int32_t fork = 1;        // launch classes of a call side by side on auxiliary streams
  int32_t wave_max_rows = 64;       // domains longer than this leave the one-wave kernels
  int32_t coop_min_entries = 6400;  // stored entries per cooperative wavepacket at least
};

struct rm_plan {
  int64_t n = 0, nnz = 0, npair = 0, nterm = 0, nq = 0;
  int32_t nfield = 0, nleaf = 0;
  int device = 0;
  // host copies of the structure (the invariant-subspace path is built from them)
  std::vector<int64_t> h_row_ptr;
  std::vector<int32_t> h_col;
  std::vector<int64_t> h_pair_ptr;
  std::vector<int32_t> h_pair_term;
  std::vector<double2> h_pair_kval;
  std::vector<int64_t> h_term_ptr;
  std::vector<int32_t> h_term_field;
  std::vector<double2> h_term_mval;
  int64_t* row_ptr = nullptr;
  int32_t* col = nullptr;
  int64_t* pair_ptr = nullptr;
  int32_t* pair_term = nullptr;
  double2* pair_kval = nullptr;
  int64_t* term_ptr = nullptr;
  int32_t* term_field = nullptr;
  double2* term_mval = nullptr;
  int32_t* term_leaf = nullptr;
  double2* mf = nullptr;     // nterm
  double2* aval = nullptr;   // nnz
  int32_t* flag = nullptr;   // id of the last field step with a non-zero MF
  int32_t step_id = 0;
  void* ws = nullptr;        // Krylov workspace
  size_t ws_bytes = 0;
  int32_t variant = RM_KERNEL_AUTO;
  int32_t precision = RM_PREC_FP64;
  RmTuning tune;
  // ELL copy of the matrix: width = max entries per row, column-major
  // (entry j of row i at j*n + i), padding entries (value 0, column i)
  int32_t ell_w = 0;
  int32_t* ell_col = nullptr;
  double2* ell_val = nullptr;
  int32_t* ell_pos = nullptr;  // CSR entry -> ELL slot
  int n_cu = 0;
  // invariant subspaces (connected components of the sparsity graph), built
  // on first use: rows permuted so that each subspace is contiguous, and the
  // operator split per field slot, A(F) = sum_s F_s B_s (no thresholds)
  int sub_ready = 0;             // 0 not built, 1 built
  int32_t nsub = 0, max_sub = 0;
  int32_t sub_w = 0;             // ELL width in permuted order
  int32_t* sub_perm = nullptr;   // permuted row -> original row
  int32_t* sub_label = nullptr;  // original row -> subspace
  int32_t* sub_iperm = nullptr;  // original row -> permuted row
  int32_t* sub_lo = nullptr;     // subspace -> first permuted row
  int32_t* sub_hi = nullptr;     // subspace -> one past its last permuted row
  int32_t* sub_col = nullptr;    // [sub_w][n] permuted column of each entry
  double2* sub_val = nullptr;    // [nfield][sub_w][n] B_s values
  std::vector<int32_t> h_sub_start, h_sub_end;   // per subspace: first / one past last permuted row
  std::vector<int32_t> h_sub_label_perm;         // permuted row -> subspace
  std::vector<int32_t> h_sub_perm;               // permuted row -> original row
  std::vector<int32_t> h_sub_col;                // host copy of sub_col ([w][n], permuted columns)
  // the B_s as CSR in permuted order, for domains longer than a wave's
  // (built when some subspace has more than RM_SUBSPACE_MAX_ROWS rows)
  int64_t* sub_rp = nullptr;     // [n+1]
  int32_t* sub_ccol = nullptr;   // [nnz] permuted column
  double2* sub_cval = nullptr;   // [nfield][nnz]
  float2* sub_cval32 = nullptr;  // [nfield][nnz] the same in FP32 (FP32-mode matvecs read it)
  std::vector<int64_t> h_sub_rp;
  // per-call scratch of the subspace path
  int32_t* dom = nullptr;        // 2 per column: domain lo, hi
  int64_t dom_cap = 0;
  int32_t* sflag = nullptr;      // per step: 1 if any MF element is non-zero
  int64_t sflag_cap = 0;
  int32_t* sflag_part = nullptr; // per step and chunk of terms: partial flags
  int64_t sflag_part_cap = 0;
  uint32_t* sflag_arrive = nullptr;   // per step: chunk arrivals (only grows)
  int64_t sflag_arrive_cap = 0;
  std::vector<int32_t> h_dom;
  int32_t* h_dom_pin = nullptr;  // pinned landing buffer of the domains' copy to the host
  int64_t h_dom_pin_cap = 0;
  void* sub_scratch = nullptr;   // launch staging and workspace (rm_subspace.hip)
  // domain reuse (rm_plan_set_domain_reuse): the domains of the last
  // rm_propagate call, kept for its next call on the same wavepackets; each
  // such call checks them on the device and stores its call id into `stale`
  // when they no longer hold
  int dom_reuse = 0;
  int32_t* dom_keep = nullptr;   // device: the kept domains, 2 per column
  int64_t dom_keep_cap = 0;
  int32_t* stale = nullptr;      // device int32[1]
  int32_t call_id = 0;
  // status of calls whose caller passed none (cooperative timeouts are reported)
  int32_t* int_status = nullptr;
  int64_t int_status_cap = 0;
  int32_t* h_status_pin = nullptr;
  int64_t h_status_pin_cap = 0;
  int32_t coop_abort = 0;        // test hook: cooperative launches start with the abort flag raised
  // the matrix of the last rm_plan_set_field: an id unique over every plan
  // and call (observable caches compare it)
  int64_t field_gen = 0;
  // observable of rm_propagate_expect in this plan's permuted ELL form,
  // built from plan obs_src's matrix obs_gen
  const rm_plan* obs_src = nullptr;
  int64_t obs_gen = -1;
  int32_t obs_w = 0;
  int32_t* obs_col = nullptr;
  double2* obs_val = nullptr;
  int64_t obs_cap = 0, obs_vcap = 0;
  double2* ex_buf = nullptr;     // per-column values when the caller wants only the sums
  int64_t ex_cap = 0;
  // propagator of rm_propagate (rm_plan_set_propagator) and the Expokit
  // path's scratch: workgroup regions, per-step norms, work items
  int32_t propag = RM_PROPAG_LANCZOS;
  void* expo_ws = nullptr;
  size_t expo_ws_bytes = 0;
  double* expo_anorm = nullptr;
  int64_t expo_anorm_cap = 0;
  int4* expo_items = nullptr;
  int64_t expo_items_cap = 0;
  // optional HIP-event timing of the Krylov step launches
  int timing = 0;
  int nrec = 0;
  std::vector<hipEvent_t> ev0, ev1;
};

// One propagation job of a launch: an operator (a plan's subspace arrays),
// the split-operator phases and field table of the job, and its wavepackets.
struct SubOp {
  const int32_t* col;      // [w][n] permuted column of each ELL entry
  const double2* val;      // [nslot][w][n] B_s
  const int64_t* rp;       // the same B_s as permuted CSR: row pointers [n+1],
  const int32_t* ccol;     //   columns [nnz] and
  const double2* cval;     //   values [nslot][nnz]
  const float2* cval32;    //   and their FP32 rounding
  int64_t nnz;
  const int32_t* perm;     // permuted row -> original row
  const double2* phase;    // original order, or NULL
  const double2* table;    // [nsteps][nslot] field products of the launch's steps
  const int32_t* sflag;    // [nsteps] 1: some MF element non-zero
  double2* psi;            // column b at psi + b*ld (original order)
  int64_t ld;
  int32_t* kused;          // kused[step*kused_ld + b - kcol0] or NULL
  int64_t kused_ld;
  int32_t* status;         // status[b - kcol0] or NULL: failures only
  const int32_t* stale;    // device flag: == call_id when the reused domains failed their check
  // observable read after every step (ow == 0: none): the operator in the
  // permuted ELL form [ow][n] (permuted columns; padding: own row, value 0);
  // expect[step * expect_ld + b - kcol0] = <psi_b|O|psi_b> after that step
  const int32_t* ocol;
  const double2* oval;
  double2* expect;
  int64_t expect_ld;
  int32_t ow;
  double2 f;
  int32_t n, w, nslot, nsteps, kcol0, call_id;
};

// Whether step `step` runs the expv (the `len(H.mfmat) > 0` test of tdse.py:377).
__device__ __forceinline__ bool step_on(const SubOp& op, int step) { return op.sflag[step] != 0; }

template <typename T>
static int upload(T** dst, const T* src, int64_t count) {
  *dst = nullptr;
  size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
  HIPCHK(hipMalloc((void**)dst, bytes));
  if (count > 0 && src) HIPCHK(hipMemcpy(*dst, src, sizeof(T) * (size_t)count, hipMemcpyHostToDevice));
  return RM_OK;
}

// Grow a device scratch array to hold at least count elements (synchronises
// the device before freeing a smaller one).
template <typename T>
static int grow(T** buf, int64_t* cap, int64_t count) {
  if (*cap >= count && *buf) return RM_OK;
  // headroom (at least 64 KB, at least twice the old size) so that calls of
  // slightly different sizes do not reallocate
  const int64_t want = std::max<int64_t>(std::max<int64_t>(count, 2 * *cap), (int64_t)(65536 / sizeof(T)));
  if (*buf) {
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
  }
  HIPCHK(hipMalloc((void**)buf, sizeof(T) * (size_t)want));
  *cap = want;
  return RM_OK;
}

// Workgroups of `kern` a CU holds with `lds` bytes of dynamic LDS, asked
// once per (kernel, LDS size) -- the query costs more than a launch.
static inline hipError_t occupancy(const void* kern, int threads, size_t lds, int* occ) {
  struct Entry { const void* k; size_t lds; int occ; };
  static std::mutex mu;
  static std::vector<Entry> cache;
  std::lock_guard<std::mutex> lock(mu);
  for (const Entry& e : cache)
    if (e.k == kern && e.lds == lds) {
      *occ = e.occ;
      return hipSuccess;
    }
  int o = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kern, threads, lds);
  if (e != hipSuccess) return e;
  if (o < 1) o = 1;
  cache.push_back({kern, lds, o});
  *occ = o;
  return hipSuccess;
}

// Krylov workspace of the plan, at least `bytes` (defined with the step kernels).
int ensure_ws(rm_plan* p, size_t bytes);
// Free a plan's subspace-launch scratch (rm_subspace.hip).
void sub_scratch_release(void* sc);
// Free the long-domain kernels' workspace (rm_coop.hip).
void coop_scratch_release(void* sc);
// A long-domain wavepacket: (column, lo, len, op), its domain's stored
// entries, longest row and field slots.
struct LongItem {
  int4 item;
  int64_t nent;
  int32_t maxrow, nslot;
};
// Long-domain wavepackets (rm_coop.hip): blk one workgroup each in one launch,
// coop the whole device each (one cooperative launch each).  d_ops, d_blk,
// d_coop: device copies of the jobs and of the two item lists; op0 (host,
// may be NULL): the one job by value instead of d_ops.  abort_preset: test
// hook, the cooperative launches start with their abort flag raised.
int long_run(void** scratch, const SubOp* d_ops, const int4* d_blk, const std::vector<LongItem>& blk,
             const int4* d_coop, const std::vector<LongItem>& coop, int32_t kcap, int32_t fixed_m,
             int32_t maxorder, double tol, bool f32, int n_cu, hipStream_t s, int32_t abort_preset = 0,
             const SubOp* op0 = nullptr, int32_t coop_grid = 0);
// Expokit propagation (rm_expokit.hip) of the items (column, lo, len) of one
// job with Krylov size m and tolerance tol; scratch grown in place.
int expokit_run(const SubOp& op, const std::vector<int4>& items, int32_t m, double tol, int n_cu, void** ws,
                size_t* ws_bytes, double** anorm, int64_t* anorm_cap, int4** d_items, int64_t* items_cap,
                hipStream_t s);
// sum[s] = sum over b of ex[s * ncols + b] (device, fixed order), on s.
int launch_expect_sum(int64_t nsteps, int64_t ncols, const double2* ex, double2* sum, hipStream_t s);
// Event-bracketed launch helpers for the plan's optional timing.
hipError_t timing_begin(rm_plan* p, hipStream_t s);
hipError_t timing_end(rm_plan* p, hipStream_t s);

// ---------------------------------------------------------------------------
// wave-level reductions (64-lane wavefronts)

constexpr int KMAX = RM_MAX_KRYLOV;
constexpr int TAYLOR_TERMS = 18;   // |h f T|_1 <= 1/2: truncation < 0.5^19/19! ~ 1.6e-23

// DPP row rotation of a double (both 32-bit halves); CTRL is the DPP control.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// v_permlane16_swap / v_permlane32_swap with the same value in both operands:
// every lane receives the (even, odd) row pair of its half, or the (lower,
// upper) half pair of the wave.
__device__ __forceinline__ void swap16_d(double v, double* even, double* odd) {
  const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  *even = __hiloint2double(hi[0], lo[0]);
  *odd = __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ void swap32_d(double v, double* lower, double* upper) {
  const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  *lower = __hiloint2double(hi[0], lo[0]);
  *upper = __hiloint2double(hi[1], lo[1]);
}

// Sum over the 64 lanes of a wave (all lanes active), bitwise identical on
// every lane: rotations by 8, 4, 2, 1 inside each 16-lane row pair every lane
// with partners whose partial sums are the same two operands (addition is
// commutative), then the row sums are combined in a fixed order,
// (row 0 + row 1) + (row 2 + row 3), through lane swaps between rows and
// halves (no scalar-register round trip).
__device__ __forceinline__ double wave_allsum(double v) {
  v += dpp_mov<0x128>(v);   // row_ror:8
  v += dpp_mov<0x124>(v);   // row_ror:4
  v += dpp_mov<0x122>(v);   // row_ror:2
  v += dpp_mov<0x121>(v);   // row_ror:1
  double e, o, lo, hi;
  swap16_d(v, &e, &o);
  swap32_d(e + o, &lo, &hi);
  return lo + hi;
}

// wave_allsum for values that are zero on lanes 32..63: row 0 + row 1 (adding
// the two zero rows changes nothing), on every lane.
__device__ __forceinline__ double wave_allsum32(double v) {
  v += dpp_mov<0x128>(v);
  v += dpp_mov<0x124>(v);
  v += dpp_mov<0x122>(v);
  v += dpp_mov<0x121>(v);
  double e, o, lo, hi;
  swap16_d(v, &e, &o);
  swap32_d(e + o, &lo, &hi);
  return lo;
}

// Sums over each half of the wave (lanes 0..31 and 32..63) for two problems
// held side by side: the operations of wave_allsum32 on each half, the lane's
// own half's sum returned.
// The two row sums of each half meet through v_permlane16_swap (the odd
// rows of one operand exchanged with the even rows of the other; with the
// same value in both, every lane gets its half's even-row and odd-row sums),
// so each half's sum is even + odd -- the readlane form's operands and order,
// without the round trip through scalar registers.
__device__ __forceinline__ double half_allsum32(double v) {
  v += dpp_mov<0x128>(v);
  v += dpp_mov<0x124>(v);
  v += dpp_mov<0x122>(v);
  v += dpp_mov<0x121>(v);
  double e, o;
  swap16_d(v, &e, &o);
  return e + o;
}

// sqrt(s) and 1/sqrt(s) for s > 0: the hardware reciprocal square root and
// two Newton steps.  Outside [2^-900, 2^900] s is first scaled by an exact
// power of two (2^1000 below, 2^-1000 above) so that they stay in range, and
// the results are scaled back exactly; inside it the scale is 1, so that
// range is computed straight away and the scaled form redone only on lanes
// outside it (the range test is off the dependent chain).  s == 0 gives no
// meaningful result (the caller's breakdown path replaces it).
__device__ __forceinline__ void sqrt_rcp_newton(double t, double* b, double* r_out) {
  double r = __builtin_amdgcn_rsq(t);
  double e = fma(-t * r, r, 1.0);
  r = fma(0.5 * r, e, r);
  e = fma(-t * r, r, 1.0);
  r = fma(0.5 * r, e, r);
  *b = t * r;
  *r_out = r;
}
__device__ __forceinline__ void sqrt_rcp_nb(double s, double* beta, double* ib) {
  sqrt_rcp_newton(s, beta, ib);
  const bool tiny = s < 0x1p-900, huge = s > 0x1p+900;
  if (tiny || huge) {
    const double up = tiny ? 0x1p+1000 : 0x1p-1000;
    const double hup = tiny ? 0x1p+500 : 0x1p-500;     // sqrt(up)
    const double hdn = tiny ? 0x1p-500 : 0x1p+500;     // 1 / sqrt(up)
    double b, r;
    sqrt_rcp_newton(s * up, &b, &r);
    *beta = b * hdn;
    *ib = r * hup;
  }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ double2 shfl2(double2 v, int src) {
  return make_double2(__shfl(v.x, src), __shfl(v.y, src));
}

// Order LDS and memory operations of the lanes of one wave: every lane's
// earlier writes are visible to every lane's later reads (no s_barrier, so it
// is safe where the waves of a workgroup follow different control flow).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// small dense Krylov pieces, run by the 64 lanes of one wave

// 1/j for the Taylor terms (j = 1..TAYLOR_TERMS; rounded at compile time), so
// a term costs a multiplication instead of a division on the serial chain.
__device__ __forceinline__ double taylor_rinv(int j) {
  constexpr double r[TAYLOR_TERMS + 1] = {0.0,        1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,
                                          1.0 / 5,    1.0 / 6,    1.0 / 7,  1.0 / 8,  1.0 / 9,
                                          1.0 / 10,   1.0 / 11,   1.0 / 12, 1.0 / 13, 1.0 / 14,
                                          1.0 / 15,   1.0 / 16,   1.0 / 17, 1.0 / 18};
  return r[j];
}

// Taylor terms for a scaled norm x <= 1/2: the smallest j with x^j / j! < 2^-60
// (x_j = (2^-60 j!)^(1/j) increasing: x^j / j! < 2^-60 iff x < x_j), at most
// TAYLOR_TERMS -- independent comparisons instead of a chain of divisions.
// They compare the high 32 bits of x and x_j as integers (for x >= 0 the bit
// pattern orders as the value), i.e. against x_j rounded down to 20 mantissa
// bits: at most one term more than the exact count, never fewer.  As integer
// comparisons they run on the scalar unit when x is the same on every lane
// (taylor_terms_hi of a lane read).
constexpr uint32_t TAYLOR_THR_HI[TAYLOR_TERMS - 1] = {
    0x3c300000u, 0x3e16a09eu, 0x3ebd12edu, 0x3f11b4f8u, 0x3f44d763u, 0x3f67f34au,
    0x3f823257u, 0x3f954b5cu, 0x3fa4e681u, 0x3fb21d6bu, 0x3fbca8e5u, 0x3fc527c8u,
    0x3fcd98d1u, 0x3fd3d6a8u, 0x3fd9b196u, 0x3fe02ca0u, 0x3fe3e3dbu};
__device__ __forceinline__ int taylor_terms_hi(uint32_t xh) {
  // xh >= t iff (t - 1 - xh) < 0: its sign bit (both below 2^31), summed as
  // plain integer arithmetic (no condition codes, no vector round trip)
  uint32_t n = 1;
#pragma unroll
  for (int j = 0; j < TAYLOR_TERMS - 1; ++j) n += (TAYLOR_THR_HI[j] - 1u - xh) >> 31;
  return (int)n;
}
// the high word of x without its sign (x = -0 counts as +0)
__device__ __forceinline__ uint32_t abs_hi(double x) { return (uint32_t)__double2hiint(x) & 0x7fffffffu; }
__device__ __forceinline__ int taylor_terms(double x) { return taylor_terms_hi(abs_hi(x)); }
// the counts of lane 0's and lane 32's x (the same on every lane of a half),
// on the scalar unit
__device__ __forceinline__ void taylor_terms_halves(double x, int* c0, int* c1) {
  *c0 = taylor_terms_hi((uint32_t)__builtin_amdgcn_readlane((int)abs_hi(x), 0));
  *c1 = taylor_terms_hi((uint32_t)__builtin_amdgcn_readlane((int)abs_hi(x), 32));
}

// c = exp(f*T) e_0 for the s x s tridiagonal T (diag alpha[0..s-1], off-diagonal
// beta[1..s-1], T[i-1,i] = T[i,i-1] = beta[i]).  Lane l owns rows l and l+64
// (s <= 128).  Scaling and stepping: exp(A)e0 = (exp(A/nsub))^nsub e0 with
// |A/nsub|_1 <= 1/2, each factor a Taylor sum.
static __device__ void expm_e0(int s, const double2* alpha, const double* beta, double2 f, double2* c) {
  const int l = threadIdx.x & 63;
  double2 a[2], bl[2], bu[2];   // A_ii, A_{i,i-1}, A_{i,i+1}
  double col1 = 0.0;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    int i = l + 64 * r;
    a[r] = make_double2(0.0, 0.0);
    bl[r] = make_double2(0.0, 0.0);
    bu[r] = make_double2(0.0, 0.0);
    if (i < s) {
      a[r] = cmul(f, alpha[i]);
      if (i >= 1) bl[r] = cscale(beta[i], f);
      if (i + 1 < s) bu[r] = cscale(beta[i + 1], f);
      // column i of |A|: |A_ii| + |A_{i-1,i}| + |A_{i+1,i}|, A symmetric off-diagonal
      col1 = fmax(col1, hypot(a[r].x, a[r].y) + hypot(bl[r].x, bl[r].y) + hypot(bu[r].x, bu[r].y));
    }
  }
  double norm1 = wave_max(col1);
  int nsub = 1;
  if (norm1 > 0.5) nsub = (int)ceil(2.0 * norm1);
  if (!(nsub >= 1) || nsub > (1 << 20)) nsub = 1;   // NaN / overflow: result becomes NaN anyway
  const double h = 1.0 / (double)nsub;
  double2 y[2];
  y[0] = make_double2(l == 0 ? 1.0 : 0.0, 0.0);
  y[1] = make_double2(0.0, 0.0);
  for (int sub = 0; sub < nsub; ++sub) {
    double2 term[2] = {y[0], y[1]};
    double2 acc[2] = {y[0], y[1]};
    for (int j = 1; j <= TAYLOR_TERMS; ++j) {
      // neighbours: row i-1 and i+1
      double2 t0_dn = shfl2(term[0], (l + 63) & 63);   // lane l-1, row l-1 (wraps at l=0)
      double2 t1_dn = shfl2(term[1], (l + 63) & 63);
      double2 t0_up = shfl2(term[0], (l + 1) & 63);
      double2 t1_up = shfl2(term[1], (l + 1) & 63);
      double2 prev0 = (l == 0) ? make_double2(0.0, 0.0) : t0_dn;
      double2 prev1 = (l == 0) ? t0_dn /* lane 63 row 63 */ : t1_dn;
      double2 next0 = (l == 63) ? t1_up /* lane 0 row 64 */ : t0_up;
      double2 next1 = (l == 63) ? make_double2(0.0, 0.0) : t1_up;
      const double sc = h * taylor_rinv(j);
      double2 n0 = cmul(a[0], term[0]);
      n0 = cfma(bl[0], prev0, n0);
      n0 = cfma(bu[0], next0, n0);
      double2 n1 = cmul(a[1], term[1]);
      n1 = cfma(bl[1], prev1, n1);
      n1 = cfma(bu[1], next1, n1);
      term[0] = cscale(sc, n0);
      term[1] = cscale(sc, n1);
      acc[0] = cadd(acc[0], term[0]);
      acc[1] = cadd(acc[1], term[1]);
    }
    y[0] = acc[0];
    y[1] = acc[1];
  }
  if (l < s) c[l] = y[0];
  if (l + 64 < s) c[l + 64] = y[1];
}

// DPP move of a double within 16-lane rows; lanes whose source falls outside
// the row receive 0 (bound control: no zero old value to set up first).  CTRL 0x111 (row_shr:1): lane l gets lane l-1;
// CTRL 0x101 (row_shl:1): lane l gets lane l+1.
template <int CTRL>
__device__ __forceinline__ double dpp_row_shift0(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Maximum over each 16-lane row (every lane of the row receives it).
__device__ __forceinline__ double row_max16(double v) {
  v = fmax(v, dpp_mov<0x128>(v));   // row_ror:8
  v = fmax(v, dpp_mov<0x124>(v));   // row_ror:4
  v = fmax(v, dpp_mov<0x122>(v));   // row_ror:2
  v = fmax(v, dpp_mov<0x121>(v));   // row_ror:1
  return v;
}

// Terms J.. of a factor of expm_e0_16, unrolled at compile time (1/J a
// constant of the term): term <- h/J (f T) term, acc += term on the lanes
// whose own count reaches J; nmax (uniform) ends the sequence.  The same
// operations in the same order as a loop over j <= nterm.
template <int J>
__device__ __forceinline__ void cplx_terms(int nmax, int nterm, double h, double2 a, double2 bl, double2 bu,
                                           double2& term, double2& acc) {
  if constexpr (J <= TAYLOR_TERMS) {
    if (J > nmax) return;
    const double2 prev = make_double2(dpp_row_shift0<0x111>(term.x), dpp_row_shift0<0x111>(term.y));
    const double2 next = make_double2(dpp_row_shift0<0x101>(term.x), dpp_row_shift0<0x101>(term.y));
    double2 nt = cmul(a, term);
    nt = cfma(bl, prev, nt);
    nt = cfma(bu, next, nt);
    constexpr double r = 1.0 / J;
    term = cscale(h * r, nt);
    if (J <= nterm) acc = cadd(acc, term);
    cplx_terms<J + 1>(nmax, nterm, h, a, bl, bu, term, acc);
  }
}

// exp(f*T) e_0 as expm_e0, for s <= 16: lane i (< s) of each 16-lane row owns
// row i and its neighbours arrive by DPP row shifts (no LDS traffic); the
// Taylor sum of each factor stops at the first term below 2^-60 of the
// bound (h |f T|_1)^j / j!, at most TAYLOR_TERMS terms.  Lanes 0..s-1 write c.
// PAIR: the two halves of the wave (lanes 0..31, 32..63) hold two independent
// problems; each lane passes its own half's alpha, beta and c, and the result
// is written by lanes (l & 31) < s of each half.  Every row computes the same
// values as in the single form (the row operations never cross rows).
template <bool PAIR = false>
static __device__ void expm_e0_16(int s, const double2* alpha, const double* beta, double2 f, double2* c) {
  const int l = threadIdx.x & 63;
  const int i = l & 15;
  double2 a = make_double2(0.0, 0.0), bl = a, bu = a;
  double col1 = 0.0;
  if (i < s) {
    a = cmul(f, alpha[i]);
    if (i >= 1) bl = cscale(beta[i], f);
    if (i + 1 < s) bu = cscale(beta[i + 1], f);
    // |re| + |im| >= |z|: a (slightly larger) bound for the scaling choice
    col1 = fabs(a.x) + fabs(a.y) + fabs(bl.x) + fabs(bl.y) + fabs(bu.x) + fabs(bu.y);
  }
  const double norm1 = row_max16(col1);
  int nsub = 1;
  if (norm1 > 0.5) nsub = (int)ceil(2.0 * norm1);
  if (!(nsub >= 1) || nsub > (1 << 20)) nsub = 1;
  const double h = 1.0 / (double)nsub;
  // each half's count (h norm1 is the same on its lanes); the larger one runs
  // (uniform, unrolled) and each lane adds only the terms of its own count
  int c0, c1;
  taylor_terms_halves(h * norm1, &c0, &c1);
  const int nterm = l < 32 ? c0 : c1;
  const int nmax = max(c0, c1);
  double2 y = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
  for (int sub = 0; sub < nsub; ++sub) {
    double2 term = y, acc = y;
    cplx_terms<1>(nmax, nterm, h, a, bl, bu, term, acc);
    y = acc;
  }
  if (PAIR) {
    if ((l & 31) < s) c[l & 31] = y;
  } else if (l < s) {
    c[l] = y;
  }
}

// exp(i phi T) e0 for a real symmetric tridiagonal T (diagonal alpha[].x,
// off-diagonal beta) and real phi: the case f = i phi of expm_e0_16 with a
// real Krylov matrix.  Same scaling rule, number of terms per problem and
// output lanes as expm_e0_16.  Without sub-steps (|phi T|_1 <= 1/2) the
// powers w_j = (phi T)^j e0 are real, and the sum is
// y = sum_j i^j w_j / j!: one real recursion, one DPP shift per neighbour and
// two FMAs on the serial chain per term, the weights 1/j! compile-time
// constants (the loop is unrolled); both halves run the larger number of
// terms and each adds only its own.  A half with |phi T|_1 > 1/2 takes
// sub-steps instead, each applying i h phi T to a complex vector with real
// coefficients; what one half computes never depends on the other.
// 1/j! rounded at compile time
constexpr double rfact_ct(int j) { return j <= 1 ? 1.0 : rfact_ct(j - 1) / (double)j; }

// Terms J.. of the real recursion of expm_e0_16_herm, unrolled at compile
// time (the weight 1/J! and the power of i are constants of the term): w <-
// (phi T) w, and y += i^J w / J! (one rounding) on the lanes whose own count
// reaches J; nmax (uniform) ends the sequence.
template <int J>
__device__ __forceinline__ void herm_terms(int nmax, int nterm, double a, double bl, double bu, double& w,
                                           double2& y) {
  if constexpr (J <= TAYLOR_TERMS) {
    if (J > nmax) return;
    const double prev = dpp_row_shift0<0x111>(w), next = dpp_row_shift0<0x101>(w);
    w = fma(bu, next, fma(bl, prev, a * w));
    constexpr double rf = ((J & 3) >= 2) ? -rfact_ct(J) : rfact_ct(J);   // i^J / J!: re for even J, im for odd
    double& yc = (J & 1) ? y.y : y.x;
    if (J <= nterm) yc = fma(rf, w, yc);
    herm_terms<J + 1>(nmax, nterm, a, bl, bu, w, y);
  }
}

// The exponential from this lane's row i = l % 16 of phi T (a = phi alpha_i,
// bl = phi beta_i, bu = phi beta_{i+1}, each 0 outside the s x s matrix) and
// the 1-norm of phi T (the same on every lane of a problem).
template <bool PAIR = false>
static __device__ void expm_e0_16_herm_row(int s, double a, double bl, double bu, double norm1, double2* c) {
  const int l = threadIdx.x & 63;
  const int i = l & 15;
  double2 y = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
  // each half decides for itself (the rows of a half are the rows of its
  // DPP rows, so the two halves may take different branches below); its
  // count comes from the scalar unit (norm1 is the same on its lanes), 0 for
  // a half that takes sub-steps
  const bool small = !(norm1 > 0.5);
  const uint64_t sm = __ballot(small);
  int c0, c1;
  taylor_terms_halves(norm1, &c0, &c1);
  if (!(sm & 1)) c0 = 0;
  if (!((sm >> 32) & 1)) c1 = 0;
  const int nterm_r = l < 32 ? c0 : c1;
  const int nmax = max(c0, c1);
  if (nmax > 0) {
    double w = y.x;
    herm_terms<1>(nmax, nterm_r, a, bl, bu, w, y);
  }
  if (!small) {
    int nsub = (int)ceil(2.0 * norm1);
    if (!(nsub >= 1) || nsub > (1 << 20)) nsub = 1;
    const double h = 1.0 / (double)nsub;
    const int nterm = taylor_terms(h * norm1);
    for (int sub = 0; sub < nsub; ++sub) {
      double2 term = y, acc = y;
      for (int j = 1; j <= nterm; ++j) {
        const double px = dpp_row_shift0<0x111>(term.x), py = dpp_row_shift0<0x111>(term.y);
        const double nx = dpp_row_shift0<0x101>(term.x), ny = dpp_row_shift0<0x101>(term.y);
        const double tx = fma(bu, nx, fma(bl, px, a * term.x));
        const double ty = fma(bu, ny, fma(bl, py, a * term.y));
        const double sc = h * taylor_rinv(j);
        term = make_double2(-ty * sc, tx * sc);
        acc = cadd(acc, term);
      }
      y = acc;
    }
  }
  if (PAIR) {
    if ((l & 31) < s) c[l & 31] = y;
  } else if (l < s) {
    c[l] = y;
  }
}
template <bool PAIR = false>
static __device__ void expm_e0_16_herm(int s, const double2* alpha, const double* beta, double phi, double2* c) {
  const int i = threadIdx.x & 15;
  double a = 0.0, bl = 0.0, bu = 0.0, col1 = 0.0;
  if (i < s) {
    a = phi * alpha[i].x;
    if (i >= 1) bl = phi * beta[i];
    if (i + 1 < s) bu = phi * beta[i + 1];
    col1 = fabs(a) + fabs(bl) + fabs(bu);
  }
  expm_e0_16_herm_row<PAIR>(s, a, bl, bu, row_max16(col1), c);
}

// Adaptive stopping metric after iteration k (one wave; tdse.py:473-476).
// Gram column k of the Krylov basis from the three-term relation
//   <V_i|V_k> = (<H V_i|V_{k-1}> - a_{k-1} G_{i,k-1} - b_{k-1} G_{i,k-2}) / b_k,
//   H V_i = b_{i+1} V_{i+1} + a_i V_i + b_i V_{i-1}
// (skipped after an exact breakdown, whose column the caller computed
// explicitly), then c = exp(f T_{k+1}) e0 and
//   conv = |u_k - u_{k-1}|^2 = d^H G d,  d = c - [c_prev, 0];
// finally c_prev = c.  Returns conv on every lane of the wave.
// FAST: exp(fT) e0 by expm_e0_16 while the Krylov dimension is <= 16.
// exp(f T_s) e0: the DPP-row variant for s <= 16, the general one beyond.
__device__ __forceinline__ void expm_e0_any(int s, const double2* alpha, const double* beta, double2 f,
                                            double2* c) {
  if (s <= 16)
    expm_e0_16<false>(s, alpha, beta, f, c);
  else
    expm_e0(s, alpha, beta, f, c);
}

template <bool FAST = false>
static __device__ double adaptive_metric(int k, bool breakdown, double2* G, int kc, const double2* s_alpha,
                                         const double* s_beta, double2 f, double2* s_c, double2* s_cprev) {
  const int tid = threadIdx.x & 63;
  const double beta = s_beta[k];
  if (!breakdown) {
    for (int i = tid; i < k; i += 64) {
      double2 hv;
      if (i == k - 1) {
        hv = cconj(s_alpha[k - 1]);
      } else {
        hv = cscale(s_beta[i + 1], G[(i + 1) * kc + (k - 1)]);
        hv = cfma(cconj(s_alpha[i]), G[i * kc + (k - 1)], hv);
        if (i > 0) hv = cadd(hv, cscale(s_beta[i], G[(i - 1) * kc + (k - 1)]));
      }
      double2 t = csub(hv, cmul(s_alpha[k - 1], G[i * kc + (k - 1)]));
      if (k >= 2) t = csub(t, cscale(s_beta[k - 1], G[i * kc + (k - 2)]));
      t = make_double2(t.x / beta, t.y / beta);
      G[i * kc + k] = t;
      G[k * kc + i] = cconj(t);
    }
  }
  if (tid == 0) G[k * kc + k] = make_double2(1.0, 0.0);
  __threadfence_block();
  if (FAST && k + 1 <= 16)
    expm_e0_16<false>(k + 1, s_alpha, s_beta, f, s_c);
  else
    expm_e0(k + 1, s_alpha, s_beta, f, s_c);
  __threadfence_block();
  double part = 0.0;
  const int s = k + 1;
  for (int i = tid; i < s; i += 64) {
    double2 di = csub(s_c[i], i < k ? s_cprev[i] : make_double2(0.0, 0.0));
    double2 r = make_double2(0.0, 0.0);
    for (int j = 0; j < s; ++j) {
      double2 dj = csub(s_c[j], j < k ? s_cprev[j] : make_double2(0.0, 0.0));
      r = cfma(G[i * kc + j], dj, r);
    }
    part += di.x * r.x + di.y * r.y;
  }
  part = wave_sum(part);
  __threadfence_block();
  for (int i = tid; i < s; i += 64) s_cprev[i] = s_c[i];
  return part;
}
