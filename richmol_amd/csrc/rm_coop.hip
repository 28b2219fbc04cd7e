// richmol_amd: propagation of wavepackets whose domain (the union of
// invariant subspaces their support touches) is longer than one wave holds.
//
// Two launch forms share one kernel body:
//  * block: one workgroup per wavepacket, many wavepackets per launch, the
//    workgroup's barriers between the Lanczos reductions;
//  * cooperative: every workgroup of the device on one wavepacket (one
//    cooperative launch per wavepacket, all workgroups resident), grid
//    barriers between the reductions -- for a few wavepackets on long domains.
// Either way a launch runs every step of the run.  Each step forms the step's
// operator A(F) = sum_s F_s B_s on the domain's entries (permuted CSR) --
// skipped when the step's field products equal those of the operator already
// built -- then runs the Lanczos expv of tdse.py:417-486 with the matvec
// spread over the waves (LPR lanes per row) and the vector work over every
// thread.  Dot products and norms are two-level sums (wave DPP, workgroup,
// and in the cooperative form the workgroups' partials added in a fixed order
// by every workgroup, so all of them take the same decisions).  A Lanczos
// iteration needs two barriers: after the matvec (alpha) and after the
// three-term update (beta); the basis vector V_k = W/beta is formed inside
// the matvec from W, so the beta barrier also publishes the matvec input.
// The cooperative register form replaces both by slot sums (coop_sum): the
// alpha sum orders nothing but itself, the beta sum is fenced.
//
// Cooperative form with every wave holding at most one group of RPW rows and
// rows of at most LPR*UR entries: a lane keeps the columns and the step's
// values of its entries in registers for the whole step (the matvec is then
// gathers and FMAs only); otherwise the entries are read at every matvec
// (from the step's operator array, or, with one field slot, straight from
// B_0 scaled by the slot's field product).

#include <algorithm>
#include <climits>

#include "rm_common.h"

namespace {

enum { GRID_REG = 0, GRID_MEM = 1, BLOCK = 2 };

// threads per workgroup (256 VGPRs per lane)
template <int MODE>
constexpr int ct_of() { return 512; }

constexpr int LPR = 16;          // lanes per row in the matvec (one DPP row)
constexpr int RPW = 64 / LPR;    // rows per wave in the matvec
constexpr int UR = 12;           // register-resident entries per lane
// memory-mode matvec: row groups per pass and entries per lane and row
// loaded together (FP32 operands: 4 x 4; FP64 keeps 2 x 8 within 256 VGPRs)
template <bool F32>
constexpr int mg_of() { return F32 ? 4 : 2; }
template <bool F32>
constexpr int umg_of() { return F32 ? 4 : 8; }
constexpr int GL = 32;           // Gram matrix order held in LDS (adaptive)
constexpr int XL_ROWS = 2048;    // block form: matvec input staged in LDS up to this many rows
constexpr unsigned SPIN_LIMIT = 1u << 22;   // polls of a grid barrier before the launch gives up
constexpr int BAR_LINES = 8;                // arrival counters of a grid barrier
constexpr int BAR_STRIDE = 32;              // u64 between counters (one 256-byte line each)
// an empty slot of the slot sums' partials: a signalling-NaN bit pattern,
// which no floating-point operation produces
constexpr unsigned long long LSENT = 0x7FF4DEAD0000BEEFull;
constexpr int LRING = 4;   // slot sums use LRING sets of slots in turn
// forms of a cooperative sum (coop_sum)
enum { SUM_BARRIER = 0, SUM_SCALAR = 1, SUM_FENCED = 2 };

struct LongArgs {
  const SubOp* ops;        // the jobs, or NULL: the one job is op0
  SubOp op0;               // the job of a single-job launch, by value
  const int4* items;       // (column, lo, len, op)
  int32_t nitems;
  char* ws;                // workspace: one region (cooperative) or one per workgroup (block)
  int64_t wg_bytes;        // bytes per workgroup region (block form)
  int64_t offV, offWb, offWf, offY, offS, offP, offA;   // byte offsets in a region
  double* part;            // [2][grid][4] workgroup partial sums (cooperative)
  double* lpart;           // [LRING][grid][4] partials of the slot sums (LSENT at launch)
  double2* gram;           // [grid][kcap * kcap] (adaptive, orders beyond GL)
  double2* cscr;           // [grid][2 * kcap] (adaptive, orders beyond GL)
  unsigned long long* bar; // grid-barrier arrivals, BAR_LINES counters (zero at launch)
  int32_t* abort_flag;     // set when a barrier wait exceeds SPIN_LIMIT polls
  int64_t lmax;            // vector capacity of a region
  int32_t kcap, fixed_m, maxorder;
  double tol;
};

struct SyncState {
  unsigned long long target = 0;
  int epoch = 0;       // LDS half of the next sum
  int pepoch = 0;      // half of `part` of the next grid sum
  int nsl = 0;         // slot sums so far (the next one uses slot set nsl % LRING)
  int pend = -1;       // the slot sum whose slots this workgroup has not emptied yet
};

// Empties this workgroup's slots of slot sum ss.pend.  Called when every
// workgroup is past its reads of that sum: at the end of the next grid
// barrier or slot sum.
__device__ __forceinline__ void slot_release(const LongArgs& a, SyncState& ss) {
  if (ss.pend >= 0) {
    if (threadIdx.x < 4)
      __hip_atomic_store((unsigned long long*)a.lpart + (size_t)(ss.pend % LRING) * gridDim.x * 4 +
                             blockIdx.x * 4 + threadIdx.x,
                         LSENT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ss.pend = -1;
  }
}

__device__ __forceinline__ double2 xadd(double2 a, double2 b) { return cadd(a, b); }
__device__ __forceinline__ float2 xadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

// Grid barrier on BAR_LINES arrival counters that only grow, one per 256-byte
// line: workgroup g arrives on line g % BAR_LINES, so barrier number t
// completes when line c holds t * (number of workgroups on line c).
// Spreading the arrivals keeps the atomics of ~100 workgroups from queueing
// on one address.  Every wave first waits for its own memory operations; one
// thread per workgroup then publishes the workgroup's writes (agent-scope
// release) and arrives; lanes 0..BAR_LINES-1 of wave 0 each poll one line
// and the wave leaves when all lines are complete, then acquires (agent
// scope; the waves of a workgroup share the CU's caches).  A wait that
// exceeds SPIN_LIMIT polls (a workgroup that never arrives) sets the abort
// flag; from then on barriers no longer wait and the launch runs to its end,
// its results marked invalid.
__device__ __forceinline__ void grid_barrier(const LongArgs& a, SyncState& ss, int* s_ab) {
  ss.target += 1;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x < 64 && *s_ab == 0) {
    const int lane = threadIdx.x;
    if (lane == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(a.bar + BAR_STRIDE * (blockIdx.x % BAR_LINES), 1ull, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned c = (unsigned)lane % BAR_LINES;
    const unsigned long long need =
        lane < BAR_LINES ? ss.target * (unsigned long long)((gridDim.x + BAR_LINES - 1 - c) / BAR_LINES) : 0ull;
    unsigned spins = 0;
    while (!__all(lane >= BAR_LINES ||
                  __hip_atomic_load(a.bar + BAR_STRIDE * c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need)) {
      __builtin_amdgcn_s_sleep(1);
      ++spins;
      if ((spins & 255u) == 0 &&
          (spins >= SPIN_LIMIT || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        if (lane == 0) {
          __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          *s_ab = 1;
        }
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  slot_release(a, ss);
}

// Every thread of the cooperating set (the workgroup, or the grid) waits.
template <int MODE>
__device__ __forceinline__ void sync_all(const LongArgs& a, SyncState& ss, int* s_ab) {
  if constexpr (MODE == BLOCK)
    __syncthreads();
  else
    grid_barrier(a, ss, s_ab);
}

// Sum of NQ values over every thread of the cooperating set; every thread
// returns the same sums (fixed order at every level).  s_red: [2][CWV][4],
// the halves used in turn.
//
// Slot forms (cooperative form only).  Each workgroup stores its partials
// with agent-scope atomic stores into slots that hold LSENT while empty; wave
// 0 of every workgroup polls all slots (agent-scope atomic loads) until none
// is empty and adds them in the barrier form's order (same result).
// SUM_SCALAR orders nothing but the sum itself: after it a workgroup reads
// only vector entries it wrote.  SUM_FENCED also publishes the workgroups'
// writes, as a grid barrier does: every wave waits for its stores, the
// storing thread releases at agent scope before its slot store, and wave 0
// acquires at agent scope after the poll (release fence, relaxed store ->
// relaxed load, acquire fence).  Slot sums use LRING slot sets in turn.  A
// workgroup empties its slots of a sum at the end of the next grid barrier or
// slot sum (every workgroup is past its reads by then: it stored its next
// partial, which depends on them, or it arrived at the barrier), and the
// callers never run two SUM_SCALAR sums without a grid barrier or a
// SUM_FENCED sum between them, so between that emptying and the next use of
// the set (LRING sums on) lies a fenced event that publishes it.  The poll is
// bounded as the grid barrier's wait is.
template <int NQ, int CWV, int MODE, int FORM = SUM_BARRIER>
__device__ void coop_sum(double (&v)[NQ], const LongArgs& a, SyncState& ss, double (*s_red)[4],
                         double* s_tot, int* s_ab) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_allsum(v[q]);
  double (*red)[4] = s_red + ss.epoch * CWV;
  ss.epoch ^= 1;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NQ; ++q) red[wv][q] = v[q];
  if constexpr (MODE != BLOCK && FORM == SUM_FENCED) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if constexpr (MODE == BLOCK) {
    // the other half is written by the next sum; this one only after the next sum's barrier
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double t = 0.0;
      for (int w = 0; w < CWV; ++w) t += red[w][q];
      v[q] = t;
    }
  } else if constexpr (FORM != SUM_BARRIER) {
    const int n = ss.nsl++;
    unsigned long long* lp = (unsigned long long*)a.lpart + (size_t)(n % LRING) * gridDim.x * 4;
    if (tid < NQ) {
      double t = 0.0;
      for (int w = 0; w < CWV; ++w) t += red[w][tid];
      if constexpr (FORM == SUM_FENCED) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(lp + blockIdx.x * 4 + tid, (unsigned long long)__double_as_longlong(t), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    if (wv == 0) {
      double acc[NQ];
      unsigned spins = 0;
      for (;;) {
        bool full = true;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          acc[q] = 0.0;
          for (int g = lane; g < (int)gridDim.x; g += 64) {
            const unsigned long long u = __hip_atomic_load(lp + g * 4 + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            full = full && u != LSENT;
            acc[q] += __longlong_as_double((long long)u);
          }
        }
        if (__all(full) || *s_ab) break;
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if ((spins & 255u) == 0 &&
            (spins >= SPIN_LIMIT || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          if (lane == 0) {
            __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *s_ab = 1;
          }
          break;
        }
      }
      if constexpr (FORM == SUM_FENCED) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const double t = wave_allsum(acc[q]);
        if (lane == 0) s_tot[q] = t;
      }
    }
    __syncthreads();
    slot_release(a, ss);
    ss.pend = n;
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = s_tot[q];
  } else {
    double* part = a.part + (size_t)ss.pepoch * gridDim.x * 4;
    ss.pepoch ^= 1;
    if (tid < NQ) {
      double t = 0.0;
      for (int w = 0; w < CWV; ++w) t += red[w][tid];
      part[blockIdx.x * 4 + tid] = t;
    }
    grid_barrier(a, ss, s_ab);
    if (wv == 0) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        double acc = 0.0;
        for (int g = lane; g < (int)gridDim.x; g += 64) acc += part[g * 4 + q];
        acc = wave_allsum(acc);
        if (lane == 0) s_tot[q] = acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = s_tot[q];
  }
}

// Sum over the 16 lanes of a DPP row, every lane of the row receiving it.
__device__ __forceinline__ double row_sum16(double v) {
  v += dpp_mov<0x128>(v);   // row_ror:8
  v += dpp_mov<0x124>(v);   // row_ror:4
  v += dpp_mov<0x122>(v);   // row_ror:2
  v += dpp_mov<0x121>(v);   // row_ror:1
  return v;
}

__device__ __forceinline__ double inv_beta(double beta) {
  // hardware reciprocal and two Newton steps (within an ulp of the quotient)
  double ib = __builtin_amdgcn_rcp(beta);
  ib = fma(fma(-beta, ib, 1.0), ib, ib);
  ib = fma(fma(-beta, ib, 1.0), ib, ib);
  return ib;
}

// Dynamic LDS of the XL stage (matvec input and row offsets), rounded up to
// 256 bytes: the LV vectors follow it.
template <bool F32>
constexpr size_t xl_bytes() {
  return ((F32 ? sizeof(float2) : sizeof(double2)) * XL_ROWS + sizeof(int) * (XL_ROWS + 1) + 255) & ~(size_t)255;
}

// XL (block form, domains of at most XL_ROWS rows): the matvec input is
// staged in LDS by the threads that write it, and gathered from there.
// LV (with XL): the Krylov basis and the step's vectors (V, W, Y, S, P) sit
// in LDS after the stage instead of the workgroup's global region (the host
// takes it when they fit); the step's operator array stays global.
template <int MODE, bool FIXED, bool F32, bool XL = false, bool LV = false>
__global__ __launch_bounds__(ct_of<MODE>()) void k_long(LongArgs a) {
  using XT = typename XType<F32>::T;
  static_assert(!XL || MODE == BLOCK, "LDS-staged input is for the block form");
  static_assert(!LV || XL, "LDS vectors come with the LDS-staged input");
  extern __shared__ char s_dyn[];
  XT* s_x = (XT*)s_dyn;                                   // [XL_ROWS] matvec input
  int* s_rp = (int*)(s_dyn + sizeof(XT) * XL_ROWS);       // [XL_ROWS + 1] row offsets from e_lo
  constexpr int CT = ct_of<MODE>();
  constexpr int CWV = CT / 64;
  constexpr bool REG = MODE == GRID_REG;
  __shared__ double2 s_alpha[KMAX];
  __shared__ double s_beta[KMAX + 1];
  __shared__ double2 s_cf[KMAX];
  __shared__ double2 s_gram[FIXED ? 1 : GL * GL];
  __shared__ double2 s_cv[FIXED ? 1 : 2 * GL];
  __shared__ double s_red[2 * CWV][4];
  __shared__ double s_tot[4];
  __shared__ double s_conv;
  __shared__ int s_ab;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the cooperating set: this workgroup (block form) or the grid
  const int G = MODE == BLOCK ? 1 : gridDim.x;
  const int g = MODE == BLOCK ? 0 : blockIdx.x;
  const int64_t nthr = (int64_t)G * CT;
  const int64_t t0 = (int64_t)g * CT + tid;
  const int nwaves = G * CWV, gw = g * CWV + wv;
  const int sub = lane % LPR, rw = lane / LPR;
  const int kc = a.kcap;
  char* base = a.ws + (MODE == BLOCK ? (int64_t)blockIdx.x * a.wg_bytes : 0);
  char* vbase = LV ? s_dyn + xl_bytes<F32>() : base;      // the vectors' region (offsets below offA)
  double2* V = (double2*)(vbase + a.offV);
  double2* Wb = (double2*)(vbase + a.offWb);
  float2* Wf = (float2*)(vbase + a.offWf);
  double2* Y = (double2*)(vbase + a.offY);
  double2* S = (double2*)(vbase + a.offS);
  double2* P = (double2*)(vbase + a.offP);
  XT* Ast = (XT*)(base + a.offA);
  const XT* Xin;
  if constexpr (XL)
    Xin = s_x;
  else if constexpr (F32)
    Xin = Wf;
  else
    Xin = Wb;
  const int64_t lmax = a.lmax;
  const double2 zero = make_double2(0.0, 0.0);
  SyncState ss;
  if (tid == 0) s_ab = 0;
  __syncthreads();

  for (int itm = MODE == BLOCK ? blockIdx.x : 0; itm < a.nitems; itm += MODE == BLOCK ? gridDim.x : 1) {
    const int4 item = a.items[itm];
    const SubOp& op = a.ops ? a.ops[item.w] : a.op0;
    const int b = item.x, lo = item.y, len = item.z;
    // reused domains that failed their check: the wavepacket is left as it is
    // (the same decision in every thread of the cooperating set)
    if (op.stale && *op.stale == op.call_id) {
      if (t0 == 0 && op.status) op.status[b - op.kcol0] |= RM_STALE_DOMAIN;
      continue;
    }
    const int64_t e_lo = op.rp[lo];
    const int ngroups = (len + RPW - 1) / RPW;
    // per-row loops: REG runs each row on its matvec owner (lane sub == 0 of
    // row gw * RPW + rw), so that after a scalar-only sum a workgroup reads
    // only vector entries it wrote itself; the other forms stride the rows
    // over every thread of the cooperating set
    const int r_beg = !REG ? (int)t0 : (sub == 0 && gw * RPW + rw < len) ? gw * RPW + rw : len;
    const int r_step = REG ? len : (int)nthr;
    for (int64_t i = r_beg; i < len; i += r_step) {
      const int orig = op.perm[lo + i];
      S[i] = op.psi[(int64_t)b * op.ld + orig];
      P[i] = op.phase ? op.phase[orig] : make_double2(1.0, 0.0);
    }
    if constexpr (XL)
      for (int64_t i = t0; i <= len; i += nthr) s_rp[i] = (int)(op.rp[lo + i] - e_lo);
    // first entry of domain row r (XL: from LDS, published by the first barrier of a step)
    auto row_start = [&](int r) __attribute__((always_inline)) -> int64_t {
      if constexpr (XL)
        return e_lo + s_rp[r];
      else
        return op.rp[lo + r];
    };
    // this lane's row in the matvec (REG: the wave's only row group) and its entries
    const int rreg = gw * RPW + rw;
    int64_t ebeg = 0, eend = 0;
    if (REG && rreg < len) {
      ebeg = op.rp[lo + rreg];
      eend = op.rp[lo + rreg + 1];
    }
    int cI[REG ? UR : 1];
    XT Av[REG ? UR : 1];
    if constexpr (REG) {
#pragma unroll
      for (int u = 0; u < UR; ++u) {
        const int64_t e = ebeg + sub + (int64_t)u * LPR;
        cI[u] = e < eend ? op.ccol[e] - lo : 0;
        Av[u] = to_x<F32>(zero);
      }
    }
    // one field slot (memory mode): the matvec reads B_0 and scales the row sums
    const bool direct = !REG && op.nslot == 1;
    int built = -1;   // step whose operator the registers / Ast hold
    // <psi|O|psi> after step `step` (the state in S): a barrier publishes S,
    // every thread takes rows of the domain, one sum (a uniform branch: only
    // jobs with an observable)
    auto observe = [&](int step) __attribute__((always_inline)) {
      if (op.ow <= 0) return;
      sync_all<MODE>(a, ss, &s_ab);
      double d[2] = {0.0, 0.0};
      for (int64_t i = t0; i < len; i += nthr) {
        double2 y = zero;
        for (int j = 0; j < op.ow; ++j) {
          const int64_t q = (int64_t)j * op.n + lo + i;
          const int c = op.ocol[q] - lo;
          if (c >= 0 && c < len) y = cfma(op.oval[q], S[c], y);
        }
        const double2 v = S[i];
        d[0] += v.x * y.x + v.y * y.y;
        d[1] += v.x * y.y - v.y * y.x;
      }
      coop_sum<2, CWV, MODE>(d, a, ss, s_red, s_tot, &s_ab);
      if (t0 == 0 && op.expect) op.expect[(int64_t)step * op.expect_ld + b - op.kcol0] = make_double2(d[0], d[1]);
    };

    for (int step = 0; step < op.nsteps; ++step) {
      if (!step_on(op, step)) {
        // no field: the expv is skipped (tdse.py:377), phases only
        if (op.phase)
          for (int64_t i = r_beg; i < len; i += r_step) S[i] = cmul(cmul(S[i], P[i]), P[i]);
        if (t0 == 0 && op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = 0;
        observe(step);
        continue;
      }
      const double2* fp = op.table + (int64_t)step * op.nslot;
      // the step's operator on this lane's entries, unless the field products
      // are those of the operator already built
      bool same = built >= 0 || direct;
      if (built >= 0) {
        const double2* fb = op.table + (int64_t)built * op.nslot;
        for (int s = 0; s < op.nslot; ++s) same = same && fp[s].x == fb[s].x && fp[s].y == fb[s].y;
      }
      if (!same) {
        built = step;
        if constexpr (REG) {
#pragma unroll
          for (int u = 0; u < UR; ++u) {
            const int64_t e = ebeg + sub + (int64_t)u * LPR;
            double2 v = zero;
            if (e < eend)
              for (int s = 0; s < op.nslot; ++s) v = cfma(fp[s], op.cval[(int64_t)s * op.nnz + e], v);
            Av[u] = to_x<F32>(v);
          }
        } else {
          for (int rg = gw; rg < ngroups; rg += nwaves) {
            const int r = rg * RPW + rw;
            if (r < len) {
              const int64_t e1 = op.rp[lo + r + 1];
              for (int64_t e = op.rp[lo + r] + sub; e < e1; e += LPR) {
                double2 v = zero;
                for (int s = 0; s < op.nslot; ++s) v = cfma(fp[s], op.cval[(int64_t)s * op.nnz + e], v);
                Ast[e - e_lo] = to_x<F32>(v);
              }
            }
          }
        }
      }
      const XT f0 = to_x<F32>(load_or_zero(direct, fp));
      // V_0 = phi = P psi (unnormalised, tdse.py:443-446), the first matvec input
      for (int64_t i = r_beg; i < len; i += r_step) {
        const double2 v = op.phase ? cmul(S[i], P[i]) : S[i];
        V[i] = v;
        Wb[i] = v;
        if constexpr (F32) Wf[i] = to_x<true>(v);
        if constexpr (XL) s_x[i] = to_x<F32>(v);
      }
      sync_all<MODE>(a, ss, &s_ab);

      // Gram matrix and coefficient scratch: LDS up to order GL, then global
      double2* Gm = s_gram;
      double2* cv = s_cv;
      int g_ld = GL;
      int k = 1;
      double conv = 1.0, beta = 0.0, ib = 1.0;
      bool brk = false;
      for (int it = 0;; ++it) {
        double2* Vit = V + (int64_t)it * lmax;
        // w = A (W / beta) = A V_it on this wave's rows; alpha = vdot(w, V_it)
        double red[3] = {0.0, 0.0, 0.0};
        auto finish_row = [&](int r, XT acc) __attribute__((always_inline)) {
          double2 w = to_d2(acc);
          w.x = row_sum16(w.x);
          w.y = row_sum16(w.y);
          if (sub == 0 && r < len) {
            w = cscale(ib, w);
            double2 wr;
            if constexpr (XL && !F32)
              wr = s_x[r];   // W itself
            else
              wr = Wb[r];
            const double2 v = cscale(ib, wr);
            Vit[r] = v;
            Y[r] = w;
            red[0] += w.x * v.x + w.y * v.y;
            red[1] += w.x * v.y - w.y * v.x;
            red[2] += v.x * v.x + v.y * v.y;
          }
        };
        if constexpr (REG) {
          if (gw < ngroups) {
            XT acc = to_x<F32>(zero);
#pragma unroll
            for (int u = 0; u < UR; ++u) acc = xfma(Av[u], Xin[cI[u]], acc);
            finish_row(rreg, acc);
          }
        } else {
          auto ecol = [&](int64_t q, int64_t e1) __attribute__((always_inline)) -> int {
            return q < e1 ? op.ccol[q] - lo : 0;
          };
          auto eval = [&](int64_t q, int64_t e1) __attribute__((always_inline)) -> XT {
            if (q >= e1) return to_x<F32>(make_double2(0.0, 0.0));
            if (!direct) return Ast[q - e_lo];
            if constexpr (F32)
              return op.cval32[q];   // the FP32 rounding of B_0, prepared with the plan
            else
              return op.cval[q];
          };
          // MG row groups per pass, their loads in flight together; UMG entries
          // per lane and row at a time
          constexpr int MG = mg_of<F32>(), UMG = umg_of<F32>();
          for (int rg = gw; rg < ngroups; rg += MG * nwaves) {
            int rr[MG];
            int64_t r0[MG], r1[MG];
            XT acc[MG];
            int64_t lm = 0;
#pragma unroll
            for (int g = 0; g < MG; ++g) {
              rr[g] = (rg + g * nwaves) * RPW + rw;
              r0[g] = rr[g] < len ? row_start(rr[g]) : 0;
              r1[g] = rr[g] < len ? row_start(rr[g] + 1) : 0;
              acc[g] = to_x<F32>(zero);
              lm = std::max(lm, r1[g] - r0[g]);
            }
            // the wave's longest row: entries of a pass beyond it are skipped
            // (a wave-uniform bound, so short rows cost no idle slots)
            int lmw = (int)lm;
            lmw = max(lmw, __shfl_xor(lmw, 16));
            lmw = max(lmw, __shfl_xor(lmw, 32));
            lmw = __builtin_amdgcn_readfirstlane(lmw);
            for (int64_t t = sub; t < lm; t += (int64_t)LPR * UMG) {
              const int tw = (int)(t - sub);   // the same on every lane
              int cc[MG][UMG];
              XT vv[MG][UMG];
#pragma unroll
              for (int u = 0; u < UMG; ++u)
                if (tw + u * LPR < lmw)
#pragma unroll
                  for (int g = 0; g < MG; ++g) {
                    cc[g][u] = ecol(r0[g] + t + (int64_t)u * LPR, r1[g]);
                    vv[g][u] = eval(r0[g] + t + (int64_t)u * LPR, r1[g]);
                  }
              XT pp[MG];
#pragma unroll
              for (int g = 0; g < MG; ++g) pp[g] = to_x<F32>(zero);
#pragma unroll
              for (int u = 0; u < UMG; ++u)
                if (tw + u * LPR < lmw)
#pragma unroll
                  for (int g = 0; g < MG; ++g) pp[g] = xfma(vv[g][u], Xin[cc[g][u]], pp[g]);
#pragma unroll
              for (int g = 0; g < MG; ++g) acc[g] = xadd(acc[g], pp[g]);
            }
#pragma unroll
            for (int g = 0; g < MG; ++g) {
              if (direct) acc[g] = xfma(f0, acc[g], to_x<F32>(zero));
              finish_row(rr[g], acc[g]);
            }
          }
        }
        // alpha (and |V_0|^2): REG reads after it only its own rows
        constexpr int FA = REG ? SUM_SCALAR : SUM_BARRIER;
        if (it == 0 && !FIXED) {
          coop_sum<3, CWV, MODE, FA>(red, a, ss, s_red, s_tot, &s_ab);
        } else {
          double r2[2] = {red[0], red[1]};
          coop_sum<2, CWV, MODE, FA>(r2, a, ss, s_red, s_tot, &s_ab);
          red[0] = r2[0];
          red[1] = r2[1];
        }
        const double2 al = make_double2(red[0], red[1]);
        if (tid < 64) s_alpha[it] = al;
        // W = w - alpha V_it - beta V_{it-1}, and |W|^2
        const double2* Vprev = V + (int64_t)(it > 0 ? it - 1 : 0) * lmax;
        double bs = 0.0;
        for (int64_t i = r_beg; i < len; i += r_step) {
          double2 wn = csub(Y[i], cmul(al, Vit[i]));
          if (it > 0) wn = csub(wn, cscale(beta, Vprev[i]));
          Wb[i] = wn;
          if constexpr (F32) Wf[i] = to_x<true>(wn);
          if constexpr (XL) s_x[i] = to_x<F32>(wn);
          bs += wn.x * wn.x + wn.y * wn.y;
        }
        k = it + 1;
        if (!FIXED) {
          if (it == 0) {
            if (tid == 0) {
              Gm[0] = make_double2(red[2], 0.0);
              cv[g_ld] = make_double2(1.0, 0.0);   // c_prev = [1]
            }
          } else if (wv == 0) {
            wave_sync();
            const double c = adaptive_metric<true>(it, brk, Gm, g_ld, s_alpha, s_beta, op.f, cv, cv + g_ld);
            if (lane == 0) s_conv = c;
          }
          __syncthreads();
          if (it > 0) conv = s_conv;
        }
        const bool live = FIXED ? k < a.fixed_m : (k < a.maxorder && conv > a.tol);
        if (!live) break;
        if (!FIXED && k == GL) {
          // the next Gram column has index GL: matrix and coefficients move to global memory
          double2* G2 = a.gram + (int64_t)blockIdx.x * kc * kc;
          double2* c2 = a.cscr + (int64_t)blockIdx.x * 2 * kc;
          for (int q = tid; q < GL * GL; q += CT) G2[(q / GL) * kc + q % GL] = s_gram[q];
          for (int q = tid; q < GL; q += CT) {
            c2[q] = s_cv[q];
            c2[kc + q] = s_cv[GL + q];
          }
          Gm = G2;
          cv = c2;
          g_ld = kc;
          __threadfence_block();
          __syncthreads();
        }
        // beta_k = |W| (tdse.py:453); publishes W, the next matvec input
        double bsum[1] = {bs};
        coop_sum<1, CWV, MODE, REG ? SUM_FENCED : SUM_BARRIER>(bsum, a, ss, s_red, s_tot, &s_ab);
        beta = sqrt(bsum[0]);
        if (tid < 64) s_beta[k] = beta;
        brk = !(beta != 0.0);
        if (!brk) {
          ib = inv_beta(beta);
          continue;
        }
        // exact breakdown: Gram-Schmidt of ones against the basis (tdse.py:459-465);
        // the restart vector becomes the next matvec input with ib = 1
        ib = 1.0;
        for (int64_t i = t0; i < len; i += nthr) Wb[i] = make_double2(1.0, 0.0);
        for (int j = 0; j < k; ++j) {
          const double2* Vj = V + (int64_t)j * lmax;
          double d[2] = {0.0, 0.0};
          for (int64_t i = t0; i < len; i += nthr) {
            const double2 x = Vj[i], y = Wb[i];
            d[0] += x.x * y.x + x.y * y.y;
            d[1] += x.x * y.y - x.y * y.x;
          }
          coop_sum<2, CWV, MODE>(d, a, ss, s_red, s_tot, &s_ab);
          const double2 proj = make_double2(d[0], d[1]);
          for (int64_t i = t0; i < len; i += nthr) Wb[i] = csub(Wb[i], cmul(proj, Vj[i]));
        }
        double nv[1] = {0.0};
        for (int64_t i = t0; i < len; i += nthr) nv[0] += Wb[i].x * Wb[i].x + Wb[i].y * Wb[i].y;
        coop_sum<1, CWV, MODE>(nv, a, ss, s_red, s_tot, &s_ab);
        const double nrm = sqrt(nv[0]);
        // nrm == 0: the basis spans the domain; the reference's restart vector
        // lies outside it and enters with coefficient exactly zero
        double2* Vk = V + (int64_t)k * lmax;
        for (int64_t i = t0; i < len; i += nthr) {
          const double2 v = nrm > 0.0 ? make_double2(Wb[i].x / nrm, Wb[i].y / nrm) : zero;
          Wb[i] = v;
          Vk[i] = v;
          if constexpr (F32) Wf[i] = to_x<true>(v);
          if constexpr (XL) s_x[i] = to_x<F32>(v);
        }
        if (!FIXED) {
          // Gram column of the restart vector, explicit dots
          for (int j = 0; j < k; ++j) {
            const double2* Vj = V + (int64_t)j * lmax;
            double d[2] = {0.0, 0.0};
            for (int64_t i = t0; i < len; i += nthr) {
              const double2 x = Vj[i], y = Vk[i];
              d[0] += x.x * y.x + x.y * y.y;
              d[1] += x.x * y.y - x.y * y.x;
            }
            coop_sum<2, CWV, MODE>(d, a, ss, s_red, s_tot, &s_ab);
            if (tid == 0) {
              Gm[j * g_ld + k] = make_double2(d[0], d[1]);
              Gm[k * g_ld + j] = make_double2(d[0], -d[1]);
            }
          }
        }
        sync_all<MODE>(a, ss, &s_ab);   // the restart vector is the next matvec's input
      }

      // u = sum_j c_j V_j with c = exp(f T) e0; psi = P u
      if (wv == 0) {
        wave_sync();
        if (k == 1) {
          if (lane == 0) s_cf[0] = make_double2(1.0, 0.0);
        } else if (FIXED) {
          expm_e0_any(k, s_alpha, s_beta, op.f, s_cf);
        } else {
          for (int j = lane; j < k; j += 64) s_cf[j] = cv[j];   // exp(f T) e0 of the last metric
        }
      }
      __syncthreads();
      for (int64_t i = r_beg; i < len; i += r_step) {
        double2 u = zero;
        for (int j = 0; j < k; ++j) u = cfma(s_cf[j], V[(int64_t)j * lmax + i], u);
        S[i] = op.phase ? cmul(u, P[i]) : u;
      }
      if (t0 == 0) {
        if (op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = k;
        if (!FIXED && k >= a.maxorder && op.status) op.status[b - op.kcol0] |= RM_NOT_CONVERGED;
      }
      observe(step);
      // the next step's V_0 is written by the same threads; s_cf and the LDS
      // Gram matrix are rewritten only after the next step's first barrier
    }
    for (int64_t i = r_beg; i < len; i += r_step) op.psi[(int64_t)b * op.ld + op.perm[lo + i]] = S[i];
    if (MODE != BLOCK && t0 == 0 && op.status &&
        (s_ab || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
      op.status[b - op.kcol0] |= RM_DEVICE_TIMEOUT;
    // block form: the next item rewrites this workgroup's region only after
    // these reads (same threads) and the workgroup barrier of its first step
    __syncthreads();
  }
}

__global__ void k_long_reset(unsigned long long* bar, int32_t* abort_flag, unsigned long long* lpart, int n_lpart,
                             int32_t abort_preset) {
  if (threadIdx.x < BAR_LINES) bar[BAR_STRIDE * threadIdx.x] = 0ull;
  if (threadIdx.x == 0) *abort_flag = abort_preset ? 1 : 0;
  for (int i = threadIdx.x; i < n_lpart; i += blockDim.x) lpart[i] = LSENT;
}

struct LongScratch {
  void* ws = nullptr;
  size_t bytes = 0;
};

template <int MODE, bool FIXED, bool F32, bool XL, bool LV = false>
hipError_t launch_r(LongArgs& a, int grid, size_t lds, hipStream_t s) {
  if (MODE == BLOCK) {
    hipLaunchKernelGGL((k_long<MODE, FIXED, F32, XL, LV>), dim3(grid), dim3(ct_of<MODE>()), lds, s, a);
    return hipGetLastError();
  }
  void* args[] = {&a};
  return hipLaunchCooperativeKernel((const void*)k_long<MODE, FIXED, F32, XL>, dim3(grid), dim3(ct_of<MODE>()),
                                    args, lds, s);
}

template <int MODE, bool XL = false, bool LV = false>
hipError_t launch_m(LongArgs& a, int grid, bool f32, size_t lds, hipStream_t s) {
  const bool fixed = a.fixed_m > 0;
  if (fixed)
    return f32 ? launch_r<MODE, true, true, XL, LV>(a, grid, lds, s)
               : launch_r<MODE, true, false, XL, LV>(a, grid, lds, s);
  return f32 ? launch_r<MODE, false, true, XL, LV>(a, grid, lds, s)
             : launch_r<MODE, false, false, XL, LV>(a, grid, lds, s);
}

// Workgroups of one form the device holds at once (the smallest over its
// instantiations) with lds bytes of dynamic LDS each; 0 when one of them
// cannot run with that much.
template <int MODE, bool XL = false, bool LV = false>
hipError_t resident(int n_cu, size_t lds, int* grid) {
  const void* ks[] = {(const void*)k_long<MODE, true, true, XL, LV>, (const void*)k_long<MODE, true, false, XL, LV>,
                      (const void*)k_long<MODE, false, true, XL, LV>, (const void*)k_long<MODE, false, false, XL, LV>};
  int occ = INT32_MAX;
  for (const void* k : ks) {
    int o = 0;
    hipError_t e = occupancy(k, ct_of<MODE>(), lds, &o);
    if (e != hipSuccess) return e;
    occ = std::min(occ, o);
  }
  *grid = occ * n_cu;
  return hipSuccess;
}

template <int MODE, bool XL = false>
hipError_t max_resident(int n_cu, size_t lds, int* grid) {
  hipError_t e = resident<MODE, XL>(n_cu, lds, grid);
  *grid = std::max(*grid, n_cu);
  return e;
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Region layout for vectors of lmax rows and an operator array of emax entries.
size_t region_layout(LongArgs& a, int64_t lmax, int64_t emax, int32_t kcap, bool f32, bool need_A) {
  size_t off = 0;
  a.offV = (int64_t)off;  off += al256(sizeof(double2) * (size_t)kcap * lmax);
  a.offWb = (int64_t)off; off += al256(sizeof(double2) * (size_t)lmax);
  a.offY = (int64_t)off;  off += al256(sizeof(double2) * (size_t)lmax);
  a.offS = (int64_t)off;  off += al256(sizeof(double2) * (size_t)lmax);
  a.offP = (int64_t)off;  off += al256(sizeof(double2) * (size_t)lmax);
  a.offWf = (int64_t)off; off += f32 ? al256(sizeof(float2) * (size_t)lmax) : 0;
  a.offA = (int64_t)off;  off += need_A ? al256(sizeof(double2) * (size_t)std::max<int64_t>(emax, 1)) : 0;
  return off;
}

}  // namespace

void coop_scratch_release(void* sc) {
  LongScratch* c = (LongScratch*)sc;
  if (!c) return;
  if (c->ws) (void)hipFree(c->ws);
  delete c;
}

// Long-domain wavepackets of one launch set.  blk: one workgroup each, all in
// one launch; coop: the whole device each, one cooperative launch each (its
// grid: one wave per RPW rows of the matvec, at most one workgroup per CU --
// a function of the wavepacket's domain and the device only, so its result
// does not depend on the other wavepackets).  d_ops, d_blk, d_coop: device
// copies of the jobs and of the two item lists (same order as blk / coop).
int long_run(void** scratch, const SubOp* d_ops, const int4* d_blk, const std::vector<LongItem>& blk,
             const int4* d_coop, const std::vector<LongItem>& coop, int32_t kcap, int32_t fixed_m,
             int32_t maxorder, double tol, bool f32, int n_cu, hipStream_t s, int32_t abort_preset,
             const SubOp* op0, int32_t coop_grid) {
  if (blk.empty() && coop.empty()) return RM_OK;
  int gres_r = 1, gres_m = 1, gres_b = 1, gres_x = 1;
  HIPCHK((max_resident<GRID_REG>(n_cu, 0, &gres_r)));
  HIPCHK((max_resident<GRID_MEM>(n_cu, 0, &gres_m)));
  HIPCHK((max_resident<BLOCK>(n_cu, 0, &gres_b)));
  // block items on domains of at most XL_ROWS rows come first (staged in LDS)
  size_t nx = 0;
  while (nx < blk.size() && blk[nx].item.z <= XL_ROWS) ++nx;
  const size_t lds_x = (f32 ? sizeof(float2) : sizeof(double2)) * XL_ROWS + sizeof(int) * (XL_ROWS + 1);
  if (nx) HIPCHK((max_resident<BLOCK, true>(n_cu, lds_x, &gres_x)));
  gres_r = std::min(gres_r, n_cu);
  gres_m = std::min(gres_m, n_cu);
  const bool fixed = fixed_m > 0;
  // block form: one region per workgroup
  LongArgs ab;
  memset(&ab, 0, sizeof(ab));
  int nwg = 0;
  size_t b_bytes = 0;
  if (!blk.empty()) {
    int64_t lmax = 0, emax = 0;
    bool need_A = false;
    for (const LongItem& it : blk) {
      lmax = std::max<int64_t>(lmax, it.item.z);
      emax = std::max<int64_t>(emax, it.nent);
      need_A |= it.nslot > 1;
    }
    const size_t region = region_layout(ab, lmax, emax, kcap, f32, need_A);
    nwg = (int)std::max(std::min<int64_t>((int64_t)nx, gres_x), std::min<int64_t>((int64_t)(blk.size() - nx), gres_b));
    ab.wg_bytes = (int64_t)region;
    ab.lmax = lmax;
    b_bytes = region * (size_t)nwg;
  }
  // cooperative form: one region shared by the grid
  LongArgs ac;
  memset(&ac, 0, sizeof(ac));
  size_t c_bytes = 0;
  if (!coop.empty()) {
    int64_t lmax = 0, emax = 0;
    for (const LongItem& it : coop) {
      lmax = std::max<int64_t>(lmax, it.item.z);
      emax = std::max<int64_t>(emax, it.nent);
    }
    c_bytes = region_layout(ac, lmax, emax, kcap, f32, true);
    ac.lmax = lmax;
  }
  const int gmax = std::max(std::max(gres_r, gres_m), nwg);
  const size_t b_part = al256(sizeof(double) * 2 * 4 * (size_t)gmax);
  const size_t b_lpart = al256(sizeof(double) * LRING * 4 * (size_t)gmax);
  const size_t b_gram = fixed ? 0 : al256(sizeof(double2) * (size_t)gmax * kcap * kcap);
  const size_t b_c = fixed ? 0 : al256(sizeof(double2) * (size_t)gmax * 2 * kcap);
  const size_t b_bar = sizeof(unsigned long long) * BAR_STRIDE * BAR_LINES;
  const size_t need = al256(b_bytes) + al256(c_bytes) + b_part + b_lpart + b_gram + b_c + b_bar + 256;
  LongScratch* sc = (LongScratch*)*scratch;
  if (!sc) {
    sc = new (std::nothrow) LongScratch();
    if (!sc) return fail(RM_ERR_ALLOC, "long-domain propagation: host allocation");
    *scratch = sc;
  }
  if (sc->bytes < need) {
    if (sc->ws) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(sc->ws));
      sc->ws = nullptr;
      sc->bytes = 0;
    }
    HIPCHK(hipMalloc(&sc->ws, need));
    sc->bytes = need;
  }
  char* w = (char*)sc->ws;
  char* wb = w;            w += al256(b_bytes);
  char* wc = w;            w += al256(c_bytes);
  double* part = (double*)w; w += b_part;
  double* lpart = (double*)w; w += b_lpart;
  double2* gram = fixed ? nullptr : (double2*)w; w += b_gram;
  double2* cscr = fixed ? nullptr : (double2*)w; w += b_c;
  unsigned long long* bar = (unsigned long long*)w;
  int32_t* abort_flag = (int32_t*)(w + b_bar);
  for (LongArgs* a : {&ab, &ac}) {
    a->ops = op0 ? nullptr : d_ops;
    if (op0) a->op0 = *op0;
    a->part = part;
    a->lpart = lpart;
    a->gram = gram;
    a->cscr = cscr;
    a->bar = bar;
    a->abort_flag = abort_flag;
    a->kcap = kcap;
    a->fixed_m = fixed_m;
    a->maxorder = maxorder;
    a->tol = tol;
  }
  if (nx) {
    ab.ws = wb;
    ab.items = d_blk;
    ab.nitems = (int32_t)nx;
    // the vectors of the longest of these domains in LDS beside the stage,
    // when the kernel still fits a CU with them (fixed Krylov dimensions and
    // domains of up to ~500 rows); the region layout's vector part ends at offA
    int64_t lx = 0;
    for (size_t i = 0; i < nx; ++i) lx = std::max<int64_t>(lx, blk[i].item.z);
    LongArgs lv = ab;
    const size_t vbytes = region_layout(lv, lx, 1, kcap, f32, false);
    const size_t lds_v = (f32 ? xl_bytes<true>() : xl_bytes<false>()) + vbytes;
    int gres_v = 0;
    if (lds_v <= 160 * 1024) HIPCHK((resident<BLOCK, true, true>(n_cu, lds_v, &gres_v)));
    if (gres_v > 0) {
      // same offsets as the global region's vectors (lmax differs: a layout of their own)
      lv.ws = wb;
      lv.offA = ab.offA;
      lv.wg_bytes = ab.wg_bytes;
      lv.lmax = lx;
      HIPCHK((launch_m<BLOCK, true, true>(lv, (int)std::min<int64_t>((int64_t)nx, gres_v), f32, lds_v, s)));
    } else {
      HIPCHK((launch_m<BLOCK, true>(ab, (int)std::min<int64_t>((int64_t)nx, gres_x), f32, lds_x, s)));
    }
  }
  if (blk.size() > nx) {
    ab.ws = wb;
    ab.items = d_blk + nx;
    ab.nitems = (int32_t)(blk.size() - nx);
    HIPCHK((launch_m<BLOCK>(ab, (int)std::min<int64_t>((int64_t)ab.nitems, gres_b), f32, 0, s)));
  }
  ac.ws = wc;
  ac.wg_bytes = 0;
  for (size_t t = 0; t < coop.size(); ++t) {
    const LongItem& it = coop[t];
    ac.items = d_coop + t;
    ac.nitems = 1;
    const int64_t groups = (it.item.z + RPW - 1) / RPW;
    // registers hold a lane's entries when every wave has one row group and
    // no row of the domain is longer than LPR * UR entries
    // (coop_grid > 0 caps the grid; the memory form loops over the row groups)
    constexpr int wr = ct_of<GRID_REG>() / 64, wm = ct_of<GRID_MEM>() / 64;
    const int64_t cap_r = coop_grid > 0 ? std::min<int64_t>(gres_r, coop_grid) : gres_r;
    const int64_t cap_m = coop_grid > 0 ? std::min<int64_t>(gres_m, coop_grid) : gres_m;
    const bool reg = groups <= cap_r * wr && it.maxrow <= LPR * UR;
    const int wpg = reg ? wr : wm;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((groups + wpg - 1) / wpg, reg ? cap_r : cap_m));
    hipLaunchKernelGGL(k_long_reset, dim3(1), dim3(256), 0, s, bar, abort_flag, (unsigned long long*)lpart,
                       LRING * 4 * gmax, abort_preset);
    HIPCHK(hipGetLastError());
    if (reg)
      HIPCHK((launch_m<GRID_REG>(ac, grid, f32, 0, s)));
    else
      HIPCHK((launch_m<GRID_MEM>(ac, grid, f32, 0, s)));
  }
  return RM_OK;
}
