// richmol_amd: invariant-subspace propagation.
//
// The sparsity graph of a plan's operator (union over every term and field
// component the plan holds; H0 is diagonal) splits the basis into connected
// components.  Each component is invariant under every H(F) the plan can
// form, so the Krylov vectors of a wavepacket never leave the union of the
// components its initial support touches -- its "domain".  Outside the domain
// the reference's vectors stay exactly zero (sums of products with zeros), and
// after an exact breakdown the restart vector only enters with a coefficient
// that is exactly zero, so the restricted computation gives the full-vector
// result up to the order of floating-point sums.
//
// Rows are permuted so every component is contiguous; a domain is then the
// row range [lo, hi) from its first to its last component.  One wave owns one
// wavepacket: it keeps the wavepacket, its split-operator phases and (up to
// two field slots) its operator rows in registers through every step of the
// run, assembles A(F_step) = sum_s F_s B_s per step, and runs the Lanczos
// expv (tdse.py:417-486) with wave-level reductions only (no workgroup
// barriers).  The gathered vector of the matvec lives in LDS, and so does the
// Krylov basis when it fits; otherwise the basis is in global memory.

#include <algorithm>
#include <climits>
#include <mutex>
#include <numeric>

#include "rm_common.h"

namespace {

constexpr int NSR = 2;                         // field slots held in registers
constexpr size_t SUB_LDS_BASIS = 32 * 1024;    // LDS budget of a wave's Krylov basis

struct SubArgs {
  const SubOp* ops;        // the jobs, or NULL: the one job is op0
  SubOp op0;               // the job of a single-job launch, by value (no staging)
  const int4* items;       // (column, lo, len, op), one wavepacket each
  int32_t nitems;
  double2* basis;          // per wave kcap * 64R (global-basis variant)
  double2* gram;           // per wave kcap * kcap (adaptive)
  double2* cscr;           // per wave 2 * kcap (adaptive)
  int32_t kcap, fixed_m, maxorder;
  double tol;
  int32_t max_waves;       // > 0: at most this many waves (the launch's workspace holds no more)
  int32_t obs;             // 1: some job reads an observable (k_sub_prop: LDS for the state)
};

// One wave per wavepacket, R rows per lane (domain <= 64 R rows).  WM > 0:
// the ELL entries of the lane's rows (width <= WM) are held in registers
// (their B_s too when nslot <= NSR); WM == 0: they are read (L2) at every
// matvec.  FIXED: fixed Krylov dimension.  F32: FP32 matvec (operator and
// gathered vector rounded, FP32 accumulation), everything else FP64.  BL: the
// Krylov basis in LDS (kcap * 64R * 16 B <= SUB_LDS_BASIS).
template <int R, int WM, bool FIXED, bool F32, bool BL>
__global__ __launch_bounds__(64) void k_sub_prop(SubArgs a) {
  using XT = typename XType<F32>::T;
  constexpr int NR = 64 * R;
  constexpr int WMR = WM > 0 ? WM : 1;
  constexpr int UJ = R >= 8 ? 1 : 8 / R;   // entries per row loaded together (WM == 0)
  extern __shared__ char s_raw[];
  XT* s_x = (XT*)s_raw;
  __shared__ double2 s_alpha[KMAX];
  __shared__ double s_beta[KMAX + 1];
  __shared__ double2 s_cf[KMAX];
  const int l = threadIdx.x;
  const int gw = blockIdx.x;
  const int nw = gridDim.x;
  const int kc = a.kcap;
  double2* V;
  if constexpr (BL)
    V = (double2*)(s_raw + NR * sizeof(XT));
  else
    V = a.basis + (int64_t)gw * kc * NR;
  double2* G = a.gram + (int64_t)gw * kc * kc;
  double2* cv = a.cscr + (int64_t)gw * 2 * kc;
  const double2 zero = make_double2(0.0, 0.0);

  for (int itm = gw; itm < a.nitems; itm += nw) {
    const int4 item = a.items[itm];
    const int b = item.x, lo = item.y, len = item.z;
    const SubOp op = a.ops ? a.ops[item.w] : a.op0;
    if (op.stale && *op.stale == op.call_id) {
      // reused domains that failed their check: the wavepacket is left as it is
      if (l == 0 && op.status) op.status[b - op.kcol0] |= RM_STALE_DOMAIN;
      continue;
    }
    const int n = op.n, W = op.w;
    const bool breg = WM > 0 && op.nslot <= NSR;
    // domain within lanes 0..31: the upper half of every sum is zero
    const bool half = R == 1 && len <= 32;
    auto rsum = [half](double v) __attribute__((always_inline)) {
      return half ? wave_allsum32(v) : wave_allsum(v);
    };
    int gq[R], orig[R];
    bool ok[R];
    double2 st[R], ph[R];   // the wavepacket and its phases on my rows, across steps
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = l + 64 * r;
      ok[r] = i < len;
      gq[r] = lo + min(i, len - 1);
      orig[r] = op.perm[gq[r]];
      // (rvalue alternatives throughout: `c ? mem : lvalue` would select pointers into scratch)
      st[r] = ok[r] ? op.psi[(int64_t)b * op.ld + orig[r]] : make_double2(0.0, 0.0);
      ph[r] = op.phase ? op.phase[orig[r]] : make_double2(1.0, 0.0);
    }
    // <psi|O|psi> after step `step` (the state in st): the state staged in
    // LDS past the matvec input and the basis, the observable's entries read
    // from memory (a uniform branch: only launches whose job has one)
    auto observe = [&](int step) __attribute__((always_inline)) {
      if (!(a.obs && op.ow > 0)) return;
      double2* s_o = (double2*)(s_raw + NR * sizeof(XT) + (BL ? (size_t)kc * NR * sizeof(double2) : 0));
#pragma unroll
      for (int r = 0; r < R; ++r) s_o[l + 64 * r] = st[r];
      wave_sync();
      double e0 = 0.0, e1 = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double2 y = zero;
        for (int j = 0; j < op.ow; ++j) {
          const int64_t q = (int64_t)j * n + gq[r];
          const int c = op.ocol[q] - lo;
          if (ok[r] && c >= 0 && c < len) y = cfma(op.oval[q], s_o[c], y);
        }
        e0 += st[r].x * y.x + st[r].y * y.y;
        e1 += st[r].x * y.y - st[r].y * y.x;
      }
      e0 = rsum(e0);
      e1 = rsum(e1);
      if (l == 0 && op.expect) op.expect[(int64_t)step * op.expect_ld + b - op.kcol0] = make_double2(e0, e1);
      wave_sync();   // s_o is rewritten at the next step only after these reads
    };
    int cidx[WMR][R];
    double2 Bv[NSR][WMR][R];
#pragma unroll
    for (int j = 0; j < WMR; ++j)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cidx[j][r] = (WM > 0 && j < W) ? op.col[(int64_t)j * n + gq[r]] - lo : gq[r] - lo;
#pragma unroll
        for (int s = 0; s < NSR; ++s)
          Bv[s][j][r] = (breg && j < W && s < op.nslot) ? op.val[((int64_t)s * W + j) * n + gq[r]]
                                                       : make_double2(0.0, 0.0);
      }

    for (int step = 0; step < op.nsteps; ++step) {
      double2 vk[R];
#pragma unroll
      for (int r = 0; r < R; ++r) vk[r] = op.phase ? cmul(st[r], ph[r]) : st[r];
      if (!step_on(op, step)) {
        // no field: the expv is skipped (tdse.py:377), phases only
#pragma unroll
        for (int r = 0; r < R; ++r) st[r] = op.phase ? cmul(vk[r], ph[r]) : vk[r];
        if (l == 0 && op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = 0;
        observe(step);
        continue;
      }
      const double2* fp = op.table + (int64_t)step * op.nslot;
      // operator entries of my rows for this step
      XT A[WMR][R];
      if (WM > 0) {
        double2 f0 = zero, f1 = zero;
        if (breg) {
          f0 = fp[0];
          if (op.nslot > 1) f1 = fp[1];
        }
        // slots j >= W hold B = 0 (column = own row): no branch on W below
#pragma unroll
        for (int j = 0; j < WMR; ++j)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            double2 v;
            if (breg) {
              v = cmul(f0, Bv[0][j][r]);
              if (NSR > 1) v = cfma(f1, Bv[1][j][r], v);
            } else {
              v = make_double2(0.0, 0.0);
              if (j < W)
                for (int s = 0; s < op.nslot; ++s)
                  v = cfma(fp[s], op.val[((int64_t)s * W + j) * n + gq[r]], v);
            }
            A[j][r] = to_x<F32>(v);
          }
      }

      // y = A x for x in LDS (rows past the domain computed on a clamped row, then zeroed)
      auto matvec = [&](double2 (&Y)[R]) __attribute__((always_inline)) {
        const double2 z0 = make_double2(0.0, 0.0);   // a local: a captured constant would live in scratch
        XT acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = to_x<F32>(z0);
        if (WM > 0) {
#pragma unroll
          for (int j = 0; j < WMR; ++j)
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = xfma(A[j][r], s_x[cidx[j][r]], acc[r]);
        } else {
          bool done = false;
          if constexpr (R <= 4) {
            if (op.nslot <= NSR) {
              // UJ entries per row at a time, all their loads issued before the
              // first is used (same products, same order as the loop below)
              for (int j0 = 0; j0 < W; j0 += UJ) {
                int c[UJ][R];
                XT v[UJ][R];
#pragma unroll
                for (int u = 0; u < UJ; ++u)
#pragma unroll
                  for (int r = 0; r < R; ++r) {
                    const int64_t q = (int64_t)(j0 + u < W ? j0 + u : 0) * n + gq[r];
                    c[u][r] = op.col[q] - lo;
                    double2 e = cfma(fp[0], op.val[q], z0);
                    if (NSR > 1 && op.nslot > 1) e = cfma(fp[1], op.val[(int64_t)W * n + q], e);
                    v[u][r] = to_x<F32>(e);
                  }
#pragma unroll
                for (int u = 0; u < UJ; ++u)
                  if (j0 + u < W)
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r] = xfma(v[u][r], s_x[c[u][r]], acc[r]);
              }
              done = true;
            }
          }
          if (!done)
            for (int j = 0; j < W; ++j) {
#pragma unroll
              for (int r = 0; r < R; ++r) {
                const int64_t q = (int64_t)j * n + gq[r];
                const int c = op.col[q] - lo;
                double2 v = z0;
                for (int s = 0; s < op.nslot; ++s) v = cfma(fp[s], op.val[(int64_t)s * W * n + q], v);
                acc[r] = xfma(to_x<F32>(v), s_x[c], acc[r]);
              }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) Y[r] = ok[r] ? to_d2(acc[r]) : z0;
      };

      // V_0 = phi (unnormalised, tdse.py:443-446).  Rows past the domain hold
      // zeros and are written unguarded (no valid row gathers them).
#pragma unroll
      for (int r = 0; r < R; ++r) {
        V[l + 64 * r] = vk[r];
        s_x[l + 64 * r] = to_x<F32>(vk[r]);
      }
      wave_sync();
      double2 Wr[R], vp[R];
      matvec(Wr);
      double acc3[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 w = Wr[r], v = vk[r];
        acc3[0] += w.x * v.x + w.y * v.y;
        acc3[1] += w.x * v.y - w.y * v.x;
        acc3[2] += v.x * v.x + v.y * v.y;
        vp[r] = zero;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) acc3[q] = rsum(acc3[q]);
      double2 al = make_double2(acc3[0], acc3[1]);
#pragma unroll
      for (int r = 0; r < R; ++r) Wr[r] = csub(Wr[r], cmul(al, vk[r]));
      s_alpha[0] = al;   // uniform value: every lane stores it (no branch)
      if (!FIXED && l == 0) {
        G[0] = make_double2(acc3[2], 0.0);
        cv[kc] = make_double2(1.0, 0.0);   // c_prev = [1]
      }
      int k = 1;
      double conv = 1.0;
      for (int it = 1;; ++it) {
        const bool live = FIXED ? it < a.fixed_m : (it < a.maxorder && conv > a.tol);
        if (!live) break;
        // beta_it = |W| (tdse.py:453)
        double bs = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) bs += Wr[r].x * Wr[r].x + Wr[r].y * Wr[r].y;
        const double beta = sqrt(rsum(bs));
        const bool brk = !(beta != 0.0);
        if (!brk) {
          // 1/beta: hardware reciprocal and two Newton steps (within an ulp of the quotient)
          double ib = __builtin_amdgcn_rcp(beta);
          ib = fma(fma(-beta, ib, 1.0), ib, ib);
          ib = fma(fma(-beta, ib, 1.0), ib, ib);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            vp[r] = vk[r];
            vk[r] = cscale(ib, Wr[r]);
          }
        } else {
          // exact breakdown: Gram-Schmidt of ones against the basis (tdse.py:459-465)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            vp[r] = vk[r];
            Wr[r] = ok[r] ? make_double2(1.0, 0.0) : zero;
          }
          for (int j = 0; j < it; ++j) {
            const double2* Vj = V + (int64_t)j * NR;
            double d0 = 0.0, d1 = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (ok[r]) {
                const double2 x = Vj[l + 64 * r];
                d0 += x.x * Wr[r].x + x.y * Wr[r].y;
                d1 += x.x * Wr[r].y - x.y * Wr[r].x;
              }
            const double2 proj = make_double2(rsum(d0), rsum(d1));
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (ok[r]) Wr[r] = csub(Wr[r], cmul(proj, Vj[l + 64 * r]));
          }
          double nv = 0.0;
#pragma unroll
          for (int r = 0; r < R; ++r) nv += Wr[r].x * Wr[r].x + Wr[r].y * Wr[r].y;
          nv = sqrt(rsum(nv));
          // nv == 0: the basis already spans the domain, so the reference's
          // restart vector lies entirely outside it.  Its part on the domain
          // is zero; it enters the result with coefficient exactly zero
          // (T is block diagonal after beta == 0), so zero stands for it.
#pragma unroll
          for (int r = 0; r < R; ++r)
            vk[r] = nv > 0.0 ? make_double2(Wr[r].x / nv, Wr[r].y / nv) : zero;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          V[(int64_t)it * NR + l + 64 * r] = vk[r];
          s_x[l + 64 * r] = to_x<F32>(vk[r]);
        }
        s_beta[it] = beta;
        wave_sync();
        if (!FIXED && brk) {
          // Gram column of the restart vector, explicit dots
          for (int j = 0; j < it; ++j) {
            const double2* Vj = V + (int64_t)j * NR;
            double d0 = 0.0, d1 = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
              if (ok[r]) {
                const double2 x = Vj[l + 64 * r];
                d0 += x.x * vk[r].x + x.y * vk[r].y;
                d1 += x.x * vk[r].y - x.y * vk[r].x;
              }
            d0 = rsum(d0);
            d1 = rsum(d1);
            if (l == 0) {
              G[j * kc + it] = make_double2(d0, d1);
              G[it * kc + j] = make_double2(d0, -d1);
            }
          }
        }
        // w = H V_it; alpha = vdot(w, V_it); W = w - alpha V_it - beta V_{it-1}
        matvec(Wr);
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          a0 += Wr[r].x * vk[r].x + Wr[r].y * vk[r].y;
          a1 += Wr[r].x * vk[r].y - Wr[r].y * vk[r].x;
        }
        al = make_double2(rsum(a0), rsum(a1));
#pragma unroll
        for (int r = 0; r < R; ++r) Wr[r] = csub(csub(Wr[r], cmul(al, vk[r])), cscale(beta, vp[r]));
        s_alpha[it] = al;
        k = it + 1;
        if (!FIXED) {
          wave_sync();
          conv = adaptive_metric<true>(it, brk, G, kc, s_alpha, s_beta, op.f, cv, cv + kc);
        }
      }

      // coefficients of the final combination
      wave_sync();
      const double2* cf = s_cf;
      if (FIXED) {
        if (k > 1) expm_e0_any(k, s_alpha, s_beta, op.f, s_cf);
        else if (l == 0) s_cf[0] = make_double2(1.0, 0.0);
      } else if (k == 1) {
        if (l == 0) s_cf[0] = make_double2(1.0, 0.0);
      } else {
        cf = cv;   // exp(f T) e0 of the last metric evaluation
      }
      __threadfence_block();
      wave_sync();
      double2 u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = zero;
      for (int j = 0; j < k - 1; ++j) {
        const double2 cj = cf[j];
        const double2* Vj = V + (int64_t)j * NR;
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = cfma(cj, Vj[min(l + 64 * r, len - 1)], u[r]);
      }
      const double2 clast = cf[k - 1];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 uu = cfma(clast, vk[r], u[r]);
        st[r] = ok[r] ? (op.phase ? cmul(uu, ph[r]) : uu) : zero;
      }
      if (l == 0) {
        if (op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = k;
        if (!FIXED && k >= a.maxorder && op.status) op.status[b - op.kcol0] = RM_NOT_CONVERGED;
      }
      observe(step);
      // the next step rewrites s_cf / s_x / the basis only after this wave's reads above
      wave_sync();
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (ok[r]) op.psi[(int64_t)b * op.ld + orig[r]] = st[r];
  }
}

// Two wavepackets per wave (fixed Krylov dimension <= 16, domains <= 32
// rows, the same operator): lanes 32 h + i hold row i of the pair's
// wavepacket h.  Per half, every operation is the one k_sub_prop<1, WM, true,
// F32, true> performs for a domain of <= 32 rows (its lane-local arithmetic,
// the half-wave sums of wave_allsum32 and the row-local exp(fT)e0), so a
// wavepacket's result does not depend on its partner, bit for bit, and
// equals the single-wave kernel's.  Item 2p + h of the launch is half h of
// pair p; an item with column -1 is an empty half.
// BAND: every entry of the domain's rows couples row i to i - 1, i or i + 1
// (a chain: the linear rotor's (m, J-parity) subspaces), so the matvec takes
// its neighbours from the adjacent lanes by wavefront DPP shifts -- no LDS
// gather and no wave barrier on the iteration's critical path; the entries
// are summed per offset (-1, 0, +1).
// Lane l <- lane l + 1 (up, wave_shl:1) or l - 1 (wave_shr:1); the lane with
// no source (63 up, 0 down) receives 0 (bound control with a zero old value:
// never a leftover register, which a zero coefficient could turn into NaN).
__device__ __forceinline__ int wave_shift_i(int v, bool up) {
  return up ? __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true)
            : __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ double wave_shift_d(double v, bool up) {
  return __hiloint2double(wave_shift_i(__double2hiint(v), up), wave_shift_i(__double2loint(v), up));
}
template <typename XT>
__device__ __forceinline__ XT wave_shift(XT v, bool up);
template <>
__device__ __forceinline__ double2 wave_shift<double2>(double2 v, bool up) {
  return make_double2(wave_shift_d(v.x, up), wave_shift_d(v.y, up));
}
template <>
__device__ __forceinline__ float2 wave_shift<float2>(float2 v, bool up) {
  return make_float2(__int_as_float(wave_shift_i(__float_as_int(v.x), up)),
                     __int_as_float(wave_shift_i(__float_as_int(v.y), up)));
}

// no lane of half h (lanes 32 h .. 32 h + 31) has p (all 64 lanes active)
__device__ __forceinline__ bool half_none(bool p, int h) {
  const uint64_t m = __ballot(p);
  return ((h ? (m >> 32) : m) & 0xffffffffull) == 0;
}

// v of lane src (ds_bpermute on the two halves of the double, no memory)
__device__ __forceinline__ double lane_read(double v, int src) {
  const int addr = src << 2;
  return __hiloint2double(__builtin_amdgcn_ds_bpermute(addr, __double2hiint(v)),
                          __builtin_amdgcn_ds_bpermute(addr, __double2loint(v)));
}

template <int WM, bool F32, bool BAND = false, bool OBS = false>
__global__ __launch_bounds__(64) void k_sub_pair(SubArgs a) {
  using XT = typename XType<F32>::T;
  constexpr int WMR = WM > 0 ? WM : 1;
  __shared__ XT s_x[64];
  __shared__ double2 s_V[16 * 64];
  __shared__ double2 s_alpha[2][16];
  __shared__ double s_beta[2][17];
  __shared__ double2 s_cf[2][16];
  constexpr int OBK = 16;                  // observable terms buffered per flush (OBS)
  __shared__ double2 s_ob[OBS ? OBK : 1][64];
  const int l = threadIdx.x;
  const int h = l >> 5, i = l & 31;
  const double2 zero = make_double2(0.0, 0.0);
  const int npair = (a.nitems + 1) >> 1;
  for (int pr = blockIdx.x; pr < npair; pr += gridDim.x) {
    const int itm = 2 * pr + h;
    const int4 item = itm < a.nitems ? a.items[itm] : make_int4(-1, 0, 0, a.items[2 * pr].w);
    const int b = item.x, lo = item.y, len = b >= 0 ? item.z : 0;
    const SubOp op = a.ops ? a.ops[item.w] : a.op0;   // the same for both halves
    if (op.stale && *op.stale == op.call_id) {
      // reused domains that failed their check: both wavepackets are left as they are
      if (i == 0 && b >= 0 && op.status) op.status[b - op.kcol0] |= RM_STALE_DOMAIN;
      continue;
    }
    const int n = op.n, W = op.w;
    const bool breg = WM > 0 && op.nslot <= NSR;
    const bool ok = i < len;
    const int gq = len > 0 ? lo + min(i, len - 1) : 0;
    const int orig = op.perm[gq];
    double2 st = ok ? op.psi[(int64_t)b * op.ld + orig] : make_double2(0.0, 0.0);
    const double2 ph = op.phase ? op.phase[orig] : make_double2(1.0, 0.0);
    int cidx[WMR];
    double2 Bv[NSR][WMR];
#pragma unroll
    for (int j = 0; j < WMR; ++j) {
      cidx[j] = ((WM > 0 && j < W) ? op.col[(int64_t)j * n + gq] - lo : gq - lo) + 32 * h;
      if (len == 0) cidx[j] = l;
#pragma unroll
      for (int sl = 0; sl < NSR; ++sl)
        Bv[sl][j] = (breg && j < W && sl < op.nslot) ? op.val[((int64_t)sl * W + j) * n + gq]
                                                    : make_double2(0.0, 0.0);
    }
    // BAND: the entries per offset -1, 0, +1 (rows past the domain: none)
    double2 Bo[NSR][3];
#pragma unroll
    for (int sl = 0; sl < NSR; ++sl)
#pragma unroll
      for (int o = 0; o < 3; ++o) Bo[sl][o] = make_double2(0.0, 0.0);
    if (BAND && ok) {
#pragma unroll
      for (int j = 0; j < WMR; ++j) {
        const int o = cidx[j] - l + 1;   // 0, 1, 2 for columns i - 1, i, i + 1
#pragma unroll
        for (int sl = 0; sl < NSR; ++sl)
#pragma unroll
          for (int oo = 0; oo < 3; ++oo)
            if (o == oo) Bo[sl][oo] = cadd(Bo[sl][oo], Bv[sl][j]);
      }
    }
    // BAND: is every field slot's operator Hermitian on the domain (diagonal
    // real, the entry to row i + 1 the conjugate of row i + 1's entry to row
    // i, to 1e-14 relative)?  Then with real field products alpha = <v|A v>
    // is real up to rounding, and its imaginary sum is not formed.
    bool herm = false;
    if constexpr (BAND) {
      bool bad = false;
#pragma unroll
      for (int sl = 0; sl < NSR; ++sl) {
        const double2 lower_next = wave_shift<double2>(Bo[sl][0], true);   // row i+1's entry to row i
        const double2 up = Bo[sl][2];
        const double scale = fmax(fmax(fabs(up.x), fabs(up.y)), fmax(fabs(lower_next.x), fabs(lower_next.y)));
        const bool pair_ok = !ok || i + 1 >= len ||
                             (fabs(up.x - lower_next.x) <= 1e-14 * scale && fabs(up.y + lower_next.y) <= 1e-14 * scale);
        const bool diag_ok = !ok || fabs(Bo[sl][1].y) <= 1e-14 * fabs(Bo[sl][1].x);
        bad |= !(pair_ok && diag_ok);
      }
      herm = half_none(bad, h);   // per half, as the operator may be Hermitian on some domains only
    }
    // BAND: are my half's entries real (every slot and offset, imaginary
    // parts within 1e-14 of the real ones, as tensors assembled in complex
    // arithmetic leave them)?  Then with real field products the matvec
    // multiplies by the real parts alone (decided per half, so a wavepacket's
    // arithmetic never depends on the one it is paired with).
    bool breal = false;
    if constexpr (BAND) {
      bool im = false;
#pragma unroll
      for (int sl = 0; sl < NSR; ++sl)
#pragma unroll
        for (int o = 0; o < 3; ++o) im |= fabs(Bo[sl][o].y) > 1e-14 * fabs(Bo[sl][o].x);
      breal = half_none(im, h);
    }
    // OBS: the observable's entries of my row as (source lane, value), and
    // summed per offset -1, 0, +1 when every entry inside the domain lies
    // within one row of the diagonal (then the rows come by DPP shifts, as
    // the banded matvec takes them); a column outside the domain (where the
    // wavepacket is zero) or a row past it gives none.  Entries beyond OWR
    // are read from memory at every step.  Each lane's term conj(psi_i)
    // (O psi)_i goes to LDS; every OBK steps (and after the last) lane i of
    // each half adds the terms of buffered step i over its rows in order.
    constexpr int OWR = 4;
    int osrc[OBS ? OWR : 1];
    double2 oval[OBS ? OWR : 1];
    double2 o3[3] = {zero, zero, zero};
    bool oband = false;
    if constexpr (OBS) {
      bool far = false;
#pragma unroll
      for (int j = 0; j < OWR; ++j) {
        const bool e = ok && j < op.ow;
        const int c = e ? op.ocol[(int64_t)j * n + gq] - lo : -1;
        const bool in = c >= 0 && c < len;
        const double2 v = load_or_zero(in, op.oval + (int64_t)j * n + gq);
        osrc[j] = in ? 32 * h + c : l;
        oval[j] = v;
        const int d = c - i;
#pragma unroll
        for (int o = 0; o < 3; ++o)
          if (in && d == o - 1) o3[o] = cadd(o3[o], v);
        far |= in && (d < -1 || d > 1);
      }
      for (int j = OWR; j < op.ow; ++j) {
        const int c = ok ? op.ocol[(int64_t)j * n + gq] - lo : -1;
        const bool in = c >= 0 && c < len;
        const double2 v = load_or_zero(in, op.oval + (int64_t)j * n + gq);
        const int d = c - i;
#pragma unroll
        for (int o = 0; o < 3; ++o)
          if (in && d == o - 1) o3[o] = cadd(o3[o], v);
        far |= in && (d < -1 || d > 1);
      }
      oband = !__any(far);
    }
    // <psi|O|psi> of the pair's wavepackets after step `step` (the state in st)
    auto observe = [&](int step) __attribute__((always_inline)) {
      if constexpr (OBS) {
        double2 y = zero;
        if (oband) {
          y = cmul(o3[1], st);
          y = cfma(o3[0], wave_shift<double2>(st, false), y);
          y = cfma(o3[2], wave_shift<double2>(st, true), y);
        } else {
#pragma unroll
          for (int j = 0; j < OWR; ++j)
            y = cfma(oval[j], make_double2(lane_read(st.x, osrc[j]), lane_read(st.y, osrc[j])), y);
          for (int j = OWR; j < op.ow; ++j) {
            const int c = ok ? op.ocol[(int64_t)j * n + gq] - lo : -1;
            const bool in = c >= 0 && c < len;
            const double2 v = load_or_zero(in, op.oval + (int64_t)j * n + gq);
            const int src = in ? 32 * h + c : l;
            y = cfma(v, make_double2(lane_read(st.x, src), lane_read(st.y, src)), y);
          }
        }
        s_ob[step % OBK][l] = make_double2(st.x * y.x + st.y * y.y, st.x * y.y - st.y * y.x);
        if ((step + 1) % OBK == 0 || step + 1 == op.nsteps) {
          wave_sync();
          const int nk = step % OBK + 1, s0 = step + 1 - nk;
          if (i < nk && b >= 0 && op.expect) {
            double2 acc = zero;
            for (int r = 0; r < len; ++r) acc = cadd(acc, s_ob[i][32 * h + r]);
            op.expect[(int64_t)(s0 + i) * op.expect_ld + b - op.kcol0] = acc;
          }
          wave_sync();   // the buffer is rewritten after these reads
        }
      }
    };
    double2* sa = s_alpha[h];
    double* sb = s_beta[h];
    double2* sc = s_cf[h];
    // the step flag and the first two field products of the next step are
    // loaded while the current step computes (their latency is off the chain)
    int sf_next = op.nsteps > 0 ? op.sflag[0] : 0;
    double2 f0_next = load_or_zero(op.nsteps > 0 && op.nslot > 0, op.table);
    double2 f1_next = load_or_zero(op.nsteps > 0 && op.nslot > 1, op.table + 1);
    for (int step = 0; step < op.nsteps; ++step) {
      const int sf = sf_next;
      const double2 f0 = f0_next, f1 = f1_next;
      if (step + 1 < op.nsteps) {
        const double2* q = op.table + (int64_t)(step + 1) * op.nslot;
        sf_next = op.sflag[step + 1];
        if (op.nslot > 0) f0_next = q[0];
        if (op.nslot > 1) f1_next = q[1];
      }
      double2 vk = op.phase ? cmul(st, ph) : st;
      if (sf == 0) {
        st = op.phase ? cmul(vk, ph) : vk;
        if (i == 0 && b >= 0 && op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = 0;
        observe(step);
        continue;
      }
      const double2* fp = op.table + (int64_t)step * op.nslot;
      XT A[WMR];
      XT Ab[3];
      if (BAND) {
#pragma unroll
        for (int o = 0; o < 3; ++o) Ab[o] = to_x<F32>(cfma(f1, Bo[1][o], cmul(f0, Bo[0][o])));
      } else if (WM > 0) {
#pragma unroll
        for (int j = 0; j < WMR; ++j) {
          double2 v;
          if (breg) {
            v = cmul(f0, Bv[0][j]);
            if (NSR > 1) v = cfma(f1, Bv[1][j], v);
          } else {
            v = make_double2(0.0, 0.0);
            if (j < W)
              for (int sl = 0; sl < op.nslot; ++sl) v = cfma(fp[sl], op.val[((int64_t)sl * W + j) * n + gq], v);
          }
          A[j] = to_x<F32>(v);
        }
      }
      const bool realA = BAND && breal && f0.y == 0.0 && f1.y == 0.0;
      auto matvec = [&]() __attribute__((always_inline)) -> double2 {
        const double2 z0 = make_double2(0.0, 0.0);
        XT acc;   // from the first product (an added zero is an instruction)
        if (BAND) {
          // rows past the domain have no entries: their result is 0 as it is
          const XT x = to_x<F32>(vk);
          const XT xm = wave_shift<XT>(x, false), xp = wave_shift<XT>(x, true);
          if (realA) {
            acc.x = Ab[0].x * xm.x;
            acc.y = Ab[0].x * xm.y;
            acc.x = Ab[1].x * x.x + acc.x;
            acc.y = Ab[1].x * x.y + acc.y;
            acc.x = Ab[2].x * xp.x + acc.x;
            acc.y = Ab[2].x * xp.y + acc.y;
          } else {
            acc = xmul(Ab[0], xm);
            acc = xfma(Ab[1], x, acc);
            acc = xfma(Ab[2], xp, acc);
          }
          return to_d2(acc);
        } else {
          acc = xmul(A[0], s_x[cidx[0]]);
#pragma unroll
          for (int j = 1; j < WMR; ++j) acc = xfma(A[j], s_x[cidx[j]], acc);
        }
        return ok ? to_d2(acc) : z0;
      };
      s_V[l] = vk;
      if (!BAND) {
        s_x[l] = to_x<F32>(vk);
        wave_sync();
      }
      // Hermitian domain operator and real field products: alpha is real,
      // and with a real-time exponent f = i phi the Krylov exponential takes
      // the real recursion (decided per half; where the halves differ both
      // forms run under their lane masks, and the exponentials' lane reads of
      // the other half only bound the unrolled number of terms)
      const bool realal = herm && __all(f0.y == 0.0 && f1.y == 0.0);
      const bool realexp = realal && __all(op.f.x == 0.0);
      double2 Wr = matvec(), vp = zero;
      double acc3[2];
      acc3[0] = Wr.x * vk.x + Wr.y * vk.y;
      acc3[1] = Wr.x * vk.y - Wr.y * vk.x;
      double2 al = make_double2(half_allsum32(acc3[0]), 0.0);
      if (realal) {
        Wr = csub(Wr, cscale(al.x, vk));
      } else {
        al.y = half_allsum32(acc3[1]);
        Wr = csub(Wr, cmul(al, vk));
      }
      sa[0] = al;
      int k = 1;
      for (int it = 1; it < a.fixed_m; ++it) {
        // beta = |W| and 1/beta from the reciprocal square root (one chain
        // instead of a square root and a quotient)
        const double ww = half_allsum32(Wr.x * Wr.x + Wr.y * Wr.y);
        double beta = 0.0, ib = 0.0;
        if (ww != 0.0) sqrt_rcp_nb(ww, &beta, &ib);
        const bool brk = !(ww != 0.0);
        if (!brk) {
          vp = vk;
          vk = cscale(ib, Wr);
        } else {
          // exact breakdown: Gram-Schmidt of ones against the basis (tdse.py:459-465)
          vp = vk;
          Wr = ok ? make_double2(1.0, 0.0) : zero;
          for (int j = 0; j < it; ++j) {
            double d0 = 0.0, d1 = 0.0;
            if (ok) {
              const double2 x = s_V[j * 64 + l];
              d0 = x.x * Wr.x + x.y * Wr.y;
              d1 = x.x * Wr.y - x.y * Wr.x;
            }
            const double2 proj = make_double2(half_allsum32(d0), half_allsum32(d1));
            if (ok) Wr = csub(Wr, cmul(proj, s_V[j * 64 + l]));
          }
          const double nv = sqrt(half_allsum32(Wr.x * Wr.x + Wr.y * Wr.y));
          vk = nv > 0.0 ? make_double2(Wr.x / nv, Wr.y / nv) : zero;
        }
        s_V[it * 64 + l] = vk;
        sb[it] = beta;
        if (!BAND) {
          s_x[l] = to_x<F32>(vk);
          wave_sync();
        }
        Wr = matvec();
        const double a0 = Wr.x * vk.x + Wr.y * vk.y, a1 = Wr.x * vk.y - Wr.y * vk.x;
        al = make_double2(half_allsum32(a0), 0.0);
        // W - beta V_{k-1} is formed while alpha is summed, then - alpha V_k
        // (one dependent operation after the sum instead of two)
        if (realal) {
          Wr = csub(csub(Wr, cscale(beta, vp)), cscale(al.x, vk));
        } else {
          al.y = half_allsum32(a1);
          Wr = csub(csub(Wr, cscale(beta, vp)), cmul(al, vk));
        }
        sa[it] = al;
        k = it + 1;
      }
      wave_sync();
      if (k > 1) {
        if (realexp) expm_e0_16_herm<true>(k, sa, sb, op.f.y, sc);
        else expm_e0_16<true>(k, sa, sb, op.f, sc);
      } else if (i == 0) {
        sc[0] = make_double2(1.0, 0.0);
      }
      __threadfence_block();
      wave_sync();
      // two interleaved partial sums (even and odd j) halve the dependent chain
      const int vrow = len > 0 ? 32 * h + min(i, len - 1) : l;
      double2 u0 = zero, u1 = zero;
      int j = 0;
      for (; j + 1 < k - 1; j += 2) {
        u0 = cfma(sc[j], s_V[j * 64 + vrow], u0);
        u1 = cfma(sc[j + 1], s_V[(j + 1) * 64 + vrow], u1);
      }
      if (j < k - 1) u0 = cfma(sc[j], s_V[j * 64 + vrow], u0);
      const double2 uu = cfma(sc[k - 1], vk, cadd(u0, u1));
      st = ok ? (op.phase ? cmul(uu, ph) : uu) : zero;
      if (i == 0 && b >= 0 && op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = k;
      observe(step);
      wave_sync();
    }
    if (ok) op.psi[(int64_t)b * op.ld + orig] = st;
  }
}

__global__ void k_sub_fill(int32_t* x, int64_t n, int32_t v0, int32_t v1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (i & 1) ? v1 : v0;
}

// Domain of every column: [min lo, max hi) over the components of its
// non-zero rows; a non-finite entry makes the domain the whole basis.
__global__ void k_col_domain(int64_t n, int64_t ld, const double2* __restrict__ psi,
                             const int32_t* __restrict__ label, const int32_t* __restrict__ lo,
                             const int32_t* __restrict__ hi, int32_t* __restrict__ dom) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (i >= n) return;
  const double2 v = psi[b * ld + i];
  if (!(isfinite(v.x) && isfinite(v.y))) {
    atomicMin(&dom[2 * b], 0);
    atomicMax(&dom[2 * b + 1], (int32_t)n);
  } else if (v.x != 0.0 || v.y != 0.0) {
    const int32_t c = label[i];
    atomicMin(&dom[2 * b], lo[c]);
    atomicMax(&dom[2 * b + 1], hi[c]);
  }
}

// sflag[st] = 1 if any MF element of step st is non-zero (or not a number):
// the `len(H.mfmat) > 0` test of tdse.py:377 without thresholds.
__global__ void k_step_flags(int64_t nterm, int64_t nsteps, int32_t nfield,
                             const int64_t* __restrict__ term_ptr, const int32_t* __restrict__ term_field,
                             const double2* __restrict__ term_mval, const double2* __restrict__ table,
                             int32_t* __restrict__ sflag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nterm) return;
  for (int64_t st = blockIdx.y; st < nsteps; st += gridDim.y) {
    const double2* fp = table + st * nfield;
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t q = term_ptr[t]; q < term_ptr[t + 1]; ++q) acc = cfma(fp[term_field[q]], term_mval[q], acc);
    if (acc.x != 0.0 || acc.y != 0.0 || acc.x != acc.x || acc.y != acc.y) sflag[st] = 1;
  }
}

// The device work of an rm_propagate call ahead of its propagation launches,
// in one launch: blocks [0, nb_flag) the step flags -- block (step, chunk)
// tests the MF elements of its chunk of terms (no block walks all terms) and
// stores its partial flag; the last block of a step to finish (a per-step
// arrival counter that only grows, so it needs no clearing) combines them
// into the step's flag; blocks [nb_flag, nb_flag + nb_fill) zero kused and
// status; the remaining blocks (domain reuse) check every entry of psi
// against its column's kept domain and store call_id into *stale on a miss.
struct Prologue {
  int64_t nterm, nsteps;
  int32_t nfield;
  const int64_t* term_ptr;
  const int32_t* term_field;
  const double2* term_mval;
  const double2* table;
  int32_t* sflag;
  int32_t* kused;
  int64_t nk;
  int32_t* status;
  int64_t ncols;
  int64_t n, ld;
  const double2* psi;
  const int32_t* label;   // original row -> subspace
  const int32_t* clo;     // subspace -> first permuted row
  const int32_t* chi;     // subspace -> one past its last
  const int32_t* dom;     // kept domains (NULL: no check)
  int32_t* stale;
  int32_t call_id;
  int32_t nb_flag, nb_fill, cpb;   // cpb: check blocks per column
  int32_t parts;                    // flag chunks per step (nb_flag = nsteps * parts)
  int64_t chunk;                    // terms per chunk
  int32_t* part;                    // [nsteps][parts] partial flags (parts > 1)
  uint32_t* arrive;                 // [nsteps] arrivals of a step's chunk blocks, multiples of parts between calls
  double2* ex;                      // observable values [nsteps][ncols] (zero columns keep the fill), or NULL
  int64_t nex;
};

__global__ __launch_bounds__(256) void k_sub_prologue(Prologue a) {
  const int tid = threadIdx.x;
  const int64_t bid = blockIdx.x;
  if (bid < a.nb_flag) {
    const int64_t st = bid / a.parts, c = bid % a.parts;
    const double2* fp = a.table + st * a.nfield;
    const int64_t t1 = std::min(a.nterm, (c + 1) * a.chunk);
    // all products of the step zero: every MF element is zero (the M values
    // are finite); otherwise the chunk is scanned until its first non-zero
    // element (a uniform exit after each pass of 256 terms)
    bool anyf = false;
    for (int f = 0; f < a.nfield; ++f) anyf |= !(fp[f].x == 0.0 && fp[f].y == 0.0);
    int any = 0;
    if (anyf)
      for (int64_t t0 = c * a.chunk; t0 < t1; t0 += 256) {
        const int64_t t = t0 + tid;
        if (t < t1) {
          double2 acc = make_double2(0.0, 0.0);
          for (int64_t q = a.term_ptr[t]; q < a.term_ptr[t + 1]; ++q)
            acc = cfma(fp[a.term_field[q]], a.term_mval[q], acc);
          any |= (acc.x != 0.0 || acc.y != 0.0 || acc.x != acc.x || acc.y != acc.y) ? 1 : 0;
        }
        if (__syncthreads_or(any)) break;
      }
    any = __syncthreads_or(any);
    if (tid == 0) {
      if (a.parts == 1) {
        a.sflag[st] = any ? 1 : 0;
      } else {
        __hip_atomic_store(a.part + bid, any ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const uint32_t old = atomicAdd(a.arrive + st, 1u);
        if ((old + 1) % (uint32_t)a.parts == 0) {
          __threadfence();
          int all = 0;
          for (int64_t c = 0; c < a.parts; ++c)
            all |= __hip_atomic_load(a.part + st * a.parts + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          a.sflag[st] = all ? 1 : 0;
        }
      }
    }
  } else if (bid < (int64_t)a.nb_flag + a.nb_fill) {
    const int64_t stride = (int64_t)a.nb_fill * 256;
    for (int64_t i = (bid - a.nb_flag) * 256 + tid; i < a.nk; i += stride) a.kused[i] = 0;
    for (int64_t i = (bid - a.nb_flag) * 256 + tid; i < a.ncols; i += stride) a.status[i] = 0;
    for (int64_t i = (bid - a.nb_flag) * 256 + tid; i < a.nex; i += stride) a.ex[i] = make_double2(0.0, 0.0);
  } else {
    const int64_t j = bid - a.nb_flag - a.nb_fill;
    const int64_t b = j / a.cpb;
    const int64_t lo = a.dom[2 * b], hi = a.dom[2 * b + 1];
    bool miss = false;
    for (int64_t i = (j % a.cpb) * 256 + tid; i < a.n; i += (int64_t)a.cpb * 256) {
      const double2 v = a.psi[b * a.ld + i];
      if (!(isfinite(v.x) && isfinite(v.y))) {
        miss |= lo > 0 || hi < a.n;
      } else if (v.x != 0.0 || v.y != 0.0) {
        const int32_t c = a.label[i];
        miss |= a.clo[c] < lo || a.chi[c] > hi;
      }
    }
    if (miss) *a.stale = a.call_id;
  }
}

inline double2 h_cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// B_s per CSR entry: sum over pairs p of K_p * sum_{q in term_p, field_q = s} M_q
std::vector<double2> entry_values(const rm_plan* p) {
  const int64_t nnz = p->nnz;
  const int32_t nf = std::max(p->nfield, 1);
  std::vector<double2> B((size_t)nf * (size_t)nnz, make_double2(0.0, 0.0));
  for (int64_t e = 0; e < nnz; ++e)
    for (int64_t q = p->h_pair_ptr[e]; q < p->h_pair_ptr[e + 1]; ++q) {
      const int32_t t = p->h_pair_term[q];
      const double2 kv = p->h_pair_kval[q];
      for (int64_t m = p->h_term_ptr[t]; m < p->h_term_ptr[t + 1]; ++m) {
        double2& d = B[(size_t)p->h_term_field[m] * nnz + e];
        const double2 pr = h_cmul(kv, p->h_term_mval[m]);
        d.x += pr.x;
        d.y += pr.y;
      }
    }
  return B;
}

// The B_s as CSR in permuted row and column order (the cooperative kernel's
// operator); rows keep the order of their entries.
int csr_build(rm_plan* p, const std::vector<double2>& B) {
  if (p->sub_rp) return RM_OK;
  const int64_t n = p->n, nnz = p->nnz;
  const int32_t nf = std::max(p->nfield, 1);
  const std::vector<int32_t>& perm = p->h_sub_perm;
  std::vector<int32_t> iperm((size_t)n);
  for (int64_t ip = 0; ip < n; ++ip) iperm[perm[ip]] = (int32_t)ip;
  std::vector<int64_t> rp((size_t)n + 1, 0);
  for (int64_t ip = 0; ip < n; ++ip) {
    const int32_t i = perm[ip];
    rp[ip + 1] = rp[ip] + (p->h_row_ptr[i + 1] - p->h_row_ptr[i]);
  }
  std::vector<int32_t> ccol((size_t)nnz);
  std::vector<double2> cval((size_t)nf * (size_t)nnz);
  for (int64_t ip = 0; ip < n; ++ip) {
    const int32_t i = perm[ip];
    const int64_t r0 = p->h_row_ptr[i], len = p->h_row_ptr[i + 1] - r0;
    for (int64_t j = 0; j < len; ++j) {
      ccol[(size_t)(rp[ip] + j)] = iperm[p->h_col[r0 + j]];
      for (int32_t s = 0; s < nf; ++s)
        cval[(size_t)s * nnz + (size_t)(rp[ip] + j)] = B[(size_t)s * nnz + (size_t)(r0 + j)];
    }
  }
  std::vector<float2> cval32(cval.size());
  for (size_t q = 0; q < cval.size(); ++q) cval32[q] = make_float2((float)cval[q].x, (float)cval[q].y);
  int rc = RM_OK;
  if ((rc = upload(&p->sub_rp, rp.data(), n + 1)) || (rc = upload(&p->sub_ccol, ccol.data(), nnz)) ||
      (rc = upload(&p->sub_cval, cval.data(), (int64_t)nf * nnz)) ||
      (rc = upload(&p->sub_cval32, cval32.data(), (int64_t)nf * nnz)))
    return rc;
  p->h_sub_rp.swap(rp);
  return RM_OK;
}

// Connected components of the sparsity graph (host union-find), the row
// permutation that makes them contiguous, and the per-slot operator B_s in
// permuted ELL order; uploaded once per plan.
int sub_build(rm_plan* p) {
  if (p->sub_ready) return RM_OK;
  const int64_t n = p->n;
  std::vector<int32_t> parent((size_t)n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (int64_t i = 0; i < n; ++i)
    for (int64_t e = p->h_row_ptr[i]; e < p->h_row_ptr[i + 1]; ++e) {
      int32_t a = find((int32_t)i), b = find(p->h_col[e]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  std::vector<int32_t> label((size_t)n), cid((size_t)n, -1);
  int32_t nc = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t r = find((int32_t)i);
    if (cid[r] < 0) cid[r] = nc++;
    label[i] = cid[r];
  }
  std::vector<int32_t> lo((size_t)nc + 1, 0), hi((size_t)nc, 0), perm((size_t)n), iperm((size_t)n);
  for (int64_t i = 0; i < n; ++i) lo[label[i] + 1]++;
  for (int32_t c = 0; c < nc; ++c) lo[c + 1] += lo[c];
  std::vector<int32_t> pos(lo.begin(), lo.end() - 1);
  for (int64_t i = 0; i < n; ++i) {
    perm[pos[label[i]]] = (int32_t)i;
    iperm[i] = pos[label[i]]++;
  }
  int32_t mx = 0;
  for (int32_t c = 0; c < nc; ++c) {
    hi[c] = lo[c + 1];
    mx = std::max(mx, hi[c] - lo[c]);
  }
  p->nsub = nc;
  p->max_sub = mx;
  const int64_t nnz = p->nnz;
  const int32_t nf = std::max(p->nfield, 1);
  const std::vector<double2> B = entry_values(p);
  int32_t mn = INT32_MAX;
  for (int32_t c = 0; c < nc; ++c) mn = std::min(mn, hi[c] - lo[c]);
  p->h_sub_start.assign(lo.begin(), lo.end() - 1);
  p->h_sub_end = hi;
  p->h_sub_label_perm.resize((size_t)n);
  for (int64_t ip = 0; ip < n; ++ip) p->h_sub_label_perm[ip] = label[perm[ip]];
  int rc = RM_OK;
  if ((rc = upload(&p->sub_perm, perm.data(), n)) || (rc = upload(&p->sub_label, label.data(), n)) ||
      (rc = upload(&p->sub_iperm, iperm.data(), n)) || (rc = upload(&p->sub_lo, lo.data(), nc)) ||
      (rc = upload(&p->sub_hi, hi.data(), nc)))
    return rc;
  // ELL (column-major) for the one-wave kernel: only if some subspace fits a wave
  const int32_t W = p->ell_w;
  if (mn <= RM_SUBSPACE_MAX_ROWS) {
    const size_t ne = (size_t)W * (size_t)n;
    std::vector<int32_t> hcol(ne);
    std::vector<double2> hval((size_t)nf * ne, make_double2(0.0, 0.0));
    for (int64_t ip = 0; ip < n; ++ip) {
      const int32_t i = perm[ip];
      const int64_t r0 = p->h_row_ptr[i], len = p->h_row_ptr[i + 1] - r0;
      for (int32_t j = 0; j < W; ++j) {
        const size_t q = (size_t)j * n + ip;
        if (j < len) {
          hcol[q] = iperm[p->h_col[r0 + j]];
          for (int32_t s = 0; s < nf; ++s) hval[(size_t)s * ne + q] = B[(size_t)s * nnz + r0 + j];
        } else {
          hcol[q] = (int32_t)ip;
        }
      }
    }
    p->h_sub_col = hcol;
    if ((rc = upload(&p->sub_col, hcol.data(), (int64_t)ne)) ||
        (rc = upload(&p->sub_val, hval.data(), (int64_t)((size_t)nf * ne))))
      return rc;
  }
  p->h_sub_perm.swap(perm);
  // CSR (row-major, permuted) for the cooperative kernel: if some subspace does not fit
  if (mx > RM_SUBSPACE_MAX_ROWS && (rc = csr_build(p, B))) return rc;
  p->sub_w = W;
  p->sub_ready = 1;
  return RM_OK;
}

// Launch classes (one kernel instantiation each): rows per lane R and where
// the operator entries live.  A wavepacket's class depends only on its own
// domain and operator width, so its arithmetic -- and its result, bit for bit
// -- does not depend on the other wavepackets of the launch.
//   0: R=1, 1: R=2, 2: R=4 with the operator in registers (ELL width <= 4)
//   3: R=4, 4: R=8, 5: R=16 reading the operator at every matvec
//   6: two wavepackets per wave (k_sub_pair): fixed Krylov dimension <= 16,
//      domains <= 32 rows, ELL width <= 4 -- the class-0 items that qualify,
//      paired with an item of the same operator or with an empty half
//   7: the same with a banded domain (every entry within one row of the
//      diagonal, <= 2 field slots): neighbours by DPP shifts (k_sub_pair BAND)
//   8: R=1, 9: R=2 reading the operator at every matvec (ELL width > 4)
constexpr int NCLASS = 10;
constexpr int PAIR_CLASS = 6;
constexpr int PAIR_BAND_CLASS = 7;
constexpr int PAIR_MAX_M = 16;
constexpr int CLASS_R[NCLASS] = {1, 2, 4, 4, 8, 16, 1, 1, 1, 2};

int sub_class(int w, int len) {
  if (w <= 4) {
    if (len <= 64) return 0;
    if (len <= 128) return 1;
    if (len <= 256) return 2;
  } else {
    if (len <= 64) return 8;
    if (len <= 128) return 9;
    if (len <= 256) return 3;
  }
  return len <= 512 ? 4 : 5;
}

// Krylov basis in LDS: register classes whose basis fits.
bool class_basis_lds(int c, int kcap) {
  if (c == PAIR_CLASS || c == PAIR_BAND_CLASS) return true;   // static LDS of k_sub_pair
  return (c <= 2 || c >= 8) && (size_t)kcap * 64 * CLASS_R[c] * sizeof(double2) <= SUB_LDS_BASIS;
}

template <bool F32, bool BAND, bool OBS>
hipError_t launch_pair(const SubArgs& a0, int n_cu, hipStream_t s, int* waves_out, bool dry) {
  auto kern = k_sub_pair<4, F32, BAND, OBS>;
  int occ = 0;
  hipError_t e = occupancy((const void*)kern, 64, 0, &occ);
  if (e != hipSuccess) return e;
  const int64_t npair = ((int64_t)a0.nitems + 1) / 2;
  int waves = (int)std::min<int64_t>(npair, (int64_t)occ * n_cu);
  if (a0.max_waves > 0) waves = std::min(waves, a0.max_waves);
  if (waves < 1) waves = 1;
  *waves_out = waves;
  if (dry) return hipSuccess;
  hipLaunchKernelGGL(kern, dim3(waves), dim3(64), 0, s, a0);
  return hipGetLastError();
}

template <int R, int WM, bool FIXED, bool F32, bool BL>
hipError_t launch_sub_k(const SubArgs& a0, int n_cu, hipStream_t s, int* waves_out, bool dry) {
  auto kern = k_sub_prop<R, WM, FIXED, F32, BL>;
  const size_t lds = (size_t)64 * R * (F32 ? sizeof(float2) : sizeof(double2)) +
                     (BL ? (size_t)a0.kcap * 64 * R * sizeof(double2) : 0) +
                     (a0.obs ? (size_t)64 * R * sizeof(double2) : 0);
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(64 * 1024));
    if (e != hipSuccess) return e;
    attr = true;
  }
  int occ = 0;
  hipError_t e = occupancy((const void*)kern, 64, lds, &occ);
  if (e != hipSuccess) return e;
  int waves = std::min<int64_t>(a0.nitems, (int64_t)occ * n_cu);
  if (a0.max_waves > 0) waves = std::min(waves, a0.max_waves);
  if (waves < 1) waves = 1;
  *waves_out = waves;
  if (dry) return hipSuccess;
  hipLaunchKernelGGL(kern, dim3(waves), dim3(64), lds, s, a0);
  return hipGetLastError();
}

template <int R, int WM, bool FIXED, bool F32>
hipError_t launch_sub_b(int c, const SubArgs& a, int n_cu, hipStream_t s, int* waves, bool dry) {
  if (class_basis_lds(c, a.kcap)) return launch_sub_k<R, WM, FIXED, F32, true>(a, n_cu, s, waves, dry);
  return launch_sub_k<R, WM, FIXED, F32, false>(a, n_cu, s, waves, dry);
}

template <bool FIXED, bool F32>
hipError_t launch_class_p(int c, const SubArgs& a, int n_cu, hipStream_t s, int* waves, bool dry) {
  switch (c) {
    case 0: return launch_sub_b<1, 4, FIXED, F32>(c, a, n_cu, s, waves, dry);
    case 1: return launch_sub_b<2, 4, FIXED, F32>(c, a, n_cu, s, waves, dry);
    case 2: return launch_sub_b<4, 4, FIXED, F32>(c, a, n_cu, s, waves, dry);
    case 3: return launch_sub_k<4, 0, FIXED, F32, false>(a, n_cu, s, waves, dry);
    case 4: return launch_sub_k<8, 0, FIXED, F32, false>(a, n_cu, s, waves, dry);
    case 8: return launch_sub_b<1, 0, FIXED, F32>(c, a, n_cu, s, waves, dry);
    case 9: return launch_sub_b<2, 0, FIXED, F32>(c, a, n_cu, s, waves, dry);
    default: return launch_sub_k<16, 0, FIXED, F32, false>(a, n_cu, s, waves, dry);
  }
}

hipError_t launch_class(int c, const SubArgs& a, bool f32, int n_cu, hipStream_t s, int* waves, bool dry) {
  if (c == PAIR_CLASS || c == PAIR_BAND_CLASS) {
    const bool band = c == PAIR_BAND_CLASS, obs = a.obs != 0;
    if (f32)
      return band ? (obs ? launch_pair<true, true, true>(a, n_cu, s, waves, dry)
                         : launch_pair<true, true, false>(a, n_cu, s, waves, dry))
                  : (obs ? launch_pair<true, false, true>(a, n_cu, s, waves, dry)
                         : launch_pair<true, false, false>(a, n_cu, s, waves, dry));
    return band ? (obs ? launch_pair<false, true, true>(a, n_cu, s, waves, dry)
                       : launch_pair<false, true, false>(a, n_cu, s, waves, dry))
                : (obs ? launch_pair<false, false, true>(a, n_cu, s, waves, dry)
                       : launch_pair<false, false, false>(a, n_cu, s, waves, dry));
  }
  const bool fixed = a.fixed_m > 0;
  if (fixed && f32) return launch_class_p<true, true>(c, a, n_cu, s, waves, dry);
  if (fixed) return launch_class_p<true, false>(c, a, n_cu, s, waves, dry);
  if (f32) return launch_class_p<false, true>(c, a, n_cu, s, waves, dry);
  return launch_class_p<false, false>(c, a, n_cu, s, waves, dry);
}

// Job descriptors and work items staged to the device through two pinned
// buffers used in turn; a buffer is refilled only after the launches that
// read it have completed (its event), so back-to-back calls never overwrite
// descriptors in use.  ws: the Krylov workspace of the waves.
struct SubScratch {
  char* dev[2] = {nullptr, nullptr};
  char* host[2] = {nullptr, nullptr};
  size_t cap[2] = {0, 0};
  hipEvent_t done[2] = {nullptr, nullptr};
  int cur = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* long_ws = nullptr;   // the long-domain kernels' workspace (rm_coop.hip)
  // launch classes run concurrently on auxiliary streams (default priority)
  hipStream_t aux[NCLASS + 1] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[NCLASS + 1] = {};
  // launch lists of the last rm_propagate call, kept for domain reuse
  bool kept = false;
  const void* kept_psi = nullptr;
  int64_t kept_ld = 0, kept_ncols = 0;
  int kept_force = 0;
  bool kept_pairs = false;             // grouped for two wavepackets per wave
  std::vector<int4> kept_cls[NCLASS];
  std::vector<LongItem> kept_blk, kept_coop;
  int4* kept_dev = nullptr;            // the kept lists' items on the device (classes, block, cooperative)
  size_t kept_dev_cap = 0;             // in items
};

void scratch_release(SubScratch* sc) {
  if (!sc) return;
  for (int i = 0; i < 2; ++i) {
    if (sc->done[i]) (void)hipEventSynchronize(sc->done[i]);
    if (sc->dev[i]) (void)hipFree(sc->dev[i]);
    if (sc->host[i]) (void)hipHostFree(sc->host[i]);
    if (sc->done[i]) (void)hipEventDestroy(sc->done[i]);
  }
  if (sc->ws) (void)hipFree(sc->ws);
  if (sc->kept_dev) (void)hipFree(sc->kept_dev);
  coop_scratch_release(sc->long_ws);
  for (int i = 0; i <= NCLASS; ++i) {
    if (sc->aux[i]) {
      (void)hipStreamSynchronize(sc->aux[i]);
      (void)hipStreamDestroy(sc->aux[i]);
    }
    if (sc->join[i]) (void)hipEventDestroy(sc->join[i]);
  }
  if (sc->fork) (void)hipEventDestroy(sc->fork);
  delete sc;
}

// The pair classes: class-0 items of <= 32 rows under a fixed Krylov
// dimension <= 16 (the banded ones, placed in PAIR_BAND_CLASS by the caller,
// keep their own class), grouped by operator, each group padded to an even
// count; other Krylov settings send the banded items back to class 0.
void group_pairs(std::vector<int4> (&cls)[NCLASS], int32_t fixed_m, bool allow = true) {
  const bool pairs = allow && fixed_m >= 1 && fixed_m <= PAIR_MAX_M;
  if (!pairs && !cls[PAIR_BAND_CLASS].empty()) {
    cls[0].insert(cls[0].end(), cls[PAIR_BAND_CLASS].begin(), cls[PAIR_BAND_CLASS].end());
    cls[PAIR_BAND_CLASS].clear();
  }
  auto pair_up = [](std::vector<int4>& v) {
    if (v.empty()) return;
    std::stable_sort(v.begin(), v.end(), [](const int4& x, const int4& y) { return x.w < y.w; });
    std::vector<int4> out;
    for (size_t i = 0; i < v.size();) {
      size_t j = i;
      while (j < v.size() && v[j].w == v[i].w) out.push_back(v[j++]);
      if ((j - i) & 1) out.push_back(make_int4(-1, 0, 0, v[i].w));
      i = j;
    }
    v.swap(out);
  };
  if (pairs && cls[PAIR_CLASS].empty()) {
    std::vector<int4> keep;
    for (const int4& it : cls[0]) (it.z <= 32 ? cls[PAIR_CLASS] : keep).push_back(it);
    cls[0].swap(keep);
    pair_up(cls[PAIR_CLASS]);
    pair_up(cls[PAIR_BAND_CLASS]);
  }
}

// Stage ops and items (the wave classes, then the block and cooperative
// lists) and launch one kernel per non-empty wave class, the block kernel
// and the cooperative launches on s; with timing on, the launches are
// bracketed as one.
// The launches of one call: every non-empty wave class, the block kernel and
// the cooperative launches, on s (or, several launches and no cooperative
// one, each on an auxiliary stream forked from s and joined back into it,
// so they run side by side); each class with a workspace region of its own.
// a0: the jobs (a0.ops, or one job by value in a0.op0); items[c]: device
// work items of class c; d_blk, d_coop: device items of the long lists.
int launch_all(SubScratch* sc, const SubArgs& a0, const int4* const (&items)[NCLASS],
               const std::vector<int4> (&cls)[NCLASS], const int4* d_blk, const std::vector<LongItem>& blk,
               const int4* d_coop, const std::vector<LongItem>& coop, int32_t kcap, int32_t fixed_m,
               int32_t maxorder, double tol, bool f32, int n_cu, hipStream_t s, rm_plan* timed,
               int32_t abort_preset, const RmTuning& tn) {
  SubArgs a = a0;
  size_t need = 0, wsoff[NCLASS] = {};
  int ngroups = blk.empty() ? 0 : 1;
  for (int c = 0; c < NCLASS; ++c) {
    if (cls[c].empty()) continue;
    ++ngroups;
    a.nitems = (int32_t)cls[c].size();
    int waves = 0;
    HIPCHK(launch_class(c, a, f32, n_cu, s, &waves, true));
    const size_t per_wave = (class_basis_lds(c, kcap) ? 0 : (size_t)kcap * 64 * CLASS_R[c]) +
                            (fixed_m > 0 ? 0 : (size_t)kcap * kcap + 2 * (size_t)kcap);
    wsoff[c] = need;
    need += (sizeof(double2) * std::max<size_t>(per_wave, 1) * (size_t)waves + 255) & ~(size_t)255;
  }
  if (sc->ws_bytes < need) {
    if (sc->ws) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(sc->ws));
      sc->ws = nullptr;
      sc->ws_bytes = 0;
    }
    HIPCHK(hipMalloc(&sc->ws, need));
    sc->ws_bytes = need;
  }
  if (timed) HIPCHK(timing_begin(timed, s));
  const bool fork = tn.fork && ngroups > 1 && coop.empty();
  int nused = 0;
  auto next_stream = [&](hipStream_t* out) -> int {
    if (!fork) {
      *out = s;
      return RM_OK;
    }
    const int i = nused++;
    if (!sc->aux[i]) HIPCHK(hipStreamCreateWithFlags(&sc->aux[i], hipStreamNonBlocking));
    if (!sc->join[i]) HIPCHK(hipEventCreateWithFlags(&sc->join[i], hipEventDisableTiming));
    HIPCHK(hipStreamWaitEvent(sc->aux[i], sc->fork, 0));
    *out = sc->aux[i];
    return RM_OK;
  };
  if (fork) {
    if (!sc->fork) HIPCHK(hipEventCreateWithFlags(&sc->fork, hipEventDisableTiming));
    HIPCHK(hipEventRecord(sc->fork, s));
  }
  int rc = RM_OK;
  for (int c = 0; c < NCLASS; ++c) {
    if (cls[c].empty()) continue;
    a.items = items[c];
    a.nitems = (int32_t)cls[c].size();
    int waves = 0;
    HIPCHK(launch_class(c, a, f32, n_cu, s, &waves, true));
    const bool bl = class_basis_lds(c, kcap);
    a.basis = (double2*)((char*)sc->ws + wsoff[c]);
    a.gram = a.basis + (bl ? 0 : (size_t)waves * kcap * 64 * CLASS_R[c]);
    a.cscr = a.gram + (fixed_m > 0 ? 0 : (size_t)waves * kcap * kcap);
    hipStream_t cs;
    if ((rc = next_stream(&cs))) return rc;
    HIPCHK(launch_class(c, a, f32, n_cu, cs, &waves, false));
    if (fork) HIPCHK(hipEventRecord(sc->join[nused - 1], cs));
  }
  hipStream_t bs = s;
  if (!blk.empty() && (rc = next_stream(&bs))) return rc;
  rc = long_run(&sc->long_ws, a0.ops, d_blk, blk, d_coop, coop, kcap, fixed_m, maxorder, tol, f32, n_cu, bs,
                abort_preset, a0.ops ? nullptr : &a0.op0, tn.coop_grid);
  if (rc) return rc;
  if (fork && !blk.empty()) HIPCHK(hipEventRecord(sc->join[nused - 1], bs));
  for (int i = 0; i < (fork ? nused : 0); ++i) HIPCHK(hipStreamWaitEvent(s, sc->join[i], 0));
  if (timed) HIPCHK(timing_end(timed, s));
  return RM_OK;
}

// Stage ops and items (the wave classes, then the block and cooperative
// lists) through the double-buffered pinned staging and launch them
// (launch_all); with timing on, the launches are bracketed as one.
int sub_run(SubScratch* sc, const std::vector<SubOp>& ops, std::vector<int4> (&cls)[NCLASS],
            const std::vector<LongItem>& blk, const std::vector<LongItem>& coop, int32_t kcap,
            int32_t fixed_m, int32_t maxorder, double tol, bool f32, int n_cu, hipStream_t s,
            rm_plan* timed, int32_t abort_preset = 0, const RmTuning& tn = RmTuning()) {
  group_pairs(cls, fixed_m, tn.pair != 0);
  size_t nitems = blk.size() + coop.size();
  for (int c = 0; c < NCLASS; ++c) nitems += cls[c].size();
  if (nitems == 0) return RM_OK;
  const size_t ops_bytes = (ops.size() * sizeof(SubOp) + 255) & ~(size_t)255;
  const size_t bytes = ops_bytes + nitems * sizeof(int4);
  const int bi = sc->cur;
  sc->cur ^= 1;
  if (sc->done[bi]) HIPCHK(hipEventSynchronize(sc->done[bi]));
  if (sc->cap[bi] < bytes) {
    // both buffers grow together (the other one after its launches are
    // done), so that alternating calls of one size allocate only once
    const size_t cap = std::max<size_t>(bytes * 2, 64 << 10);
    for (int h : {bi, bi ^ 1}) {
      if (sc->cap[h] >= cap) continue;
      if (h != bi && sc->done[h]) HIPCHK(hipEventSynchronize(sc->done[h]));
      if (sc->dev[h]) HIPCHK(hipFree(sc->dev[h]));
      if (sc->host[h]) HIPCHK(hipHostFree(sc->host[h]));
      sc->dev[h] = nullptr;
      sc->host[h] = nullptr;
      sc->cap[h] = 0;
      HIPCHK(hipMalloc((void**)&sc->dev[h], cap));
      HIPCHK(hipHostMalloc((void**)&sc->host[h], cap, hipHostMallocDefault));
      sc->cap[h] = cap;
    }
  }
  memcpy(sc->host[bi], ops.data(), ops.size() * sizeof(SubOp));
  const int4* items[NCLASS];
  size_t pos = ops_bytes;
  for (int c = 0; c < NCLASS; ++c) {
    items[c] = (const int4*)(sc->dev[bi] + pos);
    if (!cls[c].empty()) memcpy(sc->host[bi] + pos, cls[c].data(), cls[c].size() * sizeof(int4));
    pos += cls[c].size() * sizeof(int4);
  }
  const size_t off_blk = pos;
  for (const LongItem& it : blk) {
    memcpy(sc->host[bi] + pos, &it.item, sizeof(int4));
    pos += sizeof(int4);
  }
  const size_t off_coop = pos;
  for (const LongItem& it : coop) {
    memcpy(sc->host[bi] + pos, &it.item, sizeof(int4));
    pos += sizeof(int4);
  }
  HIPCHK(hipMemcpyAsync(sc->dev[bi], sc->host[bi], bytes, hipMemcpyHostToDevice, s));
  SubArgs a;
  memset(&a, 0, sizeof(a));
  a.ops = (const SubOp*)sc->dev[bi];
  a.kcap = kcap;
  a.fixed_m = fixed_m;
  a.maxorder = maxorder;
  a.tol = tol;
  for (const SubOp& op : ops) a.obs |= op.ow > 0 ? 1 : 0;
  const int rc = launch_all(sc, a, items, cls, (const int4*)(sc->dev[bi] + off_blk), blk,
                            (const int4*)(sc->dev[bi] + off_coop), coop, kcap, fixed_m, maxorder, tol, f32,
                            n_cu, s, timed, abort_preset, tn);
  if (rc) return rc;
  if (!sc->done[bi]) HIPCHK(hipEventCreateWithFlags(&sc->done[bi], hipEventDisableTiming));
  HIPCHK(hipEventRecord(sc->done[bi], s));
  return RM_OK;
}

// Every stored entry of rows [lo, hi) couples a row only to itself or to its
// neighbours in the permuted order.
bool sub_banded(const rm_plan* p, int32_t lo, int32_t hi) {
  const int64_t n = p->n;
  if (p->h_sub_col.size() < (size_t)p->sub_w * (size_t)n) return false;
  for (int32_t r = lo; r < hi; ++r)
    for (int32_t j = 0; j < p->sub_w; ++j) {
      const int32_t c = p->h_sub_col[(size_t)j * n + r];
      if (c < r - 1 || c > r + 1) return false;
    }
  return true;
}

// Domains [lo, hi) of ncols columns into p->h_dom (synchronises s).
int sub_domains(rm_plan* p, const void* psi, int64_t ld, int64_t ncols, hipStream_t s) {
  const int64_t n = p->n;
  int rc = grow(&p->dom, &p->dom_cap, 2 * ncols);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sub_fill, dim3(grid_for(2 * ncols, 256)), dim3(256), 0, s, p->dom, 2 * ncols,
                     (int32_t)n, 0);
  hipLaunchKernelGGL(k_col_domain, dim3(grid_for(n, 256), (unsigned)ncols), dim3(256), 0, s, n, ld,
                     (const double2*)psi, p->sub_label, p->sub_lo, p->sub_hi, p->dom);
  HIPCHK(hipGetLastError());
  if (p->h_dom_pin_cap < 2 * ncols) {
    if (p->h_dom_pin) HIPCHK(hipHostFree(p->h_dom_pin));
    p->h_dom_pin = nullptr;
    p->h_dom_pin_cap = 0;
    const int64_t cap = std::max<int64_t>(2 * ncols, 16384);
    HIPCHK(hipHostMalloc((void**)&p->h_dom_pin, sizeof(int32_t) * (size_t)cap, hipHostMallocDefault));
    p->h_dom_pin_cap = cap;
  }
  HIPCHK(hipMemcpyAsync(p->h_dom_pin, p->dom, sizeof(int32_t) * 2 * ncols, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  p->h_dom.assign(p->h_dom_pin, p->h_dom_pin + 2 * ncols);
  return RM_OK;
}

void step_flags(rm_plan* p, int64_t nsteps, const double* fprod_dev, int32_t* sflag, hipStream_t s) {
  hipLaunchKernelGGL(k_sub_fill, dim3(grid_for(nsteps, 256)), dim3(256), 0, s, sflag, nsteps, 0, 0);
  if (p->nterm > 0)
    hipLaunchKernelGGL(k_step_flags, dim3(grid_for(p->nterm, 256), (unsigned)std::min<int64_t>(nsteps, 65535)),
                       dim3(256), 0, s, p->nterm, nsteps, p->nfield, p->term_ptr, p->term_field,
                       p->term_mval, (const double2*)fprod_dev, sflag);
}

SubOp plan_op(const rm_plan* p) {
  SubOp op;
  memset(&op, 0, sizeof(op));
  op.col = p->sub_col;
  op.val = p->sub_val;
  op.rp = p->sub_rp;
  op.ccol = p->sub_ccol;
  op.cval = p->sub_cval;
  op.cval32 = p->sub_cval32;
  op.nnz = p->nnz;
  op.perm = p->sub_perm;
  op.n = (int32_t)p->n;
  op.w = p->sub_w;
  op.nslot = p->nfield;
  return op;
}

}  // namespace

void sub_scratch_release(void* sc) { scratch_release((SubScratch*)sc); }

extern "C" int rm_plan_subspaces(rm_plan* p, int64_t* count, int64_t* max_rows) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_subspaces: NULL plan");
  int rc = sub_build(p);
  if (rc) return rc;
  if (count) *count = p->nsub;
  if (max_rows) *max_rows = p->max_sub;
  return RM_OK;
}

extern "C" int rm_plan_set_tuning(rm_plan* p, int32_t knob, int32_t value) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_tuning: NULL plan");
  switch (knob) {
    case RM_TUNE_PAIR:
    case RM_TUNE_FORK:
      if (value != 0 && value != 1) return fail(RM_ERR_ARG, "rm_plan_set_tuning: knob %d takes 0 or 1", knob);
      (knob == RM_TUNE_PAIR ? p->tune.pair : p->tune.fork) = value;
      break;
    case RM_TUNE_COOP_MAX:
      if (value < 0) return fail(RM_ERR_ARG, "rm_plan_set_tuning: negative cooperative column limit %d", value);
      p->tune.coop_max = value;
      break;
    case RM_TUNE_COOP_GRID:
      if (value < 0) return fail(RM_ERR_ARG, "rm_plan_set_tuning: negative cooperative grid %d", value);
      p->tune.coop_grid = value;
      break;
    case RM_TUNE_WAVE_MAX_ROWS:
      if (value < 0 || value > RM_SUBSPACE_MAX_ROWS)
        return fail(RM_ERR_ARG, "rm_plan_set_tuning: one-wave domain limit %d outside 0..%d", value,
                    RM_SUBSPACE_MAX_ROWS);
      p->tune.wave_max_rows = value;
      break;
    case RM_TUNE_COOP_MIN_ENTRIES:
      if (value < 0) return fail(RM_ERR_ARG, "rm_plan_set_tuning: negative entry count %d", value);
      p->tune.coop_min_entries = value;
      break;
    default:
      return fail(RM_ERR_ARG, "rm_plan_set_tuning: unknown knob %d", knob);
  }
  // launch lists kept for domain reuse were grouped under the old choices
  if (p->sub_scratch) ((SubScratch*)p->sub_scratch)->kept = false;
  return RM_OK;
}

extern "C" int rm_plan_set_domain_reuse(rm_plan* p, int32_t on) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_domain_reuse: NULL plan");
  p->dom_reuse = on ? 1 : 0;
  if (p->sub_scratch) ((SubScratch*)p->sub_scratch)->kept = false;
  return RM_OK;
}

extern "C" int rm_plan_set_propagator(rm_plan* p, int32_t propag) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_propagator: NULL plan");
  if (propag != RM_PROPAG_LANCZOS && propag != RM_PROPAG_EXPOKIT)
    return fail(RM_ERR_ARG, "rm_plan_set_propagator: unknown propagator %d", propag);
  p->propag = propag;
  return RM_OK;
}

extern "C" int rm_plan_set_coop_abort(rm_plan* p, int32_t preset) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_coop_abort: NULL plan");
  p->coop_abort = preset ? 1 : 0;
  return RM_OK;
}

extern "C" int rm_plan_set_precision(rm_plan* p, int32_t precision) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_set_precision: NULL plan");
  if (precision != RM_PREC_FP64 && precision != RM_PREC_FP32)
    return fail(RM_ERR_ARG, "rm_plan_set_precision: unknown precision %d", precision);
  p->precision = precision;
  return RM_OK;
}

// A long-domain item of plan p with its domain's entry count and longest row.
LongItem long_item(const rm_plan* p, int col, int lo, int len, int op) {
  LongItem it;
  it.item = make_int4(col, lo, len, op);
  it.nent = p->h_sub_rp[lo + len] - p->h_sub_rp[lo];
  int64_t mr = 0;
  for (int64_t r = lo; r < (int64_t)lo + len; ++r) mr = std::max(mr, p->h_sub_rp[r + 1] - p->h_sub_rp[r]);
  it.maxrow = (int32_t)std::min<int64_t>(mr, INT32_MAX);
  it.nslot = p->nfield;
  return it;
}

// Most work first, with the domains the block kernel stages in LDS
// (RM_XL_ROWS rows at most) ahead of the longer ones.
constexpr int RM_XL_ROWS = 2048;
void sort_long(std::vector<LongItem>& v) {
  std::stable_sort(v.begin(), v.end(), [](const LongItem& x, const LongItem& y) {
    const bool lx = x.item.z <= RM_XL_ROWS, ly = y.item.z <= RM_XL_ROWS;
    if (lx != ly) return lx;
    return x.nent > y.nent;
  });
}

// The subspace path of rm_propagate.  Arguments are validated by the caller
// (ld, ncols, Krylov parameters, pointers).  Domains of up to
// p->tune.wave_max_rows rows (64 unless tuned, at most RM_SUBSPACE_MAX_ROWS)
// go to the one-wave kernels; longer ones to the cooperative kernel when
// there are at most p->tune.coop_max of them and each holds at least
// their count times p->tune.coop_min_entries stored entries, else to the
// block kernel.  force: RM_KERNEL_COOP / RM_KERNEL_BLOCK send every
// wavepacket to that kernel (0: the rule above).  *done = false when the
// path does not apply; nothing has then been written to psi.
// The observable plan's matrix in p's permuted ELL form (p->obs_col,
// p->obs_val, width p->obs_w), rebuilt when the plan or its matrix changed.
__global__ void k_obs_ell(int64_t n, int32_t ow, const int32_t* __restrict__ perm,
                          const int32_t* __restrict__ iperm, const int64_t* __restrict__ rp,
                          const int32_t* __restrict__ col, const double2* __restrict__ val,
                          int32_t* __restrict__ ocol, double2* __restrict__ oval) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int32_t r = perm[q];
  const int64_t e0 = rp[r], e1 = rp[r + 1];
  for (int32_t j = 0; j < ow; ++j) {
    const int64_t e = e0 + j;
    ocol[(int64_t)j * n + q] = e < e1 ? iperm[col[e]] : (int32_t)q;
    oval[(int64_t)j * n + q] = e < e1 ? val[e] : make_double2(0.0, 0.0);
  }
}

int obs_prepare(rm_plan* p, const rm_plan* obs, hipStream_t s) {
  if (p->obs_src == obs && p->obs_gen == obs->field_gen && p->obs_col) return RM_OK;
  int64_t w = 1;
  for (int64_t r = 0; r < obs->n; ++r) w = std::max<int64_t>(w, obs->h_row_ptr[r + 1] - obs->h_row_ptr[r]);
  if (w * p->n > INT32_MAX) return fail(RM_ERR_ARG, "rm_propagate_expect: observable too wide (%lld entries in a row)",
                                        (long long)w);
  int rc;
  if ((rc = grow(&p->obs_col, &p->obs_cap, w * p->n)) || (rc = grow(&p->obs_val, &p->obs_vcap, w * p->n)))
    return rc;
  hipLaunchKernelGGL(k_obs_ell, dim3(grid_for(p->n, 256)), dim3(256), 0, s, p->n, (int32_t)w, p->sub_perm,
                     p->sub_iperm, obs->row_ptr, obs->col, obs->aval, p->obs_col, p->obs_val);
  HIPCHK(hipGetLastError());
  p->obs_src = obs;
  p->obs_gen = obs->field_gen;
  p->obs_w = (int32_t)w;
  return RM_OK;
}

int sub_propagate(rm_plan* p, void* psi, int64_t ld, int64_t ncols, int64_t nsteps,
                  const double* fprod_dev, double f_re, double f_im, const void* phase,
                  int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                  int32_t* status, int force, hipStream_t s, bool* done, const rm_plan* obs,
                  double2* expect) {
  *done = false;
  if (nsteps > INT32_MAX) return RM_OK;
  int rc = sub_build(p);
  if (rc) return rc;
  if (obs && (rc = obs_prepare(p, obs, s))) return rc;
  if (!p->sub_scratch) p->sub_scratch = new (std::nothrow) SubScratch();
  if (!p->sub_scratch) return fail(RM_ERR_ALLOC, "rm_propagate: host allocation");
  SubScratch* sc = (SubScratch*)p->sub_scratch;
  // the caller passed no status: an internal one, so that a cooperative
  // timeout can still be reported
  int32_t* st = status;
  if (!st) {
    if ((rc = grow(&p->int_status, &p->int_status_cap, ncols))) return rc;
    st = p->int_status;
  }
  const bool pairs = p->tune.pair && krylov_fixed >= 1 && krylov_fixed <= PAIR_MAX_M;
  // the Expokit propagator takes fresh domains at every call (its items are
  // not kept)
  const bool expo = p->propag == RM_PROPAG_EXPOKIT;
  const bool reuse = !expo && p->dom_reuse && sc->kept && sc->kept_psi == psi && sc->kept_ld == ld &&
                     sc->kept_ncols == ncols && sc->kept_force == force && sc->kept_pairs == pairs;
  // step flags in chunks of <= 1024 terms, at most 16 per step; from 64
  // steps on, or for up to 8192 terms, one block per step (a step with a
  // field ends its scan at the first non-zero element, in its first pass:
  // more blocks per step would only repeat that first pass, and a step
  // without one scans at most 32 passes)
  const int32_t parts = (nsteps >= 64 || p->nterm <= 8192)
                            ? 1
                            : (int32_t)std::max<int64_t>(1, std::min<int64_t>(16, (p->nterm + 1023) / 1024));
  const int64_t chunk = std::max<int64_t>(1, (p->nterm + parts - 1) / parts);
  if (nsteps * parts > INT32_MAX) return RM_OK;
  if ((rc = grow(&p->sflag, &p->sflag_cap, nsteps))) return rc;
  if (parts > 1) {
    if ((rc = grow(&p->sflag_part, &p->sflag_part_cap, nsteps * parts))) return rc;
    const int64_t cap0 = p->sflag_arrive_cap;
    if ((rc = grow(&p->sflag_arrive, &p->sflag_arrive_cap, nsteps))) return rc;
    if (p->sflag_arrive_cap != cap0)   // a new counter array starts at zero
      hipLaunchKernelGGL(k_sub_fill, dim3(grid_for(p->sflag_arrive_cap, 256)), dim3(256), 0, s,
                         (int32_t*)p->sflag_arrive, p->sflag_arrive_cap, 0, 0);
  }
  if (!p->stale) {
    if ((rc = upload(&p->stale, (const int32_t*)nullptr, 1))) return rc;
    hipLaunchKernelGGL(k_sub_fill, dim3(1), dim3(64), 0, s, p->stale, (int64_t)1, 0, 0);
  }
  p->call_id = p->call_id == INT32_MAX ? 1 : p->call_id + 1;
  // step flags, output fills and (reused domains) the check in one launch;
  // without reuse it runs while the host waits for the domains
  Prologue pa;
  memset(&pa, 0, sizeof(pa));
  pa.nterm = p->nterm;
  pa.nsteps = nsteps;
  pa.nfield = p->nfield;
  pa.term_ptr = p->term_ptr;
  pa.term_field = p->term_field;
  pa.term_mval = p->term_mval;
  pa.table = (const double2*)fprod_dev;
  pa.sflag = p->sflag;
  pa.kused = kused;
  pa.nk = kused ? nsteps * ncols : 0;
  pa.status = st;
  pa.ncols = ncols;
  pa.n = p->n;
  pa.ld = ld;
  pa.psi = (const double2*)psi;
  pa.label = p->sub_label;
  pa.clo = p->sub_lo;
  pa.chi = p->sub_hi;
  pa.dom = reuse ? p->dom_keep : nullptr;
  pa.stale = p->stale;
  pa.call_id = p->call_id;
  pa.parts = parts;
  pa.chunk = chunk;
  pa.part = p->sflag_part;
  pa.arrive = p->sflag_arrive;
  pa.nb_flag = (int32_t)(nsteps * parts);
  pa.ex = obs ? expect : nullptr;
  pa.nex = obs ? nsteps * ncols : 0;
  pa.nb_fill = (int32_t)std::max<int64_t>(
      1, std::min<int64_t>(grid_for(std::max(std::max(pa.nk, pa.nex), ncols), 256), 1024));
  pa.cpb = (int32_t)std::max<int64_t>(1, (int64_t)grid_for(p->n, 1024));
  const int64_t nb = (int64_t)pa.nb_flag + pa.nb_fill + (reuse ? pa.cpb * ncols : 0);
  hipLaunchKernelGGL(k_sub_prologue, dim3((unsigned)nb), dim3(256), 0, s, pa);
  HIPCHK(hipGetLastError());
  // the job of this call (one operator, every column)
  auto make_op = [&]() {
    SubOp op = plan_op(p);
    op.phase = (const double2*)phase;
    op.table = (const double2*)fprod_dev;
    op.sflag = p->sflag;
    op.psi = (double2*)psi;
    op.ld = ld;
    op.kused = kused;
    op.kused_ld = ncols;
    op.status = st;
    op.stale = reuse ? p->stale : nullptr;
    op.call_id = p->call_id;
    op.f = make_double2(f_re, f_im);
    op.nsteps = (int32_t)nsteps;
    if (obs) {
      op.ocol = p->obs_col;
      op.oval = p->obs_val;
      op.ow = p->obs_w;
      op.expect = expect;
      op.expect_ld = ncols;
    }
    return op;
  };
  if (!reuse) {
    sc->kept = false;
    if ((rc = sub_domains(p, psi, ld, ncols, s))) return rc;
    if (expo) {
      // Expokit: one workgroup per column on its domain, any length
      std::vector<int4> items;
      for (int64_t b = 0; b < ncols; ++b) {
        const int32_t lo = p->h_dom[2 * b], hi = p->h_dom[2 * b + 1];
        if (hi > lo) items.push_back(make_int4((int)b, lo, hi - lo, 0));
      }
      if (!p->sub_rp && (rc = csr_build(p, entry_values(p)))) return rc;
      *done = true;
      const SubOp op = make_op();
      if (p->timing) HIPCHK(timing_begin(p, s));
      rc = expokit_run(op, items, krylov_fixed, tol, p->n_cu, &p->expo_ws, &p->expo_ws_bytes, &p->expo_anorm,
                       &p->expo_anorm_cap, &p->expo_items, &p->expo_items_cap, s);
      if (rc) return rc;
      if (p->timing) HIPCHK(timing_end(p, s));
      return RM_OK;
    }
    for (auto& v : sc->kept_cls) v.clear();
    sc->kept_blk.clear();
    sc->kept_coop.clear();
    std::vector<int4> big;
    for (int64_t b = 0; b < ncols; ++b) {
      const int32_t lo = p->h_dom[2 * b], hi = p->h_dom[2 * b + 1];
      if (hi <= lo) continue;   // zero column: stays zero (kused 0)
      if (force || hi - lo > std::min<int32_t>(p->tune.wave_max_rows, RM_SUBSPACE_MAX_ROWS) || !p->sub_col) {
        big.push_back(make_int4((int)b, lo, hi - lo, 0));
      } else {
        const int c = sub_class(p->sub_w, hi - lo);
        sc->kept_cls[c == 0 && hi - lo <= 32 && p->nfield <= NSR && sub_banded(p, lo, hi) ? PAIR_BAND_CLASS : c]
            .push_back(make_int4((int)b, lo, hi - lo, 0));
      }
    }
    if (!big.empty() && !p->sub_rp && (rc = csr_build(p, entry_values(p)))) return rc;
    // the cooperative form runs its wavepackets one launch after another on
    // many workgroups, the block form all of them at once, one workgroup
    // each: cooperative while the count times ~80 us (a cooperative step)
    // stays below one workgroup's step on the smallest of them (~0.0125 us
    // per stored entry), i.e. count * coop_min_entries <= its entries
    std::vector<LongItem> lbig;
    int64_t emin = INT64_MAX;
    for (const int4& it : big) {
      lbig.push_back(long_item(p, it.x, it.y, it.z, 0));
      emin = std::min<int64_t>(emin, lbig.back().nent);
    }
    const bool to_coop = force == RM_KERNEL_COOP ||
                         (force == 0 && (int64_t)big.size() <= p->tune.coop_max &&
                          (int64_t)big.size() * p->tune.coop_min_entries <= emin);
    for (const LongItem& it : lbig) (to_coop ? sc->kept_coop : sc->kept_blk).push_back(it);
    sort_long(sc->kept_blk);
    for (auto& v : sc->kept_cls)
      std::stable_sort(v.begin(), v.end(), [](const int4& x, const int4& y) { return x.z > y.z; });
    group_pairs(sc->kept_cls, krylov_fixed, p->tune.pair != 0);
    sc->kept_pairs = pairs;
    if (p->dom_reuse) {
      // a copy of their own: rm_plan_domains rewrites p->dom
      if ((rc = grow(&p->dom_keep, &p->dom_keep_cap, 2 * ncols))) return rc;
      HIPCHK(hipMemcpyAsync(p->dom_keep, p->dom, sizeof(int32_t) * 2 * (size_t)ncols, hipMemcpyDeviceToDevice, s));
      // the launch lists' items uploaded once (classes in order, then the
      // block and the cooperative items); later calls launch from them
      std::vector<int4> all;
      for (int c = 0; c < NCLASS; ++c) all.insert(all.end(), sc->kept_cls[c].begin(), sc->kept_cls[c].end());
      for (const LongItem& it : sc->kept_blk) all.push_back(it.item);
      for (const LongItem& it : sc->kept_coop) all.push_back(it.item);
      if (sc->kept_dev_cap < all.size() || !sc->kept_dev) {
        if (sc->kept_dev) {
          HIPCHK(hipStreamSynchronize(s));
          HIPCHK(hipFree(sc->kept_dev));
          sc->kept_dev = nullptr;
        }
        sc->kept_dev_cap = std::max<size_t>(all.size(), 4096);
        HIPCHK(hipMalloc((void**)&sc->kept_dev, sizeof(int4) * sc->kept_dev_cap));
      }
      if (!all.empty()) {
        HIPCHK(hipMemcpyAsync(sc->kept_dev, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));   // `all` is released at the end of this block
      }
      sc->kept = true;
      sc->kept_psi = psi;
      sc->kept_ld = ld;
      sc->kept_ncols = ncols;
      sc->kept_force = force;
    }
  }
  *done = true;
  std::vector<SubOp> ops(1, make_op());
  const SubOp& op = ops[0];
  const int32_t kcap = krylov_fixed > 0 ? krylov_fixed : maxorder;
  const bool f32 = p->precision == RM_PREC_FP32;
  if (sc->kept) {
    // kept lists: nothing to stage, the job goes by value
    SubArgs a;
    memset(&a, 0, sizeof(a));
    a.ops = nullptr;
    a.op0 = op;
    a.obs = op.ow > 0 ? 1 : 0;
    a.kcap = kcap;
    a.fixed_m = krylov_fixed;
    a.maxorder = maxorder;
    a.tol = tol;
    const int4* items[NCLASS];
    size_t pos = 0;
    for (int c = 0; c < NCLASS; ++c) {
      items[c] = sc->kept_dev + pos;
      pos += sc->kept_cls[c].size();
    }
    const int4* d_blk = sc->kept_dev + pos;
    const int4* d_coop = d_blk + sc->kept_blk.size();
    rc = launch_all(sc, a, items, sc->kept_cls, d_blk, sc->kept_blk, d_coop, sc->kept_coop, kcap, krylov_fixed,
                    maxorder, tol, f32, p->n_cu, s, p, p->coop_abort, p->tune);
  } else {
    std::vector<int4> cls[NCLASS];
    for (int c = 0; c < NCLASS; ++c) cls[c] = sc->kept_cls[c];
    rc = sub_run(sc, ops, cls, sc->kept_blk, sc->kept_coop, kcap, krylov_fixed, maxorder, tol, f32, p->n_cu, s, p,
                 p->coop_abort, p->tune);
  }
  if (rc || status || sc->kept_coop.empty()) return rc;
  // no status array from the caller: read the internal one for timeouts
  if (p->h_status_pin_cap < ncols) {
    if (p->h_status_pin) HIPCHK(hipHostFree(p->h_status_pin));
    p->h_status_pin = nullptr;
    p->h_status_pin_cap = 0;
    HIPCHK(hipHostMalloc((void**)&p->h_status_pin, sizeof(int32_t) * (size_t)ncols, hipHostMallocDefault));
    p->h_status_pin_cap = ncols;
  }
  HIPCHK(hipMemcpyAsync(p->h_status_pin, st, sizeof(int32_t) * (size_t)ncols, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int64_t b = 0; b < ncols; ++b)
    if (p->h_status_pin[b] & RM_DEVICE_TIMEOUT)
      return fail(RM_DEVICE_TIMEOUT, "rm_propagate: a cooperative launch's bounded wait gave up (column %lld); "
                  "the result is invalid", (long long)b);
  return RM_OK;
}

// ---------------------------------------------------------------------------
// batches of independent jobs (continuous batching)

struct rm_batch {
  SubScratch* sc = nullptr;
  int device = 0;
  int n_cu = 256;
  int32_t block_rows = RM_SUBSPACE_MAX_ROWS;   // longer domains: one workgroup per job
};

extern "C" int rm_batch_create(rm_batch** out) {
  if (!out) return fail(RM_ERR_ARG, "rm_batch_create: out is NULL");
  *out = nullptr;
  rm_batch* b = new (std::nothrow) rm_batch();
  if (!b) return fail(RM_ERR_ALLOC, "rm_batch_create: host allocation");
  b->sc = new (std::nothrow) SubScratch();
  if (!b->sc) {
    delete b;
    return fail(RM_ERR_ALLOC, "rm_batch_create: host allocation");
  }
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess ||
      b->n_cu <= 0)
    b->n_cu = 256;
  *out = b;
  return RM_OK;
}

extern "C" int rm_batch_set_block_rows(rm_batch* b, int32_t rows) {
  if (!b) return fail(RM_ERR_ARG, "rm_batch_set_block_rows: NULL batch");
  if (rows < 0 || rows > RM_SUBSPACE_MAX_ROWS)
    return fail(RM_ERR_ARG, "rm_batch_set_block_rows: rows=%d outside [0, %d]", rows, RM_SUBSPACE_MAX_ROWS);
  b->block_rows = rows;
  return RM_OK;
}

extern "C" int rm_batch_destroy(rm_batch* b) {
  if (!b) return RM_OK;
  scratch_release(b->sc);
  delete b;
  return RM_OK;
}

extern "C" int rm_plan_domains(rm_plan* p, const void* psi, int64_t ld, int64_t ncols, int32_t* lohi,
                               void* stream) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_domains: NULL plan");
  if (ncols < 0 || ncols > 65535 || ld < p->n) return fail(RM_ERR_ARG, "rm_plan_domains: bad ld/ncols");
  if (ncols == 0 || p->n == 0) return RM_OK;
  if (!psi || !lohi) return fail(RM_ERR_ARG, "rm_plan_domains: NULL pointer");
  int rc = sub_build(p);
  if (rc) return rc;
  if ((rc = sub_domains(p, psi, ld, ncols, (hipStream_t)stream))) return rc;
  memcpy(lohi, p->h_dom.data(), sizeof(int32_t) * 2 * ncols);
  return RM_OK;
}

extern "C" int rm_plan_step_flags(rm_plan* p, int64_t nsteps, const double* fprod_dev, int32_t* sflag,
                                  void* stream) {
  if (!p) return fail(RM_ERR_ARG, "rm_plan_step_flags: NULL plan");
  if (nsteps < 0) return fail(RM_ERR_ARG, "rm_plan_step_flags: nsteps < 0");
  if (nsteps == 0) return RM_OK;
  if (!sflag || (p->nfield > 0 && !fprod_dev)) return fail(RM_ERR_ARG, "rm_plan_step_flags: NULL pointer");
  step_flags(p, nsteps, fprod_dev, sflag, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

// Jobs of a batch call as launch lists: validated, one SubOp each, wave
// classes and block-form items, longest jobs first.  fixed_plans: every plan
// must already hold its subspace arrays (built by an earlier call; nothing
// is uploaded here) and every job must fit the one-wave kernels.
static int collect_jobs(const char* who, const rm_job* jobs, int32_t njobs, int32_t job_bytes, int32_t block_rows,
                 bool fixed_plans, std::vector<SubOp>& ops, std::vector<int4> (&cls)[NCLASS],
                 std::vector<LongItem>& blk) {
  if (job_bytes != (int32_t)sizeof(rm_job))
    return fail(RM_ERR_ARG, "%s: job records of %d bytes, the ABI's rm_job has %d", who, job_bytes,
                (int)sizeof(rm_job));
  if (njobs < 0 || (njobs > 0 && !jobs)) return fail(RM_ERR_ARG, "%s: bad job list", who);
  ops.reserve((size_t)njobs);
  std::vector<int64_t> work[NCLASS];
  std::vector<int64_t> blk_work;
  for (int32_t j = 0; j < njobs; ++j) {
    const rm_job& jb = jobs[j];
    rm_plan* p = jb.plan;
    if (!p) return fail(RM_ERR_ARG, "%s: job %d has no plan", who, j);
    if (jb.nsteps < 0) return fail(RM_ERR_ARG, "%s: job %d nsteps < 0", who, j);
    if (jb.nsteps == 0) continue;
    int rc = RM_OK;
    if (fixed_plans) {
      if (!p->sub_ready || !p->sub_col)
        return fail(RM_ERR_ARG, "%s: job %d: the plan's subspace arrays are not built (call rm_plan_subspaces "
                    "on it once beforehand)", who, j);
    } else if ((rc = sub_build(p))) {
      return rc;
    }
    if (!jb.psi || jb.ld < p->n || jb.col < 0 || !jb.sflag_dev || (p->nfield > 0 && !jb.fprod_dev))
      return fail(RM_ERR_ARG, "%s: job %d: bad wavepacket, field table or flags", who, j);
    // the domain must be a union of whole subspaces, or gathers would leave it
    const int32_t lo = jb.lo, len = jb.len;
    if (len < 1 || len > RM_BLOCK_MAX_ROWS || lo < 0 || (int64_t)lo + len > p->n ||
        p->h_sub_start[p->h_sub_label_perm[lo]] != lo ||
        p->h_sub_end[p->h_sub_label_perm[lo + len - 1]] != lo + len)
      return fail(RM_ERR_ARG, "%s: job %d: [%d, %d) is not a domain of %d rows at most "
                  "(use rm_plan_domains)", who, j, lo, lo + len, RM_BLOCK_MAX_ROWS);
    const bool longdom = len > block_rows || !p->sub_col;
    if (longdom && fixed_plans)
      return fail(RM_ERR_ARG, "%s: job %d: a domain of %d rows needs the block kernel, whose workspace this "
                  "entry point does not hold (at most %d rows; use rm_batch_propagate)", who, j, len,
                  RM_SUBSPACE_MAX_ROWS);
    if (longdom && !p->sub_rp && (rc = csr_build(p, entry_values(p)))) return rc;
    SubOp op = plan_op(p);
    op.phase = (const double2*)jb.phase;
    op.table = (const double2*)jb.fprod_dev;
    op.sflag = jb.sflag_dev;
    op.psi = (double2*)jb.psi;
    op.ld = jb.ld;
    op.kused = jb.kused;
    op.kused_ld = 1;
    op.kcol0 = jb.col;
    op.status = jb.status;
    op.f = make_double2(jb.f_re, jb.f_im);
    op.nsteps = jb.nsteps;
    if (longdom) {
      blk.push_back(long_item(p, jb.col, lo, len, (int)ops.size()));
      blk_work.push_back((int64_t)jb.nsteps * blk.back().nent);
    } else {
      const int c = sub_class(p->sub_w, len);
      cls[c].push_back(make_int4(jb.col, lo, len, (int)ops.size()));
      work[c].push_back((int64_t)jb.nsteps * len);
    }
    ops.push_back(op);
  }
  // longest jobs first within a class (the waves take items in order)
  for (int c = 0; c < NCLASS; ++c) {
    std::vector<size_t> ix(cls[c].size());
    std::iota(ix.begin(), ix.end(), 0);
    std::stable_sort(ix.begin(), ix.end(), [&](size_t x, size_t y) { return work[c][x] > work[c][y]; });
    std::vector<int4> sorted(ix.size());
    for (size_t t = 0; t < ix.size(); ++t) sorted[t] = cls[c][ix[t]];
    cls[c].swap(sorted);
  }
  // long domains: one workgroup each (block kernel), LDS-staged domains
  // first, most work first
  std::vector<size_t> ix(blk.size());
  std::iota(ix.begin(), ix.end(), 0);
  std::stable_sort(ix.begin(), ix.end(), [&](size_t x, size_t y) {
    const bool lx = blk[x].item.z <= RM_XL_ROWS, ly = blk[y].item.z <= RM_XL_ROWS;
    if (lx != ly) return lx;
    return blk_work[x] > blk_work[y];
  });
  std::vector<LongItem> sorted(ix.size());
  for (size_t t = 0; t < ix.size(); ++t) sorted[t] = blk[ix[t]];
  blk.swap(sorted);
  return RM_OK;
}

static int check_krylov(const char* who, int32_t krylov_fixed, int32_t maxorder, double tol, int32_t precision) {
  if (krylov_fixed < 0 || krylov_fixed > RM_MAX_KRYLOV)
    return fail(RM_ERR_ARG, "%s: krylov_fixed=%d outside [0, %d]", who, krylov_fixed, RM_MAX_KRYLOV);
  if (krylov_fixed == 0 && (maxorder < 1 || maxorder > RM_MAX_KRYLOV))
    return fail(RM_ERR_ARG, "%s: maxorder=%d outside [1, %d]", who, maxorder, RM_MAX_KRYLOV);
  if (krylov_fixed == 0 && !(tol > 0.0 && tol <= 1.0)) return fail(RM_ERR_ARG, "%s: tol must be in (0, 1]", who);
  if (precision != RM_PREC_FP64 && precision != RM_PREC_FP32)
    return fail(RM_ERR_ARG, "%s: unknown precision %d", who, precision);
  return RM_OK;
}

extern "C" int rm_batch_propagate(rm_batch* bt, const rm_job* jobs, int32_t njobs, int32_t job_bytes,
                                  int32_t krylov_fixed, int32_t maxorder, double tol, int32_t precision,
                                  void* stream) {
  if (!bt) return fail(RM_ERR_ARG, "rm_batch_propagate: NULL batch");
  int rc = check_krylov("rm_batch_propagate", krylov_fixed, maxorder, tol, precision);
  if (rc) return rc;
  std::vector<SubOp> ops;
  std::vector<int4> cls[NCLASS];
  std::vector<LongItem> blk;
  if ((rc = collect_jobs("rm_batch_propagate", jobs, njobs, job_bytes, bt->block_rows, false, ops, cls, blk)))
    return rc;
  const std::vector<LongItem> none;
  return sub_run(bt->sc, ops, cls, blk, none, krylov_fixed > 0 ? krylov_fixed : maxorder, krylov_fixed,
                 maxorder, tol, precision == RM_PREC_FP32, bt->n_cu, (hipStream_t)stream, nullptr);
}

// ---------------------------------------------------------------------------
// Allocation-free batch calls (rm_batch_propagate_static).  The descriptors
// and the Krylov workspace live in `scratch`, one copy per device (the
// library holds it while the process runs); host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.

static constexpr size_t STATIC_STAGE_BYTES = (size_t)RM_STATIC_MAX_JOBS * (sizeof(SubOp) + sizeof(int4)) + 512;
static_assert(STATIC_STAGE_BYTES < (size_t)RM_STATIC_SCRATCH_BYTES / 4, "descriptor staging exceeds its share");
static __device__ __align__(256) char scratch[RM_STATIC_SCRATCH_BYTES];

static constexpr int STATIC_MAX_DEVICES = 64;
static std::mutex g_static_mu;
static char* g_static_ptr[STATIC_MAX_DEVICES] = {};
static int g_static_cu[STATIC_MAX_DEVICES] = {};

// The current device's copy of `scratch` and its CU count (looked up once per device).
static int static_scratch(char** ptr, int* n_cu) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= STATIC_MAX_DEVICES) return fail(RM_ERR_ARG, "rm_batch_propagate_static: device %d", dev);
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(RM_ERR_HIP, "rm_batch_propagate_static: no device address for the scratch array");
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) cu = 256;
    g_static_cu[dev] = cu;
    g_static_ptr[dev] = (char*)p;
  }
  *ptr = g_static_ptr[dev];
  *n_cu = g_static_cu[dev];
  return RM_OK;
}

// sub_run for the wave classes with everything in [dev, dev + bytes): the
// descriptors first, the Krylov workspace after them; the classes run one
// after another on s, each with as many waves as the workspace holds.
static int sub_run_static(char* dev, size_t bytes, const std::vector<SubOp>& ops, std::vector<int4> (&cls)[NCLASS],
                   int32_t kcap, int32_t fixed_m, int32_t maxorder, double tol, bool f32, int n_cu, hipStream_t s) {
  group_pairs(cls, fixed_m);
  size_t nitems = 0;
  for (int c = 0; c < NCLASS; ++c) nitems += cls[c].size();
  if (nitems == 0) return RM_OK;
  const size_t ops_bytes = (ops.size() * sizeof(SubOp) + 255) & ~(size_t)255;
  const size_t stage = (ops_bytes + nitems * sizeof(int4) + 255) & ~(size_t)255;
  if (stage > STATIC_STAGE_BYTES) return fail(RM_ERR_ARG, "rm_batch_propagate_static: %zu descriptor bytes", stage);
  static thread_local std::vector<char> host;   // host side of the staging (pageable)
  if (host.size() < stage) host.resize(STATIC_STAGE_BYTES);
  memcpy(host.data(), ops.data(), ops.size() * sizeof(SubOp));
  size_t off[NCLASS];
  size_t pos = ops_bytes;
  for (int c = 0; c < NCLASS; ++c) {
    off[c] = pos;
    if (!cls[c].empty()) memcpy(host.data() + pos, cls[c].data(), cls[c].size() * sizeof(int4));
    pos += cls[c].size() * sizeof(int4);
  }
  HIPCHK(hipMemcpyAsync(dev, host.data(), stage, hipMemcpyHostToDevice, s));
  char* ws = dev + stage;
  const size_t ws_bytes = bytes - stage;
  SubArgs a;
  memset(&a, 0, sizeof(a));
  a.ops = (const SubOp*)dev;
  a.kcap = kcap;
  a.fixed_m = fixed_m;
  a.maxorder = maxorder;
  a.tol = tol;
  for (int c = 0; c < NCLASS; ++c) {
    if (cls[c].empty()) continue;
    a.items = (const int4*)(dev + off[c]);
    a.nitems = (int32_t)cls[c].size();
    const bool bl = class_basis_lds(c, kcap);
    const size_t per_wave = sizeof(double2) * ((bl ? 0 : (size_t)kcap * 64 * CLASS_R[c]) +
                                               (fixed_m > 0 ? 0 : (size_t)kcap * kcap + 2 * (size_t)kcap));
    a.max_waves = 0;
    if (per_wave > 0) {
      const size_t cap = ws_bytes / per_wave;
      if (cap < 1)
        return fail(RM_ERR_ARG, "rm_batch_propagate_static: one wave's Krylov workspace (%zu bytes) exceeds "
                    "the %zu bytes left in the scratch array", per_wave, ws_bytes);
      a.max_waves = (int32_t)std::min<size_t>(cap, INT32_MAX);
    }
    int waves = 0;
    HIPCHK(launch_class(c, a, f32, n_cu, s, &waves, true));
    a.basis = (double2*)ws;
    a.gram = a.basis + (bl ? 0 : (size_t)waves * kcap * 64 * CLASS_R[c]);
    a.cscr = a.gram + (fixed_m > 0 ? 0 : (size_t)waves * kcap * kcap);
    HIPCHK(launch_class(c, a, f32, n_cu, s, &waves, false));
  }
  return RM_OK;
}

extern "C" int rm_batch_propagate_static(const rm_job* jobs, int32_t njobs, int32_t job_bytes, int32_t krylov_fixed,
                                         int32_t maxorder, double tol, int32_t precision, void* stream) {
  int rc = check_krylov("rm_batch_propagate_static", krylov_fixed, maxorder, tol, precision);
  if (rc) return rc;
  if (njobs > RM_STATIC_MAX_JOBS)
    return fail(RM_ERR_ARG, "rm_batch_propagate_static: %d jobs exceed the maximum of %d", njobs, RM_STATIC_MAX_JOBS);
  char* dev = nullptr;
  int n_cu = 0;
  if ((rc = static_scratch(&dev, &n_cu))) return rc;
  std::vector<SubOp> ops;
  std::vector<int4> cls[NCLASS];
  std::vector<LongItem> blk;
  if ((rc = collect_jobs("rm_batch_propagate_static", jobs, njobs, job_bytes, RM_SUBSPACE_MAX_ROWS, true, ops, cls,
                         blk)))
    return rc;
  return sub_run_static(dev, RM_STATIC_SCRATCH_BYTES, ops, cls, krylov_fixed > 0 ? krylov_fixed : maxorder,
                        krylov_fixed, maxorder, tol, precision == RM_PREC_FP32, n_cu, (hipStream_t)stream);
}
