// richmol_amd: cooperative propagation of wavepackets whose domain (the union
// of invariant subspaces their support touches) is longer than one wave can
// hold.
//
// The whole grid works on one wavepacket at a time, through every step of the
// run, in one cooperative launch (all workgroups resident).  Each step forms
// the step's operator A(F) = sum_s F_s B_s on the domain's entries (permuted
// CSR) -- skipped when the step's field products equal the previous step's --
// then runs the Lanczos expv of tdse.py:417-486 with the matvec spread over
// every wave (LPR lanes per row) and the vector work over every thread.  Dot
// products and norms are two-level sums (wave DPP, workgroup, then the
// workgroups' partials added in a fixed order by every workgroup, so all of
// them take the same decisions).  A Lanczos iteration needs two grid
// barriers: after the matvec (alpha) and after the three-term update (beta);
// the basis vector V_k = W/beta is formed inside the matvec from W, so the
// beta barrier also publishes the matvec input.
//
// When every wave holds at most one group of RPW rows and a row has at most
// LPR*UR entries, a lane keeps the columns and the step's values of its
// entries in registers for the whole step (the matvec is then gathers and
// FMAs only); otherwise the entries are read at every matvec.

#include <algorithm>

#include "rm_common.h"

namespace {

// threads per workgroup: 512 with the operator in registers (256 VGPRs per
// lane), 1024 reading it from memory
template <bool REG>
constexpr int ct_of() { return REG ? 512 : 1024; }
constexpr int LPR = 16;          // lanes per row in the matvec (one DPP row)
constexpr int RPW = 64 / LPR;    // rows per wave in the matvec
constexpr int UR = 12;           // register-resident entries per lane
constexpr int UM = 8;            // entries per lane loaded together (memory mode)
constexpr int GL = 32;           // Gram matrix order held in LDS (adaptive)
constexpr unsigned SPIN_LIMIT = 1u << 22;   // polls of a grid barrier before the launch gives up

struct CoopArgs {
  SubOp op;
  int4 item;               // (column, lo, len, -): the launch's wavepacket
  double2* V;              // [kcap][lmax] Krylov basis (domain rows)
  double2* Wb;             // [lmax] W (the next matvec input is W / beta)
  float2* Wf;              // [lmax] W rounded to FP32 (FP32 mode)
  double2* Y;              // [lmax] matvec output
  double2* S;              // [lmax] the wavepacket on the domain (permuted order)
  double2* P;              // [lmax] its split-operator phases
  void* Ast;               // [emax] the step's operator entries (memory mode)
  double* part;            // [2][grid][4] workgroup partial sums, double-buffered
  double2* gram;           // [grid][kcap * kcap] (adaptive, orders beyond GL)
  double2* cscr;           // [grid][2 * kcap] (adaptive, orders beyond GL)
  unsigned long long* bar; // grid-barrier arrivals (zero at launch)
  int32_t* abort_flag;     // set when a barrier wait exceeds SPIN_LIMIT polls
  int64_t lmax;
  int32_t kcap, fixed_m, maxorder;
  double tol;
};

struct GridSync {
  unsigned long long target = 0;
  int epoch = 0;
};

// Grid barrier on an arrival counter that only grows: barrier number t
// completes when the counter reaches t * grid.  Every wave first waits for
// its own memory operations; one thread per workgroup then publishes the
// workgroup's writes (agent-scope release), arrives, waits and acquires
// (agent-scope acquire; the waves of a workgroup share the CU's caches).  A
// wait that exceeds SPIN_LIMIT polls (a workgroup that never arrives) sets
// the abort flag; from then on barriers no longer wait and the launch runs
// to its end, its results marked invalid.
__device__ __forceinline__ void grid_barrier(const CoopArgs& a, GridSync& gs, int* s_ab) {
  gs.target += gridDim.x;
  gs.epoch ^= 1;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0 && *s_ab == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(a.bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gs.target) {
      __builtin_amdgcn_s_sleep(1);
      ++spins;
      if ((spins & 255u) == 0 &&
          (spins >= SPIN_LIMIT || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_ab = 1;
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Sum of NQ values over every thread of the grid; every workgroup returns the
// same sums (the partials are added in workgroup order by each of them).
template <int NQ, int CWV>
__device__ void grid_sum(double (&v)[NQ], const CoopArgs& a, GridSync& gs, double (*s_red)[4],
                         double* s_tot, int* s_ab) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_allsum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NQ; ++q) s_red[wv][q] = v[q];
  __syncthreads();
  double* part = a.part + (size_t)gs.epoch * gridDim.x * 4;
  if (tid < NQ) {
    double t = 0.0;
    for (int w = 0; w < CWV; ++w) t += s_red[w][tid];
    part[blockIdx.x * 4 + tid] = t;
  }
  grid_barrier(a, gs, s_ab);
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

// REG: operator entries in registers (one row group per wave, <= UR entries
// per lane); otherwise read from the step's operator array at every matvec.
template <bool FIXED, bool F32, bool REG>
__global__ __launch_bounds__(ct_of<REG>()) void k_coop(CoopArgs a) {
  using XT = typename XType<F32>::T;
  constexpr int CT = ct_of<REG>();
  constexpr int CWV = CT / 64;
  __shared__ double2 s_alpha[KMAX];
  __shared__ double s_beta[KMAX + 1];
  __shared__ double2 s_cf[KMAX];
  __shared__ double2 s_gram[FIXED ? 1 : GL * GL];
  __shared__ double2 s_cv[FIXED ? 1 : 2 * GL];
  __shared__ double s_red[CWV][4];
  __shared__ double s_tot[4];
  __shared__ double s_conv;
  __shared__ int s_ab;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t nthr = (int64_t)gridDim.x * CT;
  const int64_t t0 = (int64_t)blockIdx.x * CT + tid;
  const int nwaves = gridDim.x * CWV, gw = blockIdx.x * CWV + wv;
  const int sub = lane % LPR, rw = lane / LPR;
  const int kc = a.kcap;
  const SubOp& op = a.op;
  XT* Ast = (XT*)a.Ast;
  const XT* Xin;
  if constexpr (F32)
    Xin = a.Wf;
  else
    Xin = a.Wb;
  const double2 zero = make_double2(0.0, 0.0);
  GridSync gs;
  if (tid == 0) s_ab = 0;
  __syncthreads();

  const int b = a.item.x, lo = a.item.y, len = a.item.z;
  const int64_t e_lo = op.rp[lo];
  const int ngroups = (len + RPW - 1) / RPW;
  for (int64_t i = t0; i < len; i += nthr) {
    const int orig = op.perm[lo + i];
    a.S[i] = op.psi[(int64_t)b * op.ld + orig];
    a.P[i] = op.phase ? op.phase[orig] : make_double2(1.0, 0.0);
  }
  // this lane's row in the matvec (REG: the wave's only row group) and its entries
  const int rreg = gw * RPW + rw;
  const bool rreg_ok = REG && rreg < len;
  int64_t ebeg = 0, eend = 0;
  if (rreg_ok) {
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
  int built = -1;   // step whose operator the registers / Ast hold

  for (int step = 0; step < op.nsteps; ++step) {
    if (op.sflag[step] == 0) {
      // no field: the expv is skipped (tdse.py:377), phases only
      if (op.phase)
        for (int64_t i = t0; i < len; i += nthr) a.S[i] = cmul(cmul(a.S[i], a.P[i]), a.P[i]);
      if (t0 == 0 && op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = 0;
      continue;
    }
    const double2* fp = op.table + (int64_t)step * op.nslot;
    // the step's operator on this lane's entries, unless the field products
    // are those of the operator already built
    bool same = built >= 0;
    if (same) {
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
    // V_0 = phi = P psi (unnormalised, tdse.py:443-446), the first matvec input
    for (int64_t i = t0; i < len; i += nthr) {
      const double2 v = op.phase ? cmul(a.S[i], a.P[i]) : a.S[i];
      a.V[i] = v;
      a.Wb[i] = v;
      if constexpr (F32) a.Wf[i] = to_x<true>(v);
    }
    grid_barrier(a, gs, &s_ab);

    // Gram matrix and coefficient scratch: LDS up to order GL, then global
    double2* Gm = s_gram;
    double2* cv = s_cv;
    int gs_ld = GL;
    int k = 1;
    double conv = 1.0, beta = 0.0, ib = 1.0;
    bool brk = false;
    for (int it = 0;; ++it) {
      double2* Vit = a.V + (int64_t)it * a.lmax;
      // w = A (W / beta) = A V_it on this wave's rows; alpha = vdot(w, V_it)
      double red[3] = {0.0, 0.0, 0.0};
      auto finish_row = [&](int r, XT acc) __attribute__((always_inline)) {
        double2 w = to_d2(acc);
        w.x = row_sum16(w.x);
        w.y = row_sum16(w.y);
        if (sub == 0 && r < len) {
          w = cscale(ib, w);
          const double2 v = cscale(ib, a.Wb[r]);
          Vit[r] = v;
          a.Y[r] = w;
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
        for (int rg = gw; rg < ngroups; rg += nwaves) {
          const int r = rg * RPW + rw;
          XT acc = to_x<F32>(zero);
          if (r < len) {
            const int64_t e1 = op.rp[lo + r + 1];
            for (int64_t e = op.rp[lo + r] + sub; e < e1; e += (int64_t)LPR * UM) {
              int c[UM];
              XT av[UM];
#pragma unroll
              for (int u = 0; u < UM; ++u) {
                const int64_t q = e + (int64_t)u * LPR;
                const bool ok = q < e1;
                c[u] = ok ? op.ccol[q] - lo : 0;
                av[u] = ok ? Ast[q - e_lo] : to_x<F32>(zero);
              }
#pragma unroll
              for (int u = 0; u < UM; ++u) acc = xfma(av[u], Xin[c[u]], acc);
            }
          }
          finish_row(r, acc);
        }
      }
      if (it == 0 && !FIXED) {
        grid_sum<3, CWV>(red, a, gs, s_red, s_tot, &s_ab);
      } else {
        double r2[2] = {red[0], red[1]};
        grid_sum<2, CWV>(r2, a, gs, s_red, s_tot, &s_ab);
        red[0] = r2[0];
        red[1] = r2[1];
      }
      const double2 al = make_double2(red[0], red[1]);
      if (tid < 64) s_alpha[it] = al;
      // W = w - alpha V_it - beta V_{it-1}, and |W|^2
      const double2* Vprev = a.V + (int64_t)(it > 0 ? it - 1 : 0) * a.lmax;
      double bs = 0.0;
      for (int64_t i = t0; i < len; i += nthr) {
        double2 wn = csub(a.Y[i], cmul(al, Vit[i]));
        if (it > 0) wn = csub(wn, cscale(beta, Vprev[i]));
        a.Wb[i] = wn;
        if constexpr (F32) a.Wf[i] = to_x<true>(wn);
        bs += wn.x * wn.x + wn.y * wn.y;
      }
      k = it + 1;
      if (!FIXED) {
        if (it == 0) {
          if (tid == 0) {
            Gm[0] = make_double2(red[2], 0.0);
            cv[gs_ld] = make_double2(1.0, 0.0);   // c_prev = [1]
          }
        } else if (wv == 0) {
          wave_sync();
          const double c = adaptive_metric<true>(it, brk, Gm, gs_ld, s_alpha, s_beta, op.f, cv, cv + gs_ld);
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
        gs_ld = kc;
        __threadfence_block();
        __syncthreads();
      }
      // beta_k = |W| (tdse.py:453)
      double bsum[1] = {bs};
      grid_sum<1, CWV>(bsum, a, gs, s_red, s_tot, &s_ab);
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
      for (int64_t i = t0; i < len; i += nthr) a.Wb[i] = make_double2(1.0, 0.0);
      for (int j = 0; j < k; ++j) {
        const double2* Vj = a.V + (int64_t)j * a.lmax;
        double d[2] = {0.0, 0.0};
        for (int64_t i = t0; i < len; i += nthr) {
          const double2 x = Vj[i], y = a.Wb[i];
          d[0] += x.x * y.x + x.y * y.y;
          d[1] += x.x * y.y - x.y * y.x;
        }
        grid_sum<2, CWV>(d, a, gs, s_red, s_tot, &s_ab);
        const double2 proj = make_double2(d[0], d[1]);
        for (int64_t i = t0; i < len; i += nthr) a.Wb[i] = csub(a.Wb[i], cmul(proj, Vj[i]));
      }
      double nv[1] = {0.0};
      for (int64_t i = t0; i < len; i += nthr) nv[0] += a.Wb[i].x * a.Wb[i].x + a.Wb[i].y * a.Wb[i].y;
      grid_sum<1, CWV>(nv, a, gs, s_red, s_tot, &s_ab);
      const double nrm = sqrt(nv[0]);
      // nrm == 0: the basis spans the domain; the reference's restart vector
      // lies outside it and enters with coefficient exactly zero
      double2* Vk = a.V + (int64_t)k * a.lmax;
      for (int64_t i = t0; i < len; i += nthr) {
        const double2 v = nrm > 0.0 ? make_double2(a.Wb[i].x / nrm, a.Wb[i].y / nrm) : zero;
        a.Wb[i] = v;
        Vk[i] = v;
        if constexpr (F32) a.Wf[i] = to_x<true>(v);
      }
      if (!FIXED) {
        // Gram column of the restart vector, explicit dots
        for (int j = 0; j < k; ++j) {
          const double2* Vj = a.V + (int64_t)j * a.lmax;
          double d[2] = {0.0, 0.0};
          for (int64_t i = t0; i < len; i += nthr) {
            const double2 x = Vj[i], y = Vk[i];
            d[0] += x.x * y.x + x.y * y.y;
            d[1] += x.x * y.y - x.y * y.x;
          }
          grid_sum<2, CWV>(d, a, gs, s_red, s_tot, &s_ab);
          if (tid == 0) {
            Gm[j * gs_ld + k] = make_double2(d[0], d[1]);
            Gm[k * gs_ld + j] = make_double2(d[0], -d[1]);
          }
        }
      } else {
        grid_barrier(a, gs, &s_ab);
      }
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
    for (int64_t i = t0; i < len; i += nthr) {
      double2 u = zero;
      for (int j = 0; j < k; ++j) u = cfma(s_cf[j], a.V[(int64_t)j * a.lmax + i], u);
      a.S[i] = op.phase ? cmul(u, a.P[i]) : u;
    }
    if (t0 == 0) {
      if (op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = k;
      if (!FIXED && k >= a.maxorder && op.status) op.status[b - op.kcol0] |= RM_NOT_CONVERGED;
    }
    // the next step's V_0 is written by the same threads; s_cf and the LDS
    // Gram matrix are rewritten only after the next step's first grid barrier
  }
  for (int64_t i = t0; i < len; i += nthr) op.psi[(int64_t)b * op.ld + op.perm[lo + i]] = a.S[i];
  if (t0 == 0 && op.status &&
      (s_ab || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
    op.status[b - op.kcol0] |= RM_DEVICE_TIMEOUT;
}

__global__ void k_coop_reset(unsigned long long* bar, int32_t* abort_flag) {
  if (threadIdx.x == 0) {
    *bar = 0ull;
    *abort_flag = 0;
  }
}

struct CoopScratch {
  void* ws = nullptr;
  size_t bytes = 0;
};

template <bool FIXED, bool F32, bool REG>
hipError_t coop_launch_r(CoopArgs& a, int grid, hipStream_t s) {
  void* args[] = {&a};
  return hipLaunchCooperativeKernel((const void*)k_coop<FIXED, F32, REG>, dim3(grid), dim3(ct_of<REG>()), args,
                                    0, s);
}

hipError_t coop_launch(CoopArgs& a, int grid, bool reg, hipStream_t s) {
  const bool fixed = a.fixed_m > 0, f32 = a.Wf != nullptr;
  if (reg) {
    if (fixed) return f32 ? coop_launch_r<true, true, true>(a, grid, s) : coop_launch_r<true, false, true>(a, grid, s);
    return f32 ? coop_launch_r<false, true, true>(a, grid, s) : coop_launch_r<false, false, true>(a, grid, s);
  }
  if (fixed) return f32 ? coop_launch_r<true, true, false>(a, grid, s) : coop_launch_r<true, false, false>(a, grid, s);
  return f32 ? coop_launch_r<false, true, false>(a, grid, s) : coop_launch_r<false, false, false>(a, grid, s);
}

// Workgroups the device holds at once for every instantiation of one mode (the smallest).
template <bool REG>
hipError_t coop_max_grid(int n_cu, int* grid) {
  const void* ks[] = {(const void*)k_coop<true, true, REG>, (const void*)k_coop<true, false, REG>,
                      (const void*)k_coop<false, true, REG>, (const void*)k_coop<false, false, REG>};
  int occ = INT32_MAX;
  for (const void* k : ks) {
    int o = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k, ct_of<REG>(), 0);
    if (e != hipSuccess) return e;
    occ = std::min(occ, o);
  }
  *grid = std::min(std::max(1, occ) * n_cu, n_cu);
  return hipSuccess;
}

}  // namespace

void coop_scratch_release(void* sc) {
  CoopScratch* c = (CoopScratch*)sc;
  if (!c) return;
  if (c->ws) (void)hipFree(c->ws);
  delete c;
}

// Propagate the wavepackets `items` (all of one plan / job op) with the
// cooperative kernel, one launch each, on stream s.  A launch's grid is one
// wave per RPW rows of the matvec (at most one workgroup per CU): a function
// of the wavepacket's domain and the device only, so its result does not
// depend on the other wavepackets.
int coop_run(rm_plan* p, void** scratch, const SubOp& op, const std::vector<int4>& items, int32_t kcap,
             int32_t fixed_m, int32_t maxorder, double tol, bool f32, int n_cu, hipStream_t s) {
  if (items.empty()) return RM_OK;
  if (!op.rp) return fail(RM_ERR_ARG, "cooperative propagation: the plan has no CSR subspace operator");
  int gmax_r = 1, gmax_m = 1;
  HIPCHK(coop_max_grid<true>(n_cu, &gmax_r));
  HIPCHK(coop_max_grid<false>(n_cu, &gmax_m));
  const int gmax = std::max(gmax_r, gmax_m);
  int64_t lmax = 0, emax = 0;
  for (const int4& it : items) {
    lmax = std::max<int64_t>(lmax, it.z);
    emax = std::max<int64_t>(emax, p->h_sub_rp[it.y + it.z] - p->h_sub_rp[it.y]);
  }
  const bool fixed = fixed_m > 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_V = al(sizeof(double2) * (size_t)kcap * lmax);
  const size_t b_vec = al(sizeof(double2) * (size_t)lmax);
  const size_t b_f = f32 ? al(sizeof(float2) * (size_t)lmax) : 0;
  const size_t b_A = al(sizeof(double2) * (size_t)std::max<int64_t>(emax, 1));
  const size_t b_part = al(sizeof(double) * 2 * 4 * (size_t)gmax);
  const size_t b_gram = fixed ? 0 : al(sizeof(double2) * (size_t)gmax * kcap * kcap);
  const size_t b_c = fixed ? 0 : al(sizeof(double2) * (size_t)gmax * 2 * kcap);
  const size_t b_bar = 256;
  const size_t need = b_V + 4 * b_vec + b_f + b_A + b_part + b_gram + b_c + b_bar;
  CoopScratch* sc = (CoopScratch*)*scratch;
  if (!sc) {
    sc = new (std::nothrow) CoopScratch();
    if (!sc) return fail(RM_ERR_ALLOC, "cooperative propagation: host allocation");
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
  CoopArgs a;
  memset(&a, 0, sizeof(a));
  a.op = op;
  a.V = (double2*)w; w += b_V;
  a.Wb = (double2*)w; w += b_vec;
  a.Y = (double2*)w; w += b_vec;
  a.S = (double2*)w; w += b_vec;
  a.P = (double2*)w; w += b_vec;
  a.Wf = f32 ? (float2*)w : nullptr; w += b_f;
  a.Ast = w; w += b_A;
  a.part = (double*)w; w += b_part;
  a.gram = fixed ? nullptr : (double2*)w; w += b_gram;
  a.cscr = fixed ? nullptr : (double2*)w; w += b_c;
  a.bar = (unsigned long long*)w;
  a.abort_flag = (int32_t*)(w + 64);
  a.lmax = lmax;
  a.kcap = kcap;
  a.fixed_m = fixed_m;
  a.maxorder = maxorder;
  a.tol = tol;
  for (const int4& it : items) {
    a.item = it;
    const int64_t groups = (it.z + RPW - 1) / RPW;
    // registers hold a lane's entries when every wave has one row group and
    // no row of the domain is longer than LPR * UR entries
    constexpr int wr = ct_of<true>() / 64, wm = ct_of<false>() / 64;
    bool reg = groups <= (int64_t)gmax_r * wr;
    for (int64_t r = it.y; reg && r < (int64_t)it.y + it.z; ++r)
      reg = p->h_sub_rp[r + 1] - p->h_sub_rp[r] <= (int64_t)LPR * UR;
    const int w = reg ? wr : wm;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((groups + w - 1) / w, reg ? gmax_r : gmax_m));
    hipLaunchKernelGGL(k_coop_reset, dim3(1), dim3(64), 0, s, a.bar, a.abort_flag);
    HIPCHK(hipGetLastError());
    HIPCHK(coop_launch(a, g, reg, s));
  }
  return RM_OK;
}
