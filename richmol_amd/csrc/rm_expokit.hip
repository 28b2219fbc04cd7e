// richmol_amd: the reference's external propagator on the device -- Expokit's
// ZHEXPV (expokit/expokit.f:2783-3170) with its small exponentials by ZGPADM
// (expokit.f:763-933), as richmol/tdse.py:383-391 and 404-412 call it:
// Krylov size m (12 there), t = Im(f), the operator A = iH(F), tol, and the
// 1-norm of the whole matrix H(F) for the first step size.
//
// One workgroup per wavepacket, on its domain (the union of invariant
// subspaces its support touches: outside it every Krylov vector is zero, as
// in the Lanczos kernels).  A launch runs every step of a run: per step the
// split-operator phases, then ZHEXPV's integration -- sub-steps of variable
// length, each a Lanczos recurrence of m matvecs (vectors in an HBM region of
// the workgroup), the corrected (m + 2) x (m + 2) matrix, its exponential by
// the (6, 6) Pade approximation with scaling and squaring in LDS (products
// and the pivoted LU solve spread over the workgroup), the local error
// estimate, rejection and the next step size -- all decided identically by
// every thread (sums in a fixed order).  The whole-matrix 1-norm of each
// step comes from k_expo_norm ahead of the launch.

#include <climits>

#include "rm_common.h"

namespace {

constexpr int ET = 256;                    // threads per workgroup
constexpr int EW = ET / 64;
constexpr int MH = RM_EXPOKIT_MAX_M + 2;   // order of the corrected matrix
// expokit.f:2882-2887
constexpr int MXSTEP = 500;
constexpr int IDEG = 6;
constexpr double DELTA = 1.2;
constexpr double GAMMA = 0.9;
constexpr double BREAK_TOL = 1.0e-7;

struct ExpoArgs {
  SubOp op;
  const int4* items;      // (column, lo, len, -)
  int32_t nitems;
  double2* ws;            // per workgroup (m + 3) * lmax
  int64_t lmax;
  const double* anorm;    // [nsteps] 1-norm of H(F_s)
  int32_t m;
  double tol;
};

// Sums of two values over the workgroup, every thread receiving them (waves'
// partials added in wave order).
__device__ __forceinline__ void bsum2(double& a, double& b, double (*red)[EW]) {
  a = wave_allsum(a);
  b = wave_allsum(b);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wv] = a;
    red[1][wv] = b;
  }
  __syncthreads();
  a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
  b = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  __syncthreads();
}

__device__ __forceinline__ double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }
__device__ __forceinline__ double2 crecip(double2 z) {
  const double d = z.x * z.x + z.y * z.y;
  return make_double2(z.x / d, -z.y / d);
}

// C = alpha (A B) for mx x mx matrices in LDS (row-major, leading dimension
// MH); every thread of the workgroup, barrier at the end.
__device__ void lds_gemm(const double2* A, const double2* B, double2* C, int mx, double alpha) {
  for (int idx = threadIdx.x; idx < mx * mx; idx += ET) {
    const int i = idx / mx, j = idx % mx;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < mx; ++k) acc = cfma(A[i * MH + k], B[k * MH + j], acc);
    C[i * MH + j] = cscale(alpha, acc);
  }
  __syncthreads();
}

// Round a step size to one or two significant digits (expokit.f:2964-2965).
__device__ double round_step(double t) {
  if (!isfinite(t)) return t;
  const double p1 = pow(10.0, (double)((long long)llround(log10(t) - sqrt(0.1)) - 1));
  return trunc(t / p1 + 0.55) * p1;
}

// exp(t Hs) of the leading mx x mx block of Hs (ZGPADM, expokit.f:818-932):
// returns the LDS matrix that holds it (one of W0..W3); ns: squarings.
__device__ double2* zgpadm(const double2* Hs, int mx, double t, double2* W0, double2* W1, double2* W2,
                           double2* W3, double* s_sc, int* s_piv) {
  const int tid = threadIdx.x;
  // scaling: ||t Hs / 2^ns||_inf < 1/2 (row sums, expokit.f:833-846)
  if (tid == 0) {
    double hn = 0.0;
    for (int i = 0; i < mx; ++i) {
      double r = 0.0;
      for (int j = 0; j < mx; ++j) r += hypot(Hs[i * MH + j].x, Hs[i * MH + j].y);
      hn = fmax(hn, r);
    }
    hn = fabs(t * hn);
    const int ns = hn > 0.0 ? max(0, (int)(log(hn) / log(2.0)) + 2) : 0;
    s_sc[0] = t / ldexp(1.0, ns);
    s_sc[1] = (double)ns;
  }
  __syncthreads();
  const double scale = s_sc[0];
  const int ns = (int)s_sc[1];
  double coef[IDEG + 1];
  coef[0] = 1.0;
  for (int k = 1; k <= IDEG; ++k)
    coef[k] = coef[k - 1] * (double)(IDEG + 1 - k) / (double)(k * (2 * IDEG + 1 - k));
  // H2 = scale^2 Hs Hs (W0); p (W1) and q (W2) start as coef[5] I, coef[6] I
  lds_gemm(Hs, Hs, W0, mx, scale * scale);
  for (int idx = tid; idx < mx * mx; idx += ET) {
    const int i = idx / mx, j = idx % mx;
    W1[i * MH + j] = make_double2(i == j ? coef[IDEG - 1] : 0.0, 0.0);
    W2[i * MH + j] = make_double2(i == j ? coef[IDEG] : 0.0, 0.0);
  }
  __syncthreads();
  // Horner (expokit.f:875-889): odd passes q = q H2 + c I, even p = p H2 + c I
  double2 *p = W1, *q = W2, *fr = W3;
  bool odd = true;
  for (int k = IDEG - 1; k > 0; --k) {
    double2* used = odd ? q : p;
    lds_gemm(used, W0, fr, mx, 1.0);
    for (int i = tid; i < mx; i += ET) fr[i * MH + i].x += coef[k - 1];
    __syncthreads();
    if (odd) {
      q = fr;
    } else {
      p = fr;
    }
    fr = used;
    odd = !odd;
  }
  // the odd polynomial times scale Hs (expokit.f:893-901)
  if (odd) {
    lds_gemm(q, Hs, fr, mx, scale);
    double2* t2 = q;
    q = fr;
    fr = t2;
  } else {
    lds_gemm(p, Hs, fr, mx, scale);
    double2* t2 = p;
    p = fr;
    fr = t2;
  }
  // q = q - p; solve q X = p by LU with partial pivoting (|re| + |im|, as
  // LAPACK's izamax); X in p (expokit.f:902-904)
  for (int idx = tid; idx < mx * mx; idx += ET) {
    const int i = idx / mx, j = idx % mx;
    q[i * MH + j] = csub(q[i * MH + j], p[i * MH + j]);
  }
  __syncthreads();
  for (int k = 0; k < mx; ++k) {
    if (tid == 0) {
      int piv = k;
      double best = cabs1(q[k * MH + k]);
      for (int i = k + 1; i < mx; ++i) {
        const double v = cabs1(q[i * MH + k]);
        if (v > best) {
          best = v;
          piv = i;
        }
      }
      *s_piv = piv;
    }
    __syncthreads();
    const int piv = *s_piv;
    if (piv != k)
      for (int j = tid; j < mx; j += ET) {
        const double2 a = q[k * MH + j], b = p[k * MH + j];
        q[k * MH + j] = q[piv * MH + j];
        q[piv * MH + j] = a;
        p[k * MH + j] = p[piv * MH + j];
        p[piv * MH + j] = b;
      }
    __syncthreads();
    const double2 r = crecip(q[k * MH + k]);
    for (int i = k + 1 + tid; i < mx; i += ET) q[i * MH + k] = cmul(q[i * MH + k], r);
    __syncthreads();
    for (int idx = tid; idx < (mx - k - 1) * (2 * mx - k - 1); idx += ET) {
      const int w = 2 * mx - k - 1;
      const int i = k + 1 + idx / w, c = idx % w;
      const double2 l = q[i * MH + k];
      if (c < mx - k - 1) {
        const int j = k + 1 + c;
        q[i * MH + j] = csub(q[i * MH + j], cmul(l, q[k * MH + j]));
      } else {
        const int j = c - (mx - k - 1);
        p[i * MH + j] = csub(p[i * MH + j], cmul(l, p[k * MH + j]));
      }
    }
    __syncthreads();
  }
  for (int k = mx - 1; k >= 0; --k) {
    const double2 r = crecip(q[k * MH + k]);
    for (int j = tid; j < mx; j += ET) p[k * MH + j] = cmul(p[k * MH + j], r);
    __syncthreads();
    for (int idx = tid; idx < k * mx; idx += ET) {
      const int i = idx / mx, j = idx % mx;
      p[i * MH + j] = csub(p[i * MH + j], cmul(q[i * MH + k], p[k * MH + j]));
    }
    __syncthreads();
  }
  // (+/-)(I + 2 X), then ns squarings (expokit.f:905-929)
  const double sg = (ns == 0 && odd) ? -1.0 : 1.0;
  for (int idx = tid; idx < mx * mx; idx += ET) {
    const int i = idx / mx, j = idx % mx;
    double2 v = cscale(2.0, p[i * MH + j]);
    if (i == j) v.x += 1.0;
    p[i * MH + j] = cscale(sg, v);
  }
  __syncthreads();
  double2* cur = p;
  double2* nxt = q;
  for (int k = 0; k < ns; ++k) {
    lds_gemm(cur, cur, nxt, mx, 1.0);
    double2* t2 = cur;
    cur = nxt;
    nxt = t2;
  }
  return cur;
}

// y = i H(F) x on the domain rows (permuted CSR of the B_s, columns relative
// to lo); returns nothing, y written by every thread for its rows.
__device__ __forceinline__ void expo_matvec(const SubOp& op, const double2* fp, int lo, int len,
                                            const double2* x, double2* y) {
  for (int i = threadIdx.x; i < len; i += ET) {
    const int64_t e0 = op.rp[lo + i], e1 = op.rp[lo + i + 1];
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t e = e0; e < e1; ++e) {
      double2 a = make_double2(0.0, 0.0);
      for (int s = 0; s < op.nslot; ++s) a = cfma(fp[s], op.cval[(int64_t)s * op.nnz + e], a);
      acc = cfma(a, x[op.ccol[e] - lo], acc);
    }
    y[i] = make_double2(-acc.y, acc.x);
  }
}

__global__ __launch_bounds__(ET) void k_expokit(ExpoArgs a) {
  __shared__ double2 s_H[MH * MH];
  __shared__ double2 s_W[4][MH * MH];
  __shared__ double2 s_e1[MH];
  __shared__ double s_red[2][EW];
  __shared__ double s_sc[2];
  __shared__ int s_piv;
  const int tid = threadIdx.x;
  const SubOp& op = a.op;
  const int m = a.m, mh = m + 2;
  double2* base = a.ws + (int64_t)blockIdx.x * (m + 3) * a.lmax;
  double2* V = base;                        // V_0 .. V_m
  double2* P = base + (int64_t)(m + 1) * a.lmax;
  double2* S = base + (int64_t)(m + 2) * a.lmax;
  const double2 zero = make_double2(0.0, 0.0);
  for (int itm = blockIdx.x; itm < a.nitems; itm += gridDim.x) {
    const int4 item = a.items[itm];
    const int b = item.x, lo = item.y, len = item.z;
    for (int i = tid; i < len; i += ET) S[i] = op.psi[(int64_t)b * op.ld + op.perm[lo + i]];
    __syncthreads();
    int32_t fail = 0;
    for (int step = 0; step < op.nsteps; ++step) {
      if (op.phase)
        for (int i = tid; i < len; i += ET) S[i] = cmul(S[i], op.phase[op.perm[lo + i]]);
      int nmult = 0;
      if (step_on(op, step)) {
        const double2* fp = op.table + (int64_t)step * op.nslot;
        const double anorm = a.anorm[step];
        const double eps = 2.220446049250313e-16;   // expokit.f:2944-2948 on IEEE doubles
        const double tol = a.tol <= eps ? sqrt(eps) : a.tol;
        const double t = op.f.y;                     // Im(f) (tdse.py:388)
        const double t_out = fabs(t), sgn = t < 0.0 ? -1.0 : 1.0;
        double r0 = 0.0, r1 = 0.0;
        for (int i = tid; i < len; i += ET) r0 += S[i].x * S[i].x + S[i].y * S[i].y;
        bsum2(r0, r1, s_red);
        double beta = sqrt(r0);
        // the first step size (expokit.f:2960-2965)
        double xm = 1.0 / (double)m;
        const double p2i = tol * pow((m + 1) / 2.72, (double)(m + 1)) * sqrt(2.0 * 3.14 * (m + 1));
        double t_new = round_step((1.0 / anorm) * pow(p2i / (4.0 * beta * anorm), xm));
        double t_now = 0.0, avnorm = 0.0;
        int k1 = 2, mbrkdwn = m, nstep = 0;
        while (t_now < t_out) {
          ++nstep;
          double t_step = fmin(t_out - t_now, t_new);
          r0 = 0.0;
          r1 = 0.0;
          for (int i = tid; i < len; i += ET) r0 += S[i].x * S[i].x + S[i].y * S[i].y;
          bsum2(r0, r1, s_red);
          beta = sqrt(r0);
          const double ib = 1.0 / beta;
          for (int i = tid; i < len; i += ET) V[i] = cscale(ib, S[i]);
          for (int idx = tid; idx < mh * MH; idx += ET) s_H[idx] = zero;
          __syncthreads();
          bool broke = false;
          for (int j = 0; j < m; ++j) {
            const double2* Vj = V + (int64_t)j * a.lmax;
            ++nmult;
            expo_matvec(op, fp, lo, len, Vj, P);
            const double2 hprev = load_or_zero(j > 0, s_H + (j > 0 ? (j - 1) * MH + j : 0));
            double d0 = 0.0, d1 = 0.0;
            for (int i = tid; i < len; i += ET) {
              double2 w = P[i];
              if (j > 0) w = csub(w, cmul(hprev, V[(int64_t)(j - 1) * a.lmax + i]));
              P[i] = w;
              const double2 v = Vj[i];
              d0 += v.x * w.x + v.y * w.y;
              d1 += v.x * w.y - v.y * w.x;
            }
            bsum2(d0, d1, s_red);
            const double2 hjj = make_double2(d0, d1);
            double nn = 0.0, dz = 0.0;
            for (int i = tid; i < len; i += ET) {
              const double2 w = csub(P[i], cmul(hjj, Vj[i]));
              P[i] = w;
              nn += w.x * w.x + w.y * w.y;
            }
            bsum2(nn, dz, s_red);
            const double hj1j = sqrt(nn);
            if (tid == 0) s_H[j * MH + j] = hjj;
            if (hj1j <= BREAK_TOL) {
              // happy breakdown (expokit.f:3000-3008)
              k1 = 0;
              mbrkdwn = j + 1;
              t_step = t_out - t_now;
              broke = true;
              __syncthreads();
              break;
            }
            if (tid == 0) {
              s_H[j * MH + j + 1] = make_double2(hj1j, 0.0);
              s_H[(j + 1) * MH + j] = make_double2(hj1j, 0.0);
            }
            const double ih = 1.0 / hj1j;
            for (int i = tid; i < len; i += ET) V[(int64_t)(j + 1) * a.lmax + i] = cscale(ih, P[i]);
            __syncthreads();
          }
          if (!broke) {
            ++nmult;
            expo_matvec(op, fp, lo, len, V + (int64_t)m * a.lmax, P);
            double nn = 0.0, dz = 0.0;
            for (int i = tid; i < len; i += ET) nn += P[i].x * P[i].x + P[i].y * P[i].y;
            bsum2(nn, dz, s_red);
            avnorm = sqrt(nn);
          }
          // the corrected scheme (expokit.f:3019-3021)
          if (tid == 0) {
            s_H[(m - 1) * MH + m] = zero;
            s_H[(m + 1) * MH + m] = make_double2(1.0, 0.0);
          }
          __syncthreads();
          double err_loc;
          for (;;) {
            const int mx = mbrkdwn + k1;
            const double2* E = zgpadm(s_H, mx, sgn * t_step, s_W[0], s_W[1], s_W[2], s_W[3], s_sc, &s_piv);
            for (int i = tid; i < mx; i += ET) s_e1[i] = E[i * MH];
            __syncthreads();
            if (k1 == 0) {
              err_loc = tol;
            } else {
              const double p1 = hypot(s_e1[m].x, s_e1[m].y) * beta;
              const double p2 = hypot(s_e1[m + 1].x, s_e1[m + 1].y) * beta * avnorm;
              if (p1 > 10.0 * p2) {
                err_loc = p2;
                xm = 1.0 / (double)m;
              } else if (p1 > p2) {
                err_loc = (p1 * p2) / (p1 - p2);
                xm = 1.0 / (double)m;
              } else {
                err_loc = p1;
                xm = 1.0 / (double)(m - 1);
              }
            }
            if (k1 != 0 && err_loc > DELTA * t_step * tol) {
              // rejected (expokit.f:3066-3097, mxreject = 0: no limit)
              t_step = round_step(GAMMA * t_step * pow(t_step * tol / err_loc, xm));
              __syncthreads();   // s_e1 and the LDS matrices are rewritten
              continue;
            }
            break;
          }
          // w = beta V exp(t_step Hs) e1 (expokit.f:3101-3105)
          const int mxw = mbrkdwn + max(0, k1 - 1);
          for (int i = tid; i < len; i += ET) {
            double2 u = zero;
            for (int k = 0; k < mxw; ++k) u = cfma(s_e1[k], V[(int64_t)k * a.lmax + i], u);
            S[i] = cscale(beta, u);
          }
          t_new = round_step(GAMMA * t_step * pow(t_step * tol / err_loc, xm));
          t_now += t_step;
          __syncthreads();
          if (nstep >= MXSTEP) {
            // expokit.f:3135-3136: iflag = 1 after mxstep steps
            fail = 1;
            break;
          }
        }
      }
      if (op.phase)
        for (int i = tid; i < len; i += ET) S[i] = cmul(S[i], op.phase[op.perm[lo + i]]);
      if (tid == 0 && op.kused) op.kused[(int64_t)step * op.kused_ld + b - op.kcol0] = nmult;
      __syncthreads();
      if (op.ow > 0) {
        // <psi|O|psi> after the step (rm_propagate_expect)
        double d0 = 0.0, d1 = 0.0;
        for (int i = tid; i < len; i += ET) {
          double2 y = zero;
          for (int j = 0; j < op.ow; ++j) {
            const int64_t q = (int64_t)j * op.n + lo + i;
            const int c = op.ocol[q] - lo;
            if (c >= 0 && c < len) y = cfma(op.oval[q], S[c], y);
          }
          d0 += S[i].x * y.x + S[i].y * y.y;
          d1 += S[i].x * y.y - S[i].y * y.x;
        }
        bsum2(d0, d1, s_red);
        if (tid == 0 && op.expect) op.expect[(int64_t)step * op.expect_ld + b - op.kcol0] = make_double2(d0, d1);
      }
    }
    for (int i = tid; i < len; i += ET) op.psi[(int64_t)b * op.ld + op.perm[lo + i]] = S[i];
    if (tid == 0 && fail && op.status) op.status[b - op.kcol0] |= RM_NOT_CONVERGED;
    __syncthreads();   // the next item rewrites the region after these reads
  }
}

// anorm[s] = ||H(F_s)||_1 of the whole matrix for the steps with a field:
// the largest row sum of |sum_s F_s B_s| (H(F) is Hermitian: rows and
// columns give the same norm), one workgroup per step.
__global__ __launch_bounds__(ET) void k_expo_norm(int64_t n, int64_t nsteps, int32_t nslot, int64_t nnz,
                                                  const int64_t* __restrict__ rp,
                                                  const double2* __restrict__ cval,
                                                  const double2* __restrict__ table,
                                                  const int32_t* __restrict__ sflag, double* __restrict__ anorm) {
  __shared__ double red[EW];
  for (int64_t st = blockIdx.x; st < nsteps; st += gridDim.x) {
    if (!sflag[st]) continue;
    const double2* fp = table + st * nslot;
    double mx = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += ET) {
      double rs = 0.0;
      for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
        double2 v = make_double2(0.0, 0.0);
        for (int s = 0; s < nslot; ++s) v = cfma(fp[s], cval[(int64_t)s * nnz + e], v);
        rs += hypot(v.x, v.y);
      }
      mx = fmax(mx, rs);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) anorm[st] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    __syncthreads();
  }
}

}  // namespace

// The Expokit propagation of the items (column, lo, len) of one job: the
// per-step norms, then one launch of k_expokit (one workgroup per item, at
// most 4 per CU).  ws / anorm: device scratch grown here.
int expokit_run(const SubOp& op, const std::vector<int4>& items, int32_t m, double tol, int n_cu, void** ws,
                size_t* ws_bytes, double** anorm, int64_t* anorm_cap, int4** d_items, int64_t* items_cap,
                hipStream_t s) {
  if (items.empty() || op.nsteps == 0) return RM_OK;
  int64_t lmax = 1;
  for (const int4& it : items) lmax = std::max<int64_t>(lmax, it.z);
  const int64_t grid = std::min<int64_t>((int64_t)items.size(), (int64_t)std::max(1, n_cu) * 4);
  const size_t need = sizeof(double2) * (size_t)(m + 3) * (size_t)lmax * (size_t)grid;
  if (*ws_bytes < need) {
    if (*ws) {
      HIPCHK(hipStreamSynchronize(s));
      HIPCHK(hipFree(*ws));
      *ws = nullptr;
      *ws_bytes = 0;
    }
    HIPCHK(hipMalloc(ws, need));
    *ws_bytes = need;
  }
  int rc;
  if ((rc = grow(anorm, anorm_cap, op.nsteps)) || (rc = grow(d_items, items_cap, (int64_t)items.size()))) return rc;
  HIPCHK(hipMemcpyAsync(*d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_expo_norm, dim3((unsigned)std::min<int64_t>(op.nsteps, 4096)), dim3(ET), 0, s, (int64_t)op.n,
                     (int64_t)op.nsteps, op.nslot, op.nnz, op.rp, op.cval, op.table, op.sflag, *anorm);
  ExpoArgs a;
  memset(&a, 0, sizeof(a));
  a.op = op;
  a.items = *d_items;
  a.nitems = (int32_t)items.size();
  a.ws = (double2*)*ws;
  a.lmax = lmax;
  a.anorm = *anorm;
  a.m = m;
  a.tol = tol;
  hipLaunchKernelGGL(k_expokit, dim3((unsigned)grid), dim3(ET), 0, s, a);
  HIPCHK(hipGetLastError());
  // the items' host copy is released by the caller: wait for the upload
  HIPCHK(hipStreamSynchronize(s));
  return RM_OK;
}
