// richmol_amd: rotational densities of wavepackets on a grid of Euler angles
// (the at-scale use of richmol/rotdens.py:10-283).  A wavepacket given by its
// coefficients over the primitive symmetric-top functions,
//   psi(alpha, beta, gamma) = sum_J sum_{m,k} a^J_mk e^{i m alpha} d^J_mk(beta) e^{i k gamma}
// (a^J_mk carrying sqrt((2J+1)/(8 pi^2)) and the state coefficients), is
// evaluated per beta as a two-sided transform of T(beta)_mk = sum_J a^J_mk
// d^J_mk(beta):
//   k_rd_t         T[w][b][m][k] from the coefficients and the d tables;
//   rocBLAS        U = E_alpha T and psi = U E_gamma^T as strided batched
//                  complex products (one batch entry per wavepacket and beta);
//   k_rd_acc       rho = |psi|^2 sin(beta), summed over the wavepackets in a
//                  fixed order (ensemble) or kept per wavepacket.

#include <rocblas/rocblas.h>

#include <mutex>

#include "rm_common.h"

namespace {

// T[w][b][m][k] = sum_{J >= max(|m|, |k|)} a[w][J][m][k] d[J][b][m][k]
// (m, k indexed from -jmax: M = 2 jmax + 1)
__global__ void k_rd_t(int32_t jmax, int32_t nw, int32_t nb, const double2* __restrict__ a,
                       const double* __restrict__ d, double2* __restrict__ T) {
  const int M = 2 * jmax + 1;
  const int64_t mm = (int64_t)M * M;
  const int64_t total = (int64_t)nw * nb * mm;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mk = t % mm;
    const int64_t wb = t / mm;
    const int b = (int)(wb % nb), w = (int)(wb / nb);
    const int m = (int)(mk / M) - jmax, k = (int)(mk % M) - jmax;
    const int j0 = max(abs(m), abs(k));
    double2 acc = make_double2(0.0, 0.0);
    for (int J = j0; J <= jmax; ++J) {
      const double2 c = a[((int64_t)w * (jmax + 1) + J) * mm + mk];
      const double dv = d[((int64_t)J * nb + b) * mm + mk];
      acc.x += c.x * dv;
      acc.y += c.y * dv;
    }
    T[t] = acc;
  }
}

// rho += |psi|^2 sin(beta): psi[w][b] is (na x ng) row-major; ensemble:
// rho[a][b][g] = sum_w (fixed order); otherwise rho[w][a][b][g]
__global__ void k_rd_acc(int32_t nw, int32_t na, int32_t nb, int32_t ng, const double2* __restrict__ psi,
                         const double* __restrict__ sinb, int32_t ensemble, int32_t w0,
                         double* __restrict__ rho) {
  const int64_t npt = (int64_t)na * nb * ng;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npt; p += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(p % ng);
    const int b = (int)((p / ng) % nb);
    const int al = (int)(p / ((int64_t)ng * nb));
    const double s = sinb[b];
    if (ensemble) {
      double acc = rho[p];
      for (int w = 0; w < nw; ++w) {
        const double2 v = psi[(((int64_t)w * nb + b) * na + al) * ng + g];
        acc += (v.x * v.x + v.y * v.y) * s;
      }
      rho[p] = acc;
    } else {
      for (int w = 0; w < nw; ++w) {
        const double2 v = psi[(((int64_t)w * nb + b) * na + al) * ng + g];
        rho[(int64_t)(w0 + w) * npt + p] = (v.x * v.x + v.y * v.y) * s;
      }
    }
  }
}

__global__ void k_rd_zero(double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

std::mutex g_rd_mu;
rocblas_handle g_rd_handle[64] = {};

#define RBCHK(expr)                                                                          \
  do {                                                                                       \
    rocblas_status st_ = (expr);                                                             \
    if (st_ != rocblas_status_success)                                                       \
      return fail(RM_ERR_HIP, "%s failed: %s", #expr, rocblas_status_to_string(st_));       \
  } while (0)

}  // namespace

extern "C" int rm_rotdens(int32_t jmax, int32_t nwp, const void* coef, int32_t na, const void* ealpha, int32_t nb,
                          const double* dtab, const double* sinb, int32_t ng, const void* egamma,
                          int32_t per_wavepacket, void* rho, void* stream) {
  if (jmax < 0 || nwp < 0 || na < 1 || nb < 1 || ng < 1)
    return fail(RM_ERR_ARG, "rm_rotdens: bad sizes (jmax %d, %d wavepackets, grid %d x %d x %d)", jmax, nwp, na, nb,
                ng);
  if (!coef || !ealpha || !dtab || !sinb || !egamma || !rho) return fail(RM_ERR_ARG, "rm_rotdens: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  const int M = 2 * jmax + 1;
  const int64_t mm = (int64_t)M * M, npt = (int64_t)na * nb * ng;
  if (!per_wavepacket || nwp == 0)
    hipLaunchKernelGGL(k_rd_zero, dim3((unsigned)std::min<int64_t>(grid_for(npt, 256), 4096)), dim3(256), 0, s,
                       (double*)rho, npt);
  if (nwp == 0) return RM_OK;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(RM_ERR_ARG, "rm_rotdens: device %d out of range", dev);
  rocblas_handle h;
  {
    std::lock_guard<std::mutex> lock(g_rd_mu);
    if (!g_rd_handle[dev]) RBCHK(rocblas_create_handle(&g_rd_handle[dev]));
    h = g_rd_handle[dev];
    RBCHK(rocblas_set_stream(h, s));
  }
  // wavepackets in chunks: T (nb M^2), U (nb na M) and psi (nb na ng) per wavepacket
  const int64_t per_w = nb * (mm + (int64_t)na * M + (int64_t)na * ng);
  const int64_t budget = (int64_t)1 << 27;   // 2 GiB of complex values at most
  const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(nwp, budget / std::max<int64_t>(per_w, 1)));
  double2* ws = nullptr;
  HIPCHK(hipMallocAsync((void**)&ws, sizeof(double2) * (size_t)per_w * chunk, s));
  double2* T = ws;
  double2* U = T + (int64_t)chunk * nb * mm;
  double2* P = U + (int64_t)chunk * nb * na * M;
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  int rc = RM_OK;
  for (int w0 = 0; w0 < nwp && rc == RM_OK; w0 += chunk) {
    const int nw = std::min(chunk, nwp - w0);
    const int64_t nt = (int64_t)nw * nb * mm;
    hipLaunchKernelGGL(k_rd_t, dim3((unsigned)std::min<int64_t>(grid_for(nt, 256), 65535)), dim3(256), 0, s, jmax,
                       nw, nb, (const double2*)coef + (int64_t)w0 * (jmax + 1) * mm, dtab, T);
    const int batch = nw * nb;
    // U_i (na x M, row-major) = E_alpha (na x M) T_i (M x M): column-major U^T = T^T E_alpha^T
    rocblas_status st = rocblas_zgemm_strided_batched(
        h, rocblas_operation_none, rocblas_operation_none, M, na, M, &one, (const rocblas_double_complex*)T, M, mm,
        (const rocblas_double_complex*)ealpha, M, 0, &zero, (rocblas_double_complex*)U, M, (int64_t)na * M, batch);
    // psi_i (na x ng, row-major) = U_i E_gamma^T: column-major psi^T = E_gamma U^T
    if (st == rocblas_status_success)
      st = rocblas_zgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, ng, na, M, &one,
                                         (const rocblas_double_complex*)egamma, M, 0,
                                         (const rocblas_double_complex*)U, M, (int64_t)na * M, &zero,
                                         (rocblas_double_complex*)P, ng, (int64_t)na * ng, batch);
    if (st != rocblas_status_success) {
      rc = fail(RM_ERR_HIP, "rm_rotdens: rocblas_zgemm_strided_batched failed: %s", rocblas_status_to_string(st));
      break;
    }
    hipLaunchKernelGGL(k_rd_acc, dim3((unsigned)std::min<int64_t>(grid_for(npt, 256), 4096)), dim3(256), 0, s, nw,
                       na, nb, ng, (const double2*)P, sinb, per_wavepacket ? 0 : 1, w0, (double*)rho);
    if (hipGetLastError() != hipSuccess) rc = fail(RM_ERR_HIP, "rm_rotdens: launch failed");
  }
  (void)hipFreeAsync(ws, s);
  return rc;
}
