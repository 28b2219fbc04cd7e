// richmol_amd: Kronecker-block operators (factored H = sum K_w (x) MF_w) for
// large rotor bases, and a batched fixed-order Lanczos propagator on them.
//
// richmol applies a tensor block by block as Y[J1] += MF_w X[J2] K_w^T
// (field.py:1235-1237).  Here, per output J1, every source block
// q = (J2, w, tensor) is stacked: Z_J1 = [MF_q X[J2]]_q (the banded m-mixing
// applied to the matvec input of all wavepackets at once), and
// Y[J1] = Z_J1 K_J1^T with K_J1 = [K_q]_q -- one dense product per J1 whose
// rows are (wavepacket, m1), the GEMM the matrix cores run (rocBLAS).  The
// complex product is done in real arithmetic on split layouts:
//   [Yr | Yi] = [Zr | Zi] [[Kr^T, Ki^T], [-Ki^T, Kr^T]].
// Vectors are kept per J block as rows (wavepacket, m) of 2 (2J+1) reals
// [re(k) | im(k)]; a wavepacket's (J) chunk is contiguous, so its dot
// products are two-pass sums over (J, wavepacket) chunks in a fixed order.
//
// The Lanczos recurrence of tdse.py:417-486 with a fixed Krylov dimension m
// (unnormalised first vector, V_{k+1} = W/beta formed where W is read, an
// exact breakdown giving a zero vector, whose coefficient in exp(fT)e0 is
// exactly zero as the reference's restart vector's is), the split-operator
// phases around it, and exp(fT)e0 per wavepacket (Taylor with scaling, one
// wave each).  FP64, or FP32 (complex64 Z, K and matvec output; the
// recurrence, norms, exp(fT), phases and the wavepacket in FP64).

#include <rocblas/rocblas.h>

#include <algorithm>
#include <vector>

#include "rm_common.h"

namespace {

// One (output J1, source block) pair: Z_J1 columns [koff, koff + dk2) of
// each half (re / im), computed from the J2 block of the input.
struct KTask {
  int32_t J1, J2, dm1, dm2, dk2, koff, ktot, nsig;
  int32_t sig0;        // first entry of the block's sigma list
  int32_t blk;
  int64_t mf_off;      // MF values of the block: [nsig][dm1] complex
  int64_t zoff;        // Z_J1 (reals) for this call's wavepackets
  int64_t xoff;        // X[J2] (reals) in the split vector layout
};

// J descriptor of the stored form's matrix-core product (k_kp_mm)
struct MJ {
  int32_t M, N, K, pad;   // rows (wavepacket, m1), columns 2 dk1, stacked reals 2 Ktot
  int64_t zoff, yoff;     // the J's Z block and output block
};

struct KBlk {          // per block: MF assembly from the field products
  int32_t dm1, nsig, ncart, slot0;
  int64_t mf_off, d_off;
};

}  // namespace

struct rm_kplan {
  int32_t nJ = 0, nslot = 0, nblk = 0;
  std::vector<int32_t> Js;              // ascending
  std::vector<int64_t> h_voff;          // per J: complex offset of the J block per wavepacket
  int64_t n = 0;                        // states per wavepacket
  std::vector<KBlk> h_blk;
  std::vector<int32_t> h_J1, h_J2, h_sig;
  std::vector<int64_t> h_sig_off;
  std::vector<std::vector<int32_t>> src;   // per J index: source blocks
  std::vector<int32_t> ktot;               // per J index
  std::vector<int64_t> kq_off;             // per block: column offset within its J1's K stack
  double2* d_mdiag = nullptr;
  int32_t* d_sig = nullptr;
  KBlk* d_blk = nullptr;
  double2* d_mf = nullptr;
  int64_t mf_count = 0;
  std::vector<double*> kbig64;            // per J index: (2 Ktot) x (2 dk) row-major
  std::vector<float*> kbig32;
  rocblas_handle rb = nullptr;
  // fused form: per J1 the stacked K with columns padded to groups of four
  // (FGK) and the rows of a group interleaved [re x 4 | im x 4], output
  // columns padded to FBN; colmap: per padded stacked column, (block << 16 | k2)
  // of the J1's source blocks, -1 for padding
  int32_t kernel = RM_KKERNEL_GEMM;
  std::vector<int32_t> kpad, npad;      // per J index
  std::vector<int64_t> kf_off, cm_off;  // per J index
  double* d_kf64 = nullptr;
  float* d_kf32 = nullptr;
  int32_t* d_colmap = nullptr;
  char* d_fz = nullptr;                 // per call: J descriptors and tiles
  int64_t fz_cap = 0;
  const int4* fz_tiles = nullptr;
  int64_t ntiles = 0;
  std::vector<char> h_fz;
  // stored form on the matrix cores (k_kp_mm): per call, the J descriptors,
  // the K-stack pointers and per matvec chunk its tiles
  char* d_mm = nullptr;
  int64_t mm_cap = 0;
  std::vector<char> h_mm;
  const MJ* mm_j = nullptr;
  const int4* mm_tiles = nullptr;       // FBM x FBN tiles (FP64, k_kp_mm)
  const int4* mm_tiles_w = nullptr;     // MBM x MBN tiles (FP32, k_kp_mm_w)
  const double* const* mm_kb64 = nullptr;
  const float* const* mm_kb32 = nullptr;
  std::vector<int64_t> mm_chunk, mm_chunk_w;   // KCHUNK + 1 tile offsets each
  int64_t* d_voff = nullptr;            // per J: complex offset (per wavepacket)
  int32_t* d_Js = nullptr;
  // per-call workspace
  void* ws = nullptr;
  size_t ws_bytes = 0;
  KTask* d_tasks = nullptr;
  int64_t tasks_cap = 0;
  int64_t* d_meta = nullptr;            // per call: row offsets and task ranges per J
  int64_t meta_cap = 0;
  int64_t ztiles = 0;
  int32_t zr = 16;                      // Z rows per tile of the Z-build kernel
  // the dense products run on their own stream, chunk c after the Z build of chunk c
  hipStream_t gs = nullptr;
  std::vector<hipEvent_t> ev;           // KCHUNK Z-done events + one products-done event
  std::vector<int32_t> chunk_j;         // per call: J index bounds of the chunks (KCHUNK + 1)
  size_t z_lds = 0;                     // its LDS bytes
  std::vector<KTask> h_tasks;
  std::vector<int64_t> h_meta;          // host copy (its source must outlive the async upload)
  double2* d_fp = nullptr;              // field products of a call's steps
  int64_t fp_cap = 0;
  double2* d_ex = nullptr;              // as an observable: per-wavepacket values when only sums are wanted
  int64_t ex_cap = 0;
};

namespace {

__host__ __device__ inline int64_t jdim(int32_t J) { return 2 * (int64_t)J + 1; }

// Chunks of output J per matvec: the Z build of chunk c + 1 overlaps the dense
// products of chunk c (memory-bound next to MFMA-bound work).
constexpr int KCHUNK = 5;

// Fused matvec tiles: FBM output rows (wavepacket, m1) x FBN output columns
// per workgroup (four multiplying waves, 2 x 2, each 2 x 4 MFMA blocks of
// 16 x 16), the stacked K dimension in chunks of FBK reals (two groups of four
// complex columns); at most FMAXT source blocks per output J.
constexpr int FBM = 64, FBN = 128, FBK = 16, FGK = 4, FMAXT = 32;

// Per-call descriptor of one output J of the fused matvec.
struct FJ {
  int32_t dm1, N, Npad, kpad;    // m rows per wavepacket, output columns (2 dk1), padded; padded stacked columns
  int32_t task0, ntask;          // its source blocks in the call's task list
  int64_t M;                     // ncols * dm1
  int64_t cm_off, kf_off, yoff;  // column map, K stack, output block
};

// Source block of the fused kernel, staged in LDS.
struct FT {
  int32_t J1, J2, dm2, dk2, nsig, sg[5];
  int64_t mf_off, xoff;
};

// MF[blk][t][m1] = sum_c fprod[slot0 + c] mdiag[blk][t][c][m1]
__global__ void k_kp_mf(const KBlk* __restrict__ blk, const double2* __restrict__ mdiag,
                        const double2* __restrict__ fp, double2* __restrict__ mf) {
  const KBlk bk = blk[blockIdx.y];
  const int64_t cnt = (int64_t)bk.nsig * bk.dm1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / bk.dm1, m1 = e % bk.dm1;
    double2 acc = make_double2(0.0, 0.0);
    for (int c = 0; c < bk.ncart; ++c)
      acc = cfma(fp[bk.slot0 + c], mdiag[bk.d_off + (t * bk.ncart + c) * bk.dm1 + m1], acc);
    mf[bk.mf_off + e] = acc;
  }
}

// Z rows of one tile (output J1, wavepacket b, rows m1 in [m0, m0 + nr)) for
// every source block q of J1:
//   Z[(b, m1)][koff_q + k2] = ib[b] sum_t MF_q[t][m1] X[J2_q](b, m2_t)[k2]
// (re and im halves), m2_t = m1 - J1 - sigma_t + J2_q.  Per block, the X
// rows the tile needs (nr + 4 at most: |sigma| <= 2) are one contiguous range
// of the input, staged in LDS with coalesced loads; the tile's MF values
// beside them.  Tiles are found from the per-J tile offsets meta[0..nJ] and
// task ranges meta[nJ+1..2nJ+1]; zr = rows per tile.
template <typename ZT>
__global__ __launch_bounds__(256) void k_kp_z(const KTask* __restrict__ tasks, const int64_t* __restrict__ meta,
                                              int32_t nJ, const int32_t* __restrict__ sig,
                                              const double2* __restrict__ mf, const double* __restrict__ X,
                                              const double* __restrict__ ib, ZT* __restrict__ Z, int32_t zr,
                                              int64_t tile0) {
  extern __shared__ double smem[];
  const int64_t tile = tile0 + blockIdx.x;
  const int64_t* tile_off = meta;
  const int64_t* task_off = meta + nJ + 1;
  int lo = 0, hi = nJ;   // tile_off[lo] <= tile < tile_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_off[mid] <= tile) lo = mid; else hi = mid;
  }
  const int64_t q0 = task_off[lo], q1 = task_off[lo + 1];
  const int dm1 = tasks[q0].dm1;
  const int nmt = (dm1 + zr - 1) / zr;
  const int64_t tt = tile - tile_off[lo];
  const int64_t b = tt / nmt;
  const int m0 = (int)(tt % nmt) * zr;
  const int nr = min(zr, dm1 - m0);
  const double s = ib[b];
  double2* s_mf = (double2*)smem;      // [5][zr]
  double* s_x = smem + 10 * zr;        // [nr + 4][2 dk2]
  int staged = -1;   // J2 whose rows s_x holds
  for (int64_t q = q0; q < q1; ++q) {
    const KTask tk = tasks[q];
    int sg[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) sg[t] = t < tk.nsig ? sig[tk.sig0 + t] : 0;
    const int64_t xld = 2 * (int64_t)tk.dk2;
    const int lo2 = max(0, m0 - tk.J1 + tk.J2 - 2);
    const int hi2 = min(tk.dm2 - 1, m0 + nr - 1 - tk.J1 + tk.J2 + 2);
    if (tk.J2 != staged) {
      // blocks of one J2 are consecutive: their rows are staged once
      const int64_t nx = hi2 >= lo2 ? (int64_t)(hi2 - lo2 + 1) * xld : 0;
      const double* src = X + tk.xoff + (b * tk.dm2 + lo2) * xld;
      for (int64_t i = threadIdx.x; i < nx; i += blockDim.x) s_x[i] = src[i];
      staged = tk.J2;
    }
    for (int i = threadIdx.x; i < tk.nsig * nr; i += blockDim.x) {
      const int t = i / nr, r = i % nr;
      s_mf[t * zr + r] = mf[tk.mf_off + (int64_t)t * tk.dm1 + m0 + r];
    }
    __syncthreads();
    ZT* zt = Z + tk.zoff + (b * tk.dm1 + m0) * 2 * (int64_t)tk.ktot;
    // one row per wave at a time, 64 consecutive columns per pass: the row's
    // taps (MF value, source row, in range or not) are the same on every lane
    const int lane = threadIdx.x & 63;
    for (int r = threadIdx.x >> 6; r < nr; r += blockDim.x >> 6) {
      const int m1 = m0 + r;
      double2 f[5];
      const double* xr[5];
      bool in[5];
#pragma unroll
      for (int t = 0; t < 5; ++t) {
        const int m2 = m1 - tk.J1 - sg[t] + tk.J2;
        in[t] = t < tk.nsig && m2 >= 0 && m2 < tk.dm2;
        f[t] = in[t] ? s_mf[t * zr + r] : make_double2(0.0, 0.0);
        xr[t] = s_x + (in[t] ? m2 - lo2 : 0) * xld;
      }
      ZT* zrow = zt + (int64_t)r * 2 * tk.ktot + tk.koff;
      for (int k2 = lane; k2 < tk.dk2; k2 += 64) {
        double zre = 0.0, zim = 0.0;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
          if (!in[t]) continue;
          const double a = xr[t][k2], c = xr[t][tk.dk2 + k2];
          zre += f[t].x * a - f[t].y * c;
          zim += f[t].x * c + f[t].y * a;
        }
        zrow[k2] = (ZT)(zre * s);
        zrow[tk.ktot + k2] = (ZT)(zim * s);
      }
    }
    __syncthreads();
  }
}

template <typename T>
struct FAcc;
template <>
struct FAcc<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mfma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of D held in item v of lane l (column l % 16)
  static __device__ __forceinline__ int row(int l, int v) { return (l >> 4) + 4 * v; }
};
template <>
struct FAcc<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mfma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
};

// The matvec Y[J1] = Z_J1 K_J1^T (real split form) with Z never stored:
// every chunk of FBK stacked reals is built in LDS from the input and the MF
// diagonals (as k_kp_z does, scaled by ib[b]) next to the matching rows of
// the K stack, and multiplied on the matrix cores (16 x 16 x 4 MFMA; lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; item v of lane l holds
// D[l / 16 + 4 v][l % 16] in FP64 and D[4 (l / 16) + v][l % 16] in FP32).
// One workgroup of eight waves per (output J, FBM rows, FBN columns) tile of
// the tile list: waves 0-3 multiply, waves 4-7 gather the next chunk into the
// other half of the double-buffered LDS tiles meanwhile (each SIMD holds one
// wave of each kind); every wave passes the same barriers, one per chunk.  A
// gathering thread builds one stacked column (t % 8) of two fixed rows; its
// column map entry is read one chunk ahead.  Offsets into the input, output
// and K stack are 32-bit (the host checks the sizes).
template <typename T>
__global__ __launch_bounds__(512) void k_kp_fused(const int4* __restrict__ tiles, const FJ* __restrict__ jd,
                                                  const KTask* __restrict__ tasks, const int32_t* __restrict__ sig,
                                                  const int32_t* __restrict__ colmap,
                                                  const double2* __restrict__ mf, const double* __restrict__ X,
                                                  const double* __restrict__ ib, const T* __restrict__ Kf,
                                                  T* __restrict__ Y) {
  using V = typename FAcc<T>::V;
  constexpr int RU = (FBM * 8) / 256;            // rows per gathering thread
  constexpr int KU = (FBK * FBN) / 256;          // K values per gathering thread
  __shared__ T sZ[2][FBK][FBM];
  __shared__ T sK[2][FBK][FBN];
  __shared__ FT sT[FMAXT];
  const int4 tl = tiles[blockIdx.x];
  const FJ j = jd[tl.x];
  const int r0 = tl.y * FBM;
  const int n0 = tl.z * FBN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const bool prod = w >= 4;
  const int pt = t & 255;                        // gathering thread index
  for (int q = t; q < j.ntask; q += blockDim.x) {
    const KTask tk = tasks[j.task0 + q];
    FT f;
    f.J1 = tk.J1;
    f.J2 = tk.J2;
    f.dm2 = tk.dm2;
    f.dk2 = tk.dk2;
    f.nsig = tk.nsig;
    for (int u = 0; u < 5; ++u) f.sg[u] = u < tk.nsig ? sig[tk.sig0 + u] : 0;
    f.mf_off = tk.mf_off;
    f.xoff = tk.xoff;
    sT[q] = f;
  }
  const int jc = pt & 7;
  const int kk = (jc >> 2) * 8 + (jc & 3);
  int rb[RU], rm[RU];
  bool rv[RU];
#pragma unroll
  for (int u = 0; u < RU; ++u) {
    const int r = r0 + (pt >> 3) + 32 * u;
    rv[u] = r < j.M;
    rb[u] = rv[u] ? r / j.dm1 : 0;
    rm[u] = rv[u] ? r - rb[u] * j.dm1 : 0;
  }
  __syncthreads();
  const int nchunk = 2 * j.kpad / FBK;
  int cm_next = (prod && nchunk > 0) ? colmap[j.cm_off + jc] : -1;
  // gather chunk ch into LDS buffer buf (gathering waves only)
  auto produce = [&](int ch, int buf) __attribute__((always_inline)) {
    const int cm = cm_next;
    if (ch + 1 < nchunk) cm_next = colmap[j.cm_off + 8 * (ch + 1) + jc];
    const T* kr = Kf + j.kf_off + (int64_t)(FBK * ch) * j.Npad + n0;
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int e = pt + 256 * u;
      sK[buf][e / FBN][e % FBN] = kr[(e / FBN) * j.Npad + (e % FBN)];
    }
    const bool live = cm >= 0;
    const FT& f = sT[live ? (cm >> 16) : 0];
    const int k2 = live ? (cm & 0xffff) : 0;
    const int xld = 2 * f.dk2;
    double zre[RU], zim[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      zre[u] = 0.0;
      zim[u] = 0.0;
    }
#pragma unroll
    for (int s5 = 0; s5 < 5; ++s5) {
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        const int m2 = rm[u] - f.J1 - f.sg[s5] + f.J2;
        const bool ok = live && rv[u] && s5 < f.nsig && m2 >= 0 && m2 < f.dm2;
        const int xo = (int)f.xoff + (rb[u] * f.dm2 + (ok ? m2 : 0)) * xld + k2;
        const double2 fv = ok ? mf[(int)f.mf_off + s5 * j.dm1 + rm[u]] : make_double2(0.0, 0.0);
        const double xr = ok ? X[xo] : 0.0, xi = ok ? X[xo + f.dk2] : 0.0;
        zre[u] += fv.x * xr - fv.y * xi;
        zim[u] += fv.x * xi + fv.y * xr;
      }
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const double sc = ib[rb[u]];
      const int row = (pt >> 3) + 32 * u;
      sZ[buf][kk][row] = (T)(zre[u] * sc);
      sZ[buf][kk + 4][row] = (T)(zim[u] * sc);
    }
  };
  V acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = V{0, 0, 0, 0};
  const int wr = w & 1, wc = (w >> 1) & 1;
  if (prod && nchunk > 0) produce(0, 0);
  __syncthreads();
  for (int ch = 0; ch < nchunk; ++ch) {
    const int cur = ch & 1;
    if (prod) {
      if (ch + 1 < nchunk) produce(ch + 1, cur ^ 1);
    } else {
#pragma unroll
      for (int ks = 0; ks < FBK / 4; ++ks) {
        const int kq = ks * 4 + (lane >> 4);
        T av[2], bv[4];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = sZ[cur][kq][wr * 32 + a * 16 + (lane & 15)];
#pragma unroll
        for (int c = 0; c < 4; ++c) bv[c] = sK[cur][kq][wc * 64 + c * 16 + (lane & 15)];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[a][c] = FAcc<T>::mfma(av[a], bv[c], acc[a][c]);
      }
    }
    __syncthreads();
  }
  if (prod) return;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = n0 + wc * 64 + c * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + wr * 32 + a * 16 + FAcc<T>::row(lane, v);
        if (row < j.M && col < j.N) Y[j.yoff + (int64_t)row * j.N + col] = acc[a][c][v];
      }
    }
}

// The stored form's product Y[J1] = Z_J1 K_J1^T (real split form) on the
// matrix cores: one workgroup of eight waves per (output J, FBM rows, FBN
// columns) tile of a chunk's tile list.  Waves 0-3 multiply (16 x 16 x 4
// MFMA, operand and result layouts as in k_kp_fused, 2 x 4 blocks each);
// waves 4-7 copy FBK-deep slices of the Z rows and of the K stack (row-major
// (2 Ktot) x (2 dk1)) into the other half of the double-buffered LDS tiles,
// one slice ahead in registers (the loads of slice c + 2 are in flight while
// slice c is multiplied); every wave passes the same barriers, one per slice.

template <typename T>
__global__ __launch_bounds__(512) void k_kp_mm(const int4* __restrict__ tiles, const MJ* __restrict__ jd,
                                               const T* const* __restrict__ kb, const T* __restrict__ Z,
                                               T* __restrict__ Y) {
  using V = typename FAcc<T>::V;
  __shared__ T sA[2][FBK][FBM];
  __shared__ T sB[2][FBK][FBN];
  const int4 tl = tiles[blockIdx.x];
  const MJ j = jd[tl.x];
  const T* __restrict__ B = kb[tl.x];
  const int r0 = tl.y * FBM, n0 = tl.z * FBN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const bool prod = w >= 4;
  const int pt = t & 255;
  const int nchunk = (j.K + FBK - 1) / FBK;
  // copying thread pt: Z row r0 + pt / 4, stacked reals 4 (pt % 4) .. + 3 of
  // the slice; K-stack row pt / 16 of the slice, columns 8 (pt % 16) .. + 7
  const int ar = pt >> 2, ak = (pt & 3) * 4, bk = pt >> 4, bc = (pt & 15) * 8;
  const bool arow = r0 + ar < j.M;
  const T* __restrict__ Ar = Z + j.zoff + (int64_t)(arow ? r0 + ar : 0) * j.K;
  T ra[4], rb[8];
  auto load = [&](int ch) __attribute__((always_inline)) {
    const int k0 = ch * FBK;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + ak + u;
      ra[u] = (arow && k < j.K) ? Ar[k] : (T)0;
    }
    const int k = k0 + bk;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int c = n0 + bc + u;
      rb[u] = (k < j.K && c < j.N) ? B[(int64_t)k * j.N + c] : (T)0;
    }
  };
  auto store = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < 4; ++u) sA[buf][ak + u][ar] = ra[u];
#pragma unroll
    for (int u = 0; u < 8; ++u) sB[buf][bk][bc + u] = rb[u];
  };
  V acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = V{0, 0, 0, 0};
  const int wr = w & 1, wc = (w >> 1) & 1;
  if (prod && nchunk > 0) {
    load(0);
    store(0);
    if (nchunk > 1) load(1);
  }
  __syncthreads();
  for (int ch = 0; ch < nchunk; ++ch) {
    const int cur = ch & 1;
    if (prod) {
      if (ch + 1 < nchunk) {
        store(cur ^ 1);
        if (ch + 2 < nchunk) load(ch + 2);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < FBK / 4; ++ks) {
        const int kq = ks * 4 + (lane >> 4);
        T av[2], bv[4];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = sA[cur][kq][wr * 32 + a * 16 + (lane & 15)];
#pragma unroll
        for (int c = 0; c < 4; ++c) bv[c] = sB[cur][kq][wc * 64 + c * 16 + (lane & 15)];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[a][c] = FAcc<T>::mfma(av[a], bv[c], acc[a][c]);
      }
    }
    __syncthreads();
  }
  if (prod) return;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = n0 + wc * 64 + c * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + wr * 32 + a * 16 + FAcc<T>::row(lane, v);
        if (row < j.M && col < j.N) Y[j.yoff + (int64_t)row * j.N + col] = acc[a][c][v];
      }
    }
}

// The same product for FP32 operands, with every wave multiplying: one
// workgroup of eight waves per (output J, MBM rows, MBN columns) tile, each
// wave a 32 x 64 block (2 x 4 MFMA blocks).  Each thread issues its loads of
// slice c + 1 before the products of slice c and writes them to the other
// LDS buffer after them, one barrier per slice; the LDS rows are padded so
// that the k rows read by one instruction fall on different banks.  (For
// FP64 operands the form above, with copying waves beside the multiplying
// ones, measures faster; for FP32 this one.)
constexpr int MBM = 128, MBN = 128;
constexpr int MPAD = 16;

template <typename T>
__global__ __launch_bounds__(512) void k_kp_mm_w(const int4* __restrict__ tiles, const MJ* __restrict__ jd,
                                                 const T* const* __restrict__ kb, const T* __restrict__ Z,
                                                 T* __restrict__ Y) {
  using V = typename FAcc<T>::V;
  __shared__ T sA[2][FBK][MBM + MPAD];
  __shared__ T sB[2][FBK][MBN + MPAD];
  const int4 tl = tiles[blockIdx.x];
  const MJ j = jd[tl.x];
  const T* __restrict__ B = kb[tl.x];
  const int r0 = tl.y * MBM, n0 = tl.z * MBN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nchunk = (j.K + FBK - 1) / FBK;
  // thread t copies Z row r0 + t / 4, stacked reals 4 (t % 4) .. + 3 of a
  // slice, and K-stack row t / 32 of the slice, columns 4 (t % 32) .. + 3
  const int ar = t >> 2, ak = (t & 3) * 4, bk = t >> 5, bc = (t & 31) * 4;
  const bool arow = r0 + ar < j.M;
  const T* __restrict__ Ar = Z + j.zoff + (int64_t)(arow ? r0 + ar : 0) * j.K;
  T ra[4], rb[4];
  auto load = [&](int ch) __attribute__((always_inline)) {
    const int k0 = ch * FBK;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + ak + u;
      ra[u] = (arow && k < j.K) ? Ar[k] : (T)0;
    }
    const int k = k0 + bk;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = n0 + bc + u;
      rb[u] = (k < j.K && c < j.N) ? B[(int64_t)k * j.N + c] : (T)0;
    }
  };
  auto store = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < 4; ++u) sA[buf][ak + u][ar] = ra[u];
#pragma unroll
    for (int u = 0; u < 4; ++u) sB[buf][bk][bc + u] = rb[u];
  };
  V acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = V{0, 0, 0, 0};
  const int wr = w & 3, wc = w >> 2;   // 4 x 2 waves of 32 x 64
  if (nchunk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int ch = 0; ch < nchunk; ++ch) {
    const int cur = ch & 1;
    if (ch + 1 < nchunk) load(ch + 1);
#pragma unroll
    for (int ks = 0; ks < FBK / 4; ++ks) {
      const int kq = ks * 4 + (lane >> 4);
      T av[2], bv[4];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = sA[cur][kq][wr * 32 + a * 16 + (lane & 15)];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = sB[cur][kq][wc * 64 + c * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = FAcc<T>::mfma(av[a], bv[c], acc[a][c]);
    }
    if (ch + 1 < nchunk) store(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = n0 + wc * 64 + c * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + wr * 32 + a * 16 + FAcc<T>::row(lane, v);
        if (row < j.M && col < j.N) Y[j.yoff + (int64_t)row * j.N + col] = acc[a][c][v];
      }
    }
}

// Per (J, wavepacket) chunk: V_it = ib W (stored), alpha partial = vdot(Y, V_it)
template <typename YT>
__global__ __launch_bounds__(256) void k_kp_dot(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                                                const double* __restrict__ W, const double* __restrict__ ib,
                                                const YT* __restrict__ Y, double* __restrict__ V,
                                                double* __restrict__ part, int32_t nJ) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;   // reals
  const double s = ib[b];
  double a0 = 0.0, a1 = 0.0;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k, pi = pr + d;
    const double vr = s * W[pr], vi = s * W[pi];
    V[pr] = vr;
    V[pi] = vi;
    const double yr = (double)Y[pr], yi = (double)Y[pi];
    a0 += yr * vr + yi * vi;
    a1 += yr * vi - yi * vr;
  }
  __shared__ double red[2][4];
  a0 = wave_allsum(a0);
  a1 = wave_allsum(a1);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a0;
    red[1][threadIdx.x >> 6] = a1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((int64_t)b * nJ + j) * 2] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    part[((int64_t)b * nJ + j) * 2 + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// W = Y - alpha V_it - beta V_{it-1} per chunk, partial |W|^2
template <typename YT>
__global__ __launch_bounds__(256) void k_kp_upd(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                                                const YT* __restrict__ Y, const double* __restrict__ Vit,
                                                const double* __restrict__ Vprev, const double2* __restrict__ alpha,
                                                const double* __restrict__ beta, int32_t it, int32_t kcap,
                                                double* __restrict__ W, double* __restrict__ part, int32_t nJ) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  const double2 al = alpha[(int64_t)b * kcap + it];
  const double be = it > 0 ? beta[(int64_t)b * (kcap + 1) + it] : 0.0;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k, pi = pr + d;
    const double vr = Vit[pr], vi = Vit[pi];
    double wr = (double)Y[pr] - (al.x * vr - al.y * vi);
    double wi = (double)Y[pi] - (al.x * vi + al.y * vr);
    if (it > 0) {
      wr -= be * Vprev[pr];
      wi -= be * Vprev[pi];
    }
    W[pr] = wr;
    W[pi] = wi;
    acc += wr * wr + wi * wi;
  }
  __shared__ double red[4];
  acc = wave_allsum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)b * nJ + j] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Second pass: per wavepacket the chunks' partials in J order.  mode 0:
// alpha (complex) -> alpha[b][it]; mode 1: |W|^2 -> beta[b][it+1] and ib.
__global__ void k_kp_fin(const double* __restrict__ part, int32_t nJ, int32_t ncols, int32_t mode, int32_t it,
                         int32_t kcap, double2* __restrict__ alpha, double* __restrict__ beta,
                         double* __restrict__ ib, double2* __restrict__ G, double2* __restrict__ cb,
                         int32_t* __restrict__ kdone) {
  const int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= ncols) return;
  double s0 = 0.0, s1 = 0.0;
  for (int j = lane; j < nJ; j += 64) {
    if (mode == 0) {
      s0 += part[((int64_t)b * nJ + j) * 2];
      s1 += part[((int64_t)b * nJ + j) * 2 + 1];
    } else {   // modes 1, 2: one partial per chunk
      s0 += part[(int64_t)b * nJ + j];
    }
  }
  s0 = wave_allsum(s0);
  s1 = wave_allsum(s1);
  if (lane != 0) return;
  if (mode == 0) {
    alpha[(int64_t)b * kcap + it] = make_double2(s0, s1);
  } else if (mode == 2) {
    // adaptive rule, start of a step: G_00 = |V_0|^2 (V_0 unnormalised), c_prev = [1]
    G[(int64_t)b * kcap * kcap] = make_double2(s0, 0.0);
    cb[(int64_t)b * 2 * kcap + kcap] = make_double2(1.0, 0.0);
    kdone[b] = 0;
  } else {
    const double be = sqrt(s0);
    beta[(int64_t)b * (kcap + 1) + it + 1] = be;
    // exact breakdown: a zero vector stands for the reference's restart (its coefficient is 0)
    ib[b] = be != 0.0 ? 1.0 / be : 0.0;
  }
}

// Start of a step: V_0 = P S (the phases of the split operator), W = V_0, ib = 1.
// (adaptive rule: part != NULL receives the chunks' |V_0|^2 partials)
__global__ void k_kp_start(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                           const double2* __restrict__ ph, const double* __restrict__ S, double* __restrict__ W,
                           double* __restrict__ ib, double* __restrict__ beta, int32_t kcap,
                           double* __restrict__ part, int32_t nJ) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k, pi = pr + d;
    const double2 p = ph ? ph[voff[j] + i] : make_double2(1.0, 0.0);
    const double sr = S[pr], si = S[pi];
    const double wr = p.x * sr - p.y * si, wi = p.x * si + p.y * sr;
    W[pr] = wr;
    W[pi] = wi;
    acc += wr * wr + wi * wi;
  }
  if (part) {
    __shared__ double red[4];
    acc = wave_allsum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[(int64_t)b * nJ + j] = (red[0] + red[1]) + (red[2] + red[3]);
  }
  if (j == 0 && threadIdx.x == 0) {
    ib[b] = 1.0;
    beta[(int64_t)b * (kcap + 1)] = 0.0;
  }
}

// A step without field: S = P P S (tdse.py:368-380 with the expv skipped)
__global__ void k_kp_phase2(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                            const double2* __restrict__ ph, double* __restrict__ S) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k, pi = pr + d;
    const double2 p = ph[voff[j] + i];
    double sr = S[pr], si = S[pi];
    double tr = p.x * sr - p.y * si, ti = p.x * si + p.y * sr;
    S[pr] = p.x * tr - p.y * ti;
    S[pi] = p.x * ti + p.y * tr;
  }
}

// exp(f T) e0 per wavepacket (one wave each)
__global__ __launch_bounds__(64) void k_kp_expm(const double2* __restrict__ alpha, const double* __restrict__ beta,
                                                int32_t m, int32_t kcap, double2 f, double2* __restrict__ coef) {
  __shared__ double2 s_a[KMAX];
  __shared__ double s_b[KMAX + 1];
  __shared__ double2 s_c[KMAX];
  const int b = blockIdx.x, l = threadIdx.x;
  for (int i = l; i < m; i += 64) {
    s_a[i] = alpha[(int64_t)b * kcap + i];
    s_b[i] = beta[(int64_t)b * (kcap + 1) + i];
  }
  wave_sync();
  if (m == 1) {
    if (l == 0) s_c[0] = make_double2(1.0, 0.0);
  } else {
    expm_e0_any(m, s_a, s_b, f, s_c);
  }
  wave_sync();
  for (int i = l; i < m; i += 64) coef[(int64_t)b * kcap + i] = s_c[i];
}

// Adaptive rule after Krylov vector k (tdse.py:466-476), one wave per
// wavepacket still running: Gram column k, c = exp(f T_{k+1}) e0 and the
// change conv = |u_k - u_{k-1}|^2 (adaptive_metric); the wavepacket stops at
// dimension k + 1 when conv <= tol or k + 1 == maxorder (then status gets
// RM_NOT_CONVERGED), keeping c as its coefficients.  After an
// exact breakdown (beta_k == 0) the zero vector that stands for the
// reference's restart vector gets the restart vector's Gram column e_k.
// Reaching maxorder sets RM_NOT_CONVERGED (the reference raises whenever its
// final dimension equals maxorder, tdse.py:480-484).
// running: incremented by every wavepacket that goes on.
__global__ __launch_bounds__(64) void k_kp_conv(const double2* __restrict__ alpha, const double* __restrict__ beta,
                                                int32_t kcap, int32_t k, double2 f, double tol, int32_t maxorder,
                                                double2* __restrict__ G, double2* __restrict__ cb,
                                                int32_t* __restrict__ kdone, double2* __restrict__ coef,
                                                int32_t* __restrict__ status, int32_t* __restrict__ running) {
  __shared__ double2 s_a[KMAX];
  __shared__ double s_b[KMAX + 1];
  const int b = blockIdx.x, l = threadIdx.x;
  if (kdone[b] != 0) return;
  for (int i = l; i <= k; i += 64) {
    s_a[i] = alpha[(int64_t)b * kcap + i];
    s_b[i] = beta[(int64_t)b * (kcap + 1) + i];
  }
  double2* Gb = G + (int64_t)b * kcap * kcap;
  double2* c = cb + (int64_t)b * 2 * kcap;
  const bool brk = !(s_b[k] != 0.0);
  if (brk)
    for (int i = l; i < k; i += 64) {
      Gb[i * kcap + k] = make_double2(0.0, 0.0);
      Gb[k * kcap + i] = make_double2(0.0, 0.0);
    }
  wave_sync();
  const double conv = adaptive_metric<false>(k, brk, Gb, kcap, s_a, s_b, f, c, c + kcap);
  wave_sync();
  const bool stop = !(k + 1 < maxorder && conv > tol);
  if (stop) {
    for (int i = l; i <= k; i += 64) coef[(int64_t)b * kcap + i] = c[kcap + i];
    if (l == 0) {
      kdone[b] = k + 1;
      if (k + 1 >= maxorder && status) status[b] |= RM_NOT_CONVERGED;
    }
  } else if (l == 0) {
    atomicAdd(running, 1);
  }
}

__global__ void k_kp_record(const int32_t* __restrict__ kdone, int32_t ncols, int32_t* __restrict__ kused) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < ncols) kused[b] = kdone[b];
}

__global__ void k_kp_izero(int32_t* __restrict__ x, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

__global__ void k_kp_fill_i(int32_t* __restrict__ x, int32_t n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Observable partials per (J, wavepacket): conj(S) . Y over the J block, S
// the wavepackets and Y = O S, both in the split layout
__global__ __launch_bounds__(256) void k_kp_odot(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                                                 const double* __restrict__ S, const double* __restrict__ Y,
                                                 double* __restrict__ part, int32_t nJ) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  double a0 = 0.0, a1 = 0.0;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k, pi = pr + d;
    const double sr = S[pr], si = S[pi], yr = Y[pr], yi = Y[pi];
    a0 += sr * yr + si * yi;
    a1 += sr * yi - si * yr;
  }
  __shared__ double red[2][4];
  a0 = wave_allsum(a0);
  a1 = wave_allsum(a1);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a0;
    red[1][threadIdx.x >> 6] = a1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((int64_t)b * nJ + j) * 2] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    part[((int64_t)b * nJ + j) * 2 + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// ex[b] = <S_b|O|S_b>: the J partials of wavepacket b added in order
__global__ void k_kp_ofin(const double* __restrict__ part, int32_t nJ, int32_t ncols, double2* __restrict__ ex) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncols) return;
  double re = 0.0, im = 0.0;
  for (int j = 0; j < nJ; ++j) {
    re += part[((int64_t)b * nJ + j) * 2];
    im += part[((int64_t)b * nJ + j) * 2 + 1];
  }
  ex[b] = make_double2(re, im);
}

__global__ void k_kp_fill_d(double* __restrict__ x, int32_t n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// S = P sum_j c_j V_j (kf != NULL: wavepacket b uses its own dimension kf[b])
__global__ void k_kp_comb(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                          const double2* __restrict__ ph, const double* __restrict__ V, int64_t vstride,
                          const double2* __restrict__ coef, int32_t m, int32_t kcap, double* __restrict__ S,
                          const int32_t* __restrict__ kf) {
  const int j = blockIdx.x, b = blockIdx.y;
  if (kf) m = kf[b];
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k, pi = pr + d;
    double ur = 0.0, ui = 0.0;
    for (int q = 0; q < m; ++q) {
      const double2 c = coef[(int64_t)b * kcap + q];
      const double vr = V[q * vstride + pr], vi = V[q * vstride + pi];
      ur += c.x * vr - c.y * vi;
      ui += c.x * vi + c.y * vr;
    }
    const double2 p = ph ? ph[voff[j] + i] : make_double2(1.0, 0.0);
    S[pr] = p.x * ur - p.y * ui;
    S[pi] = p.x * ui + p.y * ur;
  }
}

// (B, n) complex rows in the internal order <-> split layout
__global__ void k_kp_pack(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js,
                          const double2* __restrict__ psi, int64_t ld, double* __restrict__ S, int32_t to_split) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k;
    double2* p = (double2*)psi + (int64_t)b * ld + voff[j] + i;
    if (to_split) {
      const double2 v = *p;
      S[pr] = v.x;
      S[pr + d] = v.y;
    } else {
      *p = make_double2(S[pr], S[pr + d]);
    }
  }
}

// matvec output (split layout, FP64 or FP32) -> complex rows
template <typename YT>
__global__ void k_kp_out(const int64_t* __restrict__ voff, const int32_t* __restrict__ Js, const YT* __restrict__ Y,
                         double2* __restrict__ psi, int64_t ld) {
  const int j = blockIdx.x, b = blockIdx.y;
  const int64_t d = jdim(Js[j]), len = 2 * d * d;
  const int64_t base = voff[j] * 2 * gridDim.y + (int64_t)b * len;
  for (int64_t i = threadIdx.x; i < d * d; i += blockDim.x) {
    const int64_t row = i / d, k = i % d;
    const int64_t pr = base + row * 2 * d + k;
    psi[(int64_t)b * ld + voff[j] + i] = make_double2((double)Y[pr], (double)Y[pr + d]);
  }
}

#define RBCHK(expr)                                                                    \
  do {                                                                                 \
    rocblas_status st_ = (expr);                                                       \
    if (st_ != rocblas_status_success)                                                 \
      return fail(RM_ERR_HIP, "%s failed: %s", #expr, rocblas_status_to_string(st_)); \
  } while (0)

void kplan_free(rm_kplan* p) {
  if (!p) return;
  void* bufs[] = {p->d_mdiag, p->d_sig, p->d_blk, p->d_mf, p->ws, p->d_tasks, p->d_voff, p->d_Js, p->d_fp, p->d_meta, p->d_kf64, p->d_kf32, p->d_colmap, p->d_fz,
                  p->d_ex, p->d_mm};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (double* k : p->kbig64)
    if (k) (void)hipFree(k);
  for (float* k : p->kbig32)
    if (k) (void)hipFree(k);
  if (p->rb) (void)rocblas_destroy_handle(p->rb);
  for (hipEvent_t e : p->ev)
    if (e) (void)hipEventDestroy(e);
  if (p->gs) (void)hipStreamDestroy(p->gs);
  delete p;
}

}  // namespace

extern "C" int rm_kplan_create(int32_t nJ, const int32_t* Js, int32_t nslot, int32_t nblk, const int32_t* J1,
                               const int32_t* J2, const int32_t* slot0, const int32_t* ncart,
                               const int64_t* k_off, const double* K, const int64_t* sig_off,
                               const int32_t* sig, const int64_t* d_off, const double* mdiag, rm_kplan** out) {
  if (!out) return fail(RM_ERR_ARG, "rm_kplan_create: out is NULL");
  *out = nullptr;
  if (nJ < 1 || nblk < 0 || nslot < 0 || !Js || (nblk > 0 && (!J1 || !J2 || !slot0 || !ncart || !k_off || !K ||
                                                              !sig_off || !sig || !d_off || !mdiag)))
    return fail(RM_ERR_ARG, "rm_kplan_create: bad arguments");
  rm_kplan* p = new (std::nothrow) rm_kplan();
  if (!p) return fail(RM_ERR_ALLOC, "rm_kplan_create: host allocation");
  p->nJ = nJ;
  p->nslot = nslot;
  p->nblk = nblk;
  p->Js.assign(Js, Js + nJ);
  std::vector<int> jidx(1024, -1);
  int64_t pos = 0;
  for (int j = 0; j < nJ; ++j) {
    if (Js[j] < 0 || Js[j] >= 1024 || (j > 0 && Js[j] <= Js[j - 1])) {
      kplan_free(p);
      return fail(RM_ERR_ARG, "rm_kplan_create: J list must be ascending in [0, 1024)");
    }
    jidx[Js[j]] = j;
    p->h_voff.push_back(pos);
    pos += jdim(Js[j]) * jdim(Js[j]);
  }
  p->n = pos;
  // Z-build tiles: as many rows as keep the staged X rows and MF values in 64 KiB of LDS
  {
    const int64_t dkmax = jdim(Js[nJ - 1]);
    p->zr = 0;
    for (int zr = 16; zr >= 1 && !p->zr; zr >>= 1) {
      const size_t lds = sizeof(double) * (size_t)(10 * zr + (zr + 4) * 2 * dkmax);
      if (lds <= 65536) {
        p->zr = zr;
        p->z_lds = lds;
      }
    }
    if (!p->zr) {
      kplan_free(p);
      return fail(RM_ERR_ARG, "rm_kplan_create: J = %d is too large for the Z-build tiles", Js[nJ - 1]);
    }
  }
  p->src.assign(nJ, {});
  p->ktot.assign(nJ, 0);
  p->kq_off.assign(nblk, 0);
  int64_t mfpos = 0;
  for (int q = 0; q < nblk; ++q) {
    const int a = J1[q] >= 0 && J1[q] < 1024 ? jidx[J1[q]] : -1;
    const int c = J2[q] >= 0 && J2[q] < 1024 ? jidx[J2[q]] : -1;
    const int64_t ns = sig_off[q + 1] - sig_off[q];
    if (a < 0 || c < 0 || ns < 1 || ncart[q] < 1 || slot0[q] < 0 || slot0[q] + ncart[q] > nslot ||
        k_off[q + 1] - k_off[q] != jdim(J1[q]) * jdim(J2[q]) ||
        d_off[q + 1] - d_off[q] != ns * ncart[q] * jdim(J1[q])) {
      kplan_free(p);
      return fail(RM_ERR_ARG, "rm_kplan_create: block %d is inconsistent", q);
    }
    for (int64_t t = sig_off[q]; t < sig_off[q + 1]; ++t)
      if (std::abs(sig[t]) > 2) {
        kplan_free(p);
        return fail(RM_ERR_ARG, "rm_kplan_create: block %d has |sigma| > 2", q);
      }
    p->src[a].push_back(q);
    KBlk bk;
    bk.dm1 = (int32_t)jdim(J1[q]);
    bk.nsig = (int32_t)ns;
    bk.ncart = ncart[q];
    bk.slot0 = slot0[q];
    bk.mf_off = mfpos;
    bk.d_off = d_off[q];
    mfpos += ns * jdim(J1[q]);
    p->h_blk.push_back(bk);
  }
  // per output J the source blocks ordered by J2 (those of one J2 share their
  // input rows in the Z build), stacked in that order
  for (int a = 0; a < nJ; ++a) {
    std::stable_sort(p->src[a].begin(), p->src[a].end(), [&](int x, int y) { return J2[x] < J2[y]; });
    for (int q : p->src[a]) {
      p->kq_off[q] = p->ktot[a];
      p->ktot[a] += (int32_t)jdim(J2[q]);
    }
  }
  p->h_J1.assign(J1, J1 + nblk);
  p->h_J2.assign(J2, J2 + nblk);
  p->h_sig_off.assign(sig_off, sig_off + nblk + 1);
  p->h_sig.assign(sig, sig + (nblk ? sig_off[nblk] : 0));
  p->mf_count = mfpos;
  int rc = RM_OK;
  const int64_t ndiag = nblk ? d_off[nblk] : 0;
  if ((rc = upload(&p->d_mdiag, (const double2*)mdiag, ndiag)) ||
      (rc = upload(&p->d_sig, p->h_sig.data(), (int64_t)p->h_sig.size())) ||
      (rc = upload(&p->d_blk, p->h_blk.data(), (int64_t)p->h_blk.size())) ||
      (rc = upload(&p->d_mf, (const double2*)nullptr, std::max<int64_t>(mfpos, 1))) ||
      (rc = upload(&p->d_voff, p->h_voff.data(), nJ)) || (rc = upload(&p->d_Js, p->Js.data(), nJ))) {
    kplan_free(p);
    return rc;
  }
  // per J1: the real form of the stacked K, [[Kr^T, Ki^T], [-Ki^T, Kr^T]] (2 Ktot x 2 dk1)
  p->kbig64.assign(nJ, nullptr);
  p->kbig32.assign(nJ, nullptr);
  for (int a = 0; a < nJ; ++a) {
    if (p->src[a].empty()) continue;
    const int64_t dk1 = jdim(Js[a]), kt = p->ktot[a], ncol = 2 * dk1;
    std::vector<double> kb((size_t)(2 * kt * ncol), 0.0);
    for (int q : p->src[a]) {
      const int64_t dk2 = jdim(J2[q]), ko = p->kq_off[q];
      const double* Kq = K + 2 * k_off[q];   // complex (dk1 x dk2) row-major
      for (int64_t k1 = 0; k1 < dk1; ++k1)
        for (int64_t k2 = 0; k2 < dk2; ++k2) {
          const double kr = Kq[2 * (k1 * dk2 + k2)], ki = Kq[2 * (k1 * dk2 + k2) + 1];
          const int64_t r0 = ko + k2, r1 = kt + ko + k2;
          kb[(size_t)(r0 * ncol + k1)] = kr;
          kb[(size_t)(r0 * ncol + dk1 + k1)] = ki;
          kb[(size_t)(r1 * ncol + k1)] = -ki;
          kb[(size_t)(r1 * ncol + dk1 + k1)] = kr;
        }
    }
    std::vector<float> kf(kb.begin(), kb.end());
    if ((rc = upload(&p->kbig64[a], kb.data(), (int64_t)kb.size())) ||
        (rc = upload(&p->kbig32[a], kf.data(), (int64_t)kf.size()))) {
      kplan_free(p);
      return rc;
    }
  }
  // fused form: padded stacked columns (groups of FGK, the total a multiple
  // of FBK / 2), interleaved group rows, output columns padded to FBN
  p->kpad.assign(nJ, 0);
  p->npad.assign(nJ, 0);
  p->kf_off.assign(nJ, 0);
  p->cm_off.assign(nJ, 0);
  int64_t kfpos = 0, cmpos = 0;
  for (int a = 0; a < nJ; ++a) {
    const int64_t dk1 = jdim(Js[a]);
    int64_t kp = 0;
    for (int q : p->src[a]) kp += (jdim(J2[q]) + FGK - 1) / FGK * FGK;
    kp = (kp + FBK / 2 - 1) / (FBK / 2) * (FBK / 2);
    if ((int64_t)p->src[a].size() > FMAXT) {
      kplan_free(p);
      return fail(RM_ERR_ARG, "rm_kplan_create: more than %d source blocks for J = %d", FMAXT, Js[a]);
    }
    p->kpad[a] = (int32_t)kp;
    p->npad[a] = (int32_t)((2 * dk1 + FBN - 1) / FBN * FBN);
    p->kf_off[a] = kfpos;
    p->cm_off[a] = cmpos;
    kfpos += 2 * kp * p->npad[a];
    cmpos += kp;
  }
  {
    std::vector<double> kf((size_t)std::max<int64_t>(kfpos, 1), 0.0);
    std::vector<int32_t> cm((size_t)std::max<int64_t>(cmpos, 1), -1);
    for (int a = 0; a < nJ; ++a) {
      const int64_t dk1 = jdim(Js[a]), np = p->npad[a];
      double* kb = kf.data() + p->kf_off[a];
      int64_t c0 = 0;
      for (size_t li = 0; li < p->src[a].size(); ++li) {
        const int q = p->src[a][li];
        const int64_t dk2 = jdim(J2[q]);
        const double* Kq = K + 2 * k_off[q];
        for (int64_t k2 = 0; k2 < dk2; ++k2) {
          const int64_t col = c0 + k2, g = col / FGK, i = col % FGK;
          cm[(size_t)(p->cm_off[a] + col)] = (int32_t)((li << 16) | k2);
          double* rre = kb + (2 * FGK * g + i) * np;
          double* rim = kb + (2 * FGK * g + FGK + i) * np;
          for (int64_t k1 = 0; k1 < dk1; ++k1) {
            const double kr = Kq[2 * (k1 * dk2 + k2)], ki = Kq[2 * (k1 * dk2 + k2) + 1];
            rre[k1] = kr;
            rre[dk1 + k1] = ki;
            rim[k1] = -ki;
            rim[dk1 + k1] = kr;
          }
        }
        c0 += (dk2 + FGK - 1) / FGK * FGK;
      }
    }
    std::vector<float> kf32(kf.begin(), kf.end());
    if ((rc = upload(&p->d_kf64, kf.data(), (int64_t)kf.size())) ||
        (rc = upload(&p->d_kf32, kf32.data(), (int64_t)kf32.size())) ||
        (rc = upload(&p->d_colmap, cm.data(), (int64_t)cm.size()))) {
      kplan_free(p);
      return rc;
    }
  }
  if (rocblas_create_handle(&p->rb) != rocblas_status_success) {
    p->rb = nullptr;
    kplan_free(p);
    return fail(RM_ERR_HIP, "rm_kplan_create: rocblas_create_handle failed");
  }
  p->ev.assign(KCHUNK + 1, nullptr);
  if (hipStreamCreateWithFlags(&p->gs, hipStreamNonBlocking) != hipSuccess) {
    p->gs = nullptr;
    kplan_free(p);
    return fail(RM_ERR_HIP, "rm_kplan_create: stream creation failed");
  }
  for (auto& e : p->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      e = nullptr;
      kplan_free(p);
      return fail(RM_ERR_HIP, "rm_kplan_create: event creation failed");
    }
  if (rocblas_set_stream(p->rb, p->gs) != rocblas_status_success) {
    kplan_free(p);
    return fail(RM_ERR_HIP, "rm_kplan_create: rocblas_set_stream failed");
  }
  *out = p;
  return RM_OK;
}

extern "C" int rm_kplan_destroy(rm_kplan* p) {
  kplan_free(p);
  return RM_OK;
}

extern "C" int rm_kplan_info(const rm_kplan* p, int64_t* n, double* gemm_flops_per_col) {
  if (!p) return fail(RM_ERR_ARG, "rm_kplan_info: NULL plan");
  if (n) *n = p->n;
  if (gemm_flops_per_col) {
    double fl = 0.0;
    for (int a = 0; a < p->nJ; ++a) {
      const double d = (double)jdim(p->Js[a]);
      fl += 2.0 * d * (2.0 * d) * (2.0 * p->ktot[a]);
    }
    *gemm_flops_per_col = fl;
  }
  return RM_OK;
}

namespace {

struct KWork {
  double *S, *W, *V, *ib, *part, *beta;
  double2 *G, *cb;          // adaptive rule: Gram matrices, (c, c_prev) per wavepacket
  int32_t *kdone, *running;
  double2 *alpha, *coef;
  void *Y, *Z;
  int64_t vreal;   // reals per vector (all wavepackets)
};

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace for ncols wavepackets and Krylov dimension m (m = 0: a matvec only);
// adaptive: the rule's Gram matrices and coefficient buffers too.
int kwork(rm_kplan* p, int32_t ncols, int32_t m, bool f32, KWork* w, bool adaptive = false) {
  const int64_t vreal = 2 * p->n * ncols;
  int64_t zreal = 0;
  if (p->kernel != RM_KKERNEL_FUSED)
    for (int a = 0; a < p->nJ; ++a) zreal += (int64_t)ncols * jdim(p->Js[a]) * 2 * p->ktot[a];
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  const size_t mm = (size_t)std::max(m, 1);
  const size_t b_vec = al256(sizeof(double) * (size_t)vreal);
  const size_t b_y = al256(es * (size_t)vreal), b_z = al256(es * (size_t)std::max<int64_t>(zreal, 1));
  const size_t b_col = al256(sizeof(double) * (size_t)ncols);
  const size_t b_part = al256(sizeof(double) * 2 * (size_t)ncols * p->nJ);
  const size_t b_beta = al256(sizeof(double) * (size_t)ncols * (mm + 1));
  const size_t b_al = al256(sizeof(double2) * (size_t)ncols * mm);
  const size_t b_g = adaptive ? al256(sizeof(double2) * (size_t)ncols * mm * mm) : 0;
  const size_t b_cb = adaptive ? al256(sizeof(double2) * (size_t)ncols * 2 * mm) : 0;
  const size_t b_kd = adaptive ? al256(sizeof(int32_t) * ((size_t)ncols + 1)) : 0;
  const size_t need = b_vec * (2 + mm) + b_y + b_z + b_col + b_part + b_beta + 2 * b_al + b_g + b_cb + b_kd;
  if (p->ws_bytes < need) {
    if (p->ws) {
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipFree(p->ws));
      p->ws = nullptr;
      p->ws_bytes = 0;
    }
    HIPCHK(hipMalloc(&p->ws, need));
    p->ws_bytes = need;
  }
  char* c = (char*)p->ws;
  w->S = (double*)c; c += b_vec;
  w->W = (double*)c; c += b_vec;
  w->V = (double*)c; c += b_vec * mm;
  w->Y = c; c += b_y;
  w->Z = c; c += b_z;
  w->ib = (double*)c; c += b_col;
  w->part = (double*)c; c += b_part;
  w->beta = (double*)c; c += b_beta;
  w->alpha = (double2*)c; c += b_al;
  w->coef = (double2*)c; c += b_al;
  w->G = adaptive ? (double2*)c : nullptr;
  c += b_g;
  w->cb = adaptive ? (double2*)c : nullptr;
  c += b_cb;
  w->kdone = adaptive ? (int32_t*)c : nullptr;
  w->running = adaptive ? (int32_t*)c + ncols : nullptr;
  c += b_kd;
  w->vreal = vreal;
  return RM_OK;
}

// The (output J1, source block) tasks for ncols wavepackets, uploaded once per call.
int ktasks(rm_kplan* p, int32_t ncols, int64_t* ntask, std::vector<int64_t>& zoff_J, hipStream_t s) {
  std::vector<KTask>& h = p->h_tasks;
  h.clear();
  zoff_J.assign(p->nJ, 0);
  int64_t zpos = 0;
  for (int a = 0; a < p->nJ; ++a) {
    zoff_J[a] = zpos;
    if (p->src[a].empty()) continue;
    const int64_t dm1 = jdim(p->Js[a]);
    for (int q : p->src[a]) {
      KTask t;
      t.J1 = p->h_J1[q];
      t.J2 = p->h_J2[q];
      t.dm1 = (int32_t)dm1;
      t.dm2 = (int32_t)jdim(t.J2);
      t.dk2 = t.dm2;
      t.koff = (int32_t)p->kq_off[q];
      t.ktot = p->ktot[a];
      t.nsig = (int32_t)(p->h_sig_off[q + 1] - p->h_sig_off[q]);
      t.sig0 = (int32_t)p->h_sig_off[q];
      t.blk = q;
      t.mf_off = p->h_blk[q].mf_off;
      t.zoff = zpos;
      const int jb = (int)(std::lower_bound(p->Js.begin(), p->Js.end(), t.J2) - p->Js.begin());
      t.xoff = p->h_voff[jb] * 2 * ncols;
      h.push_back(t);
    }
    zpos += (int64_t)ncols * dm1 * 2 * p->ktot[a];
  }
  *ntask = (int64_t)h.size();
  if (*ntask > INT32_MAX) return fail(RM_ERR_ARG, "too many operator blocks");
  // per J: first Z row and first task (tasks are grouped by output J in J order)
  p->h_meta.assign(2 * (size_t)(p->nJ + 1), 0);
  int64_t tiles = 0, tq = 0;
  for (int a = 0; a < p->nJ; ++a) {
    p->h_meta[a] = tiles;
    p->h_meta[p->nJ + 1 + a] = tq;
    if (!p->src[a].empty()) tiles += (int64_t)ncols * ((jdim(p->Js[a]) + p->zr - 1) / p->zr);
    tq += (int64_t)p->src[a].size();
  }
  p->h_meta[p->nJ] = tiles;
  p->h_meta[2 * p->nJ + 1] = tq;
  p->ztiles = tiles;
  // chunk bounds by Z tiles (ascending J): a short first chunk (its build is
  // not overlapped), then equal shares
  p->chunk_j.assign(KCHUNK + 1, p->nJ);
  p->chunk_j[0] = 0;
  for (int c = 1; c < KCHUNK; ++c) {
    const double share = 1.0 / 16.0 + (15.0 / 16.0) * (double)(c - 1) / (double)(KCHUNK - 1);
    const int64_t target = (int64_t)(share * (double)tiles);
    int a = p->chunk_j[c - 1];
    while (a < p->nJ && p->h_meta[a + 1] <= target) ++a;
    p->chunk_j[c] = std::max(a, p->chunk_j[c - 1]);
  }
  // fused matvec: J descriptors and the tile list (longest stacks first)
  {
    std::vector<FJ> fj((size_t)p->nJ);
    std::vector<int4> tl;
    for (int a = 0; a < p->nJ; ++a) {
      FJ& d = fj[(size_t)a];
      d.dm1 = (int32_t)jdim(p->Js[a]);
      d.N = 2 * d.dm1;
      d.Npad = p->npad[a];
      d.kpad = p->kpad[a];
      d.task0 = (int32_t)p->h_meta[p->nJ + 1 + a];
      d.ntask = (int32_t)p->src[a].size();
      d.M = (int64_t)ncols * d.dm1;
      d.cm_off = p->cm_off[a];
      d.kf_off = p->kf_off[a];
      d.yoff = p->h_voff[a] * 2 * ncols;
      const int64_t nrt = (d.M + FBM - 1) / FBM, nct = (d.N + FBN - 1) / FBN;
      if (nrt > INT32_MAX) return fail(RM_ERR_ARG, "too many fused tiles");
      for (int64_t rt = 0; rt < nrt; ++rt)
        for (int64_t ct = 0; ct < nct; ++ct) tl.push_back(make_int4(a, (int)rt, (int)ct, 0));
    }
    std::stable_sort(tl.begin(), tl.end(), [&](const int4& x, const int4& y) { return p->kpad[x.x] > p->kpad[y.x]; });
    p->ntiles = (int64_t)tl.size();
    if (p->ntiles > INT32_MAX) return fail(RM_ERR_ARG, "too many fused tiles");
    const size_t fjb = (sizeof(FJ) * fj.size() + 255) & ~(size_t)255;
    p->h_fz.assign(fjb + sizeof(int4) * tl.size(), 0);
    memcpy(p->h_fz.data(), fj.data(), sizeof(FJ) * fj.size());
    if (!tl.empty()) memcpy(p->h_fz.data() + fjb, tl.data(), sizeof(int4) * tl.size());
    int rc0 = grow(&p->d_fz, &p->fz_cap, (int64_t)p->h_fz.size());
    if (rc0) return rc0;
    HIPCHK(hipMemcpyAsync(p->d_fz, p->h_fz.data(), p->h_fz.size(), hipMemcpyHostToDevice, s));
    p->fz_tiles = (const int4*)((const char*)p->d_fz + fjb);
  }
  // stored form on the matrix cores: per chunk of output J its tiles (the
  // longest stacks first), the J descriptors and the K-stack pointers
  if (p->kernel == RM_KKERNEL_MFMA) {
    std::vector<MJ> mj((size_t)p->nJ);
    std::vector<int4> tl, tw;
    p->mm_chunk.assign(KCHUNK + 1, 0);
    p->mm_chunk_w.assign(KCHUNK + 1, 0);
    for (int c = 0; c < KCHUNK; ++c) {
      p->mm_chunk[c] = (int64_t)tl.size();
      p->mm_chunk_w[c] = (int64_t)tw.size();
      const size_t first = tl.size(), first_w = tw.size();
      for (int a = p->chunk_j[c]; a < p->chunk_j[c + 1]; ++a) {
        if (p->src[a].empty()) continue;
        MJ& d = mj[(size_t)a];
        const int64_t dm = jdim(p->Js[a]);
        if ((int64_t)ncols * dm > INT32_MAX || (int64_t)ncols * dm * 2 * p->ktot[a] > INT32_MAX)
          return fail(RM_ERR_ARG, "matrix-core product: %d wavepackets exceed its 32-bit offsets", ncols);
        d.M = (int32_t)(ncols * dm);
        d.N = (int32_t)(2 * dm);
        d.K = 2 * p->ktot[a];
        d.pad = 0;
        d.zoff = zoff_J[a];
        d.yoff = p->h_voff[a] * 2 * ncols;
        const int nrt = (d.M + FBM - 1) / FBM, nct = (d.N + FBN - 1) / FBN;
        for (int rt = 0; rt < nrt; ++rt)
          for (int ct = 0; ct < nct; ++ct) tl.push_back(make_int4(a, rt, ct, 0));
        const int nrw = (d.M + MBM - 1) / MBM, ncw = (d.N + MBN - 1) / MBN;
        for (int rt = 0; rt < nrw; ++rt)
          for (int ct = 0; ct < ncw; ++ct) tw.push_back(make_int4(a, rt, ct, 0));
      }
      const auto longer = [&](const int4& x, const int4& y) { return p->ktot[x.x] > p->ktot[y.x]; };
      std::stable_sort(tl.begin() + first, tl.end(), longer);
      std::stable_sort(tw.begin() + first_w, tw.end(), longer);
    }
    p->mm_chunk[KCHUNK] = (int64_t)tl.size();
    p->mm_chunk_w[KCHUNK] = (int64_t)tw.size();
    if ((int64_t)tl.size() > INT32_MAX) return fail(RM_ERR_ARG, "too many matrix-core product tiles");
    const size_t b_j = al256(sizeof(MJ) * mj.size()), b_p = al256(sizeof(void*) * (size_t)p->nJ);
    const size_t b_t = al256(sizeof(int4) * tl.size());
    p->h_mm.assign(b_j + 2 * b_p + b_t + sizeof(int4) * tw.size(), 0);
    memcpy(p->h_mm.data(), mj.data(), sizeof(MJ) * mj.size());
    memcpy(p->h_mm.data() + b_j, p->kbig64.data(), sizeof(void*) * (size_t)p->nJ);
    memcpy(p->h_mm.data() + b_j + b_p, p->kbig32.data(), sizeof(void*) * (size_t)p->nJ);
    if (!tl.empty()) memcpy(p->h_mm.data() + b_j + 2 * b_p, tl.data(), sizeof(int4) * tl.size());
    if (!tw.empty()) memcpy(p->h_mm.data() + b_j + 2 * b_p + b_t, tw.data(), sizeof(int4) * tw.size());
    int rc0 = grow(&p->d_mm, &p->mm_cap, (int64_t)p->h_mm.size());
    if (rc0) return rc0;
    HIPCHK(hipMemcpyAsync(p->d_mm, p->h_mm.data(), p->h_mm.size(), hipMemcpyHostToDevice, s));
    p->mm_j = (const MJ*)p->d_mm;
    p->mm_kb64 = (const double* const*)(p->d_mm + b_j);
    p->mm_kb32 = (const float* const*)(p->d_mm + b_j + b_p);
    p->mm_tiles = (const int4*)(p->d_mm + b_j + 2 * b_p);
    p->mm_tiles_w = (const int4*)(p->d_mm + b_j + 2 * b_p + b_t);
  }
  int rc = grow(&p->d_tasks, &p->tasks_cap, std::max<int64_t>(*ntask, 1));
  if (rc) return rc;
  if ((rc = grow(&p->d_meta, &p->meta_cap, (int64_t)p->h_meta.size()))) return rc;
  if (*ntask) HIPCHK(hipMemcpyAsync(p->d_tasks, h.data(), sizeof(KTask) * h.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(p->d_meta, p->h_meta.data(), sizeof(int64_t) * p->h_meta.size(), hipMemcpyHostToDevice,
                        s));
  return RM_OK;
}

__global__ void k_kp_zero(double* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

// Y = H(F) (ib X) for the split-layout input X; the MF values of the step are
// in p->d_mf.  The Z build runs on s, chunk by chunk; the dense products of a
// chunk run on the plan's product stream once that chunk's Z is done, and s
// waits for the last of them.
int kmatvec(rm_kplan* p, const double* X, const double* ib, void* Y, void* Z, int32_t ncols, int64_t ntask,
            const std::vector<int64_t>& zoff_J, bool f32, hipStream_t s) {
  if (p->kernel == RM_KKERNEL_FUSED) {
    if (2 * p->n * (int64_t)ncols >= INT32_MAX || p->mf_count >= INT32_MAX)
      return fail(RM_ERR_ARG, "fused matvec: %lld wavepackets of %lld states exceed its 32-bit offsets",
                  (long long)ncols, (long long)p->n);
    if (p->ntiles > 0) {
      const FJ* jd = (const FJ*)p->d_fz;
      if (f32)
        hipLaunchKernelGGL(k_kp_fused<float>, dim3((unsigned)p->ntiles), dim3(512), 0, s, p->fz_tiles, jd,
                           p->d_tasks, p->d_sig, p->d_colmap, p->d_mf, X, ib, p->d_kf32, (float*)Y);
      else
        hipLaunchKernelGGL(k_kp_fused<double>, dim3((unsigned)p->ntiles), dim3(512), 0, s, p->fz_tiles, jd,
                           p->d_tasks, p->d_sig, p->d_colmap, p->d_mf, X, ib, p->d_kf64, (double*)Y);
      HIPCHK(hipGetLastError());
    }
    return RM_OK;
  }
  if (p->ztiles > INT32_MAX) return fail(RM_ERR_ARG, "too many Z tiles (%lld)", (long long)p->ztiles);
  for (int c = 0; c < KCHUNK; ++c) {
    const int j0 = p->chunk_j[c], j1 = p->chunk_j[c + 1];
    const int64_t t0 = p->h_meta[j0], nt = p->h_meta[j1] - t0;
    if (ntask && nt > 0) {
      if (f32)
        hipLaunchKernelGGL(k_kp_z<float>, dim3((unsigned)nt), dim3(256), p->z_lds, s, p->d_tasks, p->d_meta, p->nJ,
                           p->d_sig, p->d_mf, X, ib, (float*)Z, p->zr, t0);
      else
        hipLaunchKernelGGL(k_kp_z<double>, dim3((unsigned)nt), dim3(256), p->z_lds, s, p->d_tasks, p->d_meta,
                           p->nJ, p->d_sig, p->d_mf, X, ib, (double*)Z, p->zr, t0);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(p->ev[c], s));
    HIPCHK(hipStreamWaitEvent(p->gs, p->ev[c], 0));
    if (p->kernel == RM_KKERNEL_MFMA) {
      for (int a = j0; a < j1; ++a) {
        if (!p->src[a].empty()) continue;
        // no source block reaches this J: its output rows are zero
        const int64_t d = jdim(p->Js[a]), yoff = p->h_voff[a] * 2 * ncols;
        const int64_t cnt = f32 ? (int64_t)ncols * d * d : 2 * (int64_t)ncols * d * d;   // in doubles
        double* y0 = f32 ? (double*)((float*)Y + yoff) : (double*)Y + yoff;
        hipLaunchKernelGGL(k_kp_zero, dim3(grid_for(cnt, 256)), dim3(256), 0, p->gs, y0, cnt);
      }
      const std::vector<int64_t>& co = f32 ? p->mm_chunk_w : p->mm_chunk;
      const int64_t m0 = co[c], mt = co[c + 1] - m0;
      if (mt > 0) {
        if (f32)
          hipLaunchKernelGGL(k_kp_mm_w<float>, dim3((unsigned)mt), dim3(512), 0, p->gs, p->mm_tiles_w + m0,
                             p->mm_j, p->mm_kb32, (const float*)Z, (float*)Y);
        else
          hipLaunchKernelGGL(k_kp_mm<double>, dim3((unsigned)mt), dim3(512), 0, p->gs, p->mm_tiles + m0, p->mm_j,
                             p->mm_kb64, (const double*)Z, (double*)Y);
      }
      continue;
    }
    for (int a = j0; a < j1; ++a) {
      const int64_t d = jdim(p->Js[a]);
      const int64_t yoff = p->h_voff[a] * 2 * ncols;
      if (p->src[a].empty()) {
        // no source block reaches this J: its output rows are zero
        const int64_t cnt = f32 ? (int64_t)ncols * d * d : 2 * (int64_t)ncols * d * d;   // in doubles
        double* y0 = f32 ? (double*)((float*)Y + yoff) : (double*)Y + yoff;
        hipLaunchKernelGGL(k_kp_zero, dim3(grid_for(cnt, 256)), dim3(256), 0, p->gs, y0, cnt);
        continue;
      }
      const rocblas_int M = (rocblas_int)(ncols * d), N = (rocblas_int)(2 * d), K = (rocblas_int)(2 * p->ktot[a]);
      if (f32) {
        const float one = 1.0f, zero = 0.0f;
        RBCHK(rocblas_sgemm(p->rb, rocblas_operation_none, rocblas_operation_none, N, M, K, &one, p->kbig32[a], N,
                            (const float*)Z + zoff_J[a], K, &zero, (float*)Y + yoff, N));
      } else {
        const double one = 1.0, zero = 0.0;
        RBCHK(rocblas_dgemm(p->rb, rocblas_operation_none, rocblas_operation_none, N, M, K, &one, p->kbig64[a], N,
                            (const double*)Z + zoff_J[a], K, &zero, (double*)Y + yoff, N));
      }
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(p->ev[KCHUNK], p->gs));
  HIPCHK(hipStreamWaitEvent(s, p->ev[KCHUNK], 0));
  return RM_OK;
}

// Launch helpers of the chunk kernels: grid (nJ, ncols), 256 threads.
dim3 chunk_grid(const rm_kplan* p, int32_t ncols) { return dim3((unsigned)p->nJ, (unsigned)ncols); }

}  // namespace

// y = H(F) x for ncols wavepackets given as rows (stride ld) of complex
// values in the plan's internal order; fprod: device, nslot complex.
extern "C" int rm_kplan_matvec(rm_kplan* p, const void* x, void* y, int64_t ld, int32_t ncols,
                               const double* fprod_dev, int32_t precision, void* stream) {
  if (!p) return fail(RM_ERR_ARG, "rm_kplan_matvec: NULL plan");
  if (ncols < 1 || ncols > 65535 || ld < p->n || !x || !y || (p->nslot > 0 && !fprod_dev))
    return fail(RM_ERR_ARG, "rm_kplan_matvec: bad arguments");
  if (precision != RM_PREC_FP64 && precision != RM_PREC_FP32)
    return fail(RM_ERR_ARG, "rm_kplan_matvec: unknown precision %d", precision);
  hipStream_t s = (hipStream_t)stream;
  const bool f32 = precision == RM_PREC_FP32;
  KWork w;
  int rc = kwork(p, ncols, 0, f32, &w);
  if (rc) return rc;
  int64_t ntask = 0;
  std::vector<int64_t> zoff_J;
  if ((rc = ktasks(p, ncols, &ntask, zoff_J, s))) return rc;
  if (p->nblk)
    hipLaunchKernelGGL(k_kp_mf, dim3(8, (unsigned)p->nblk), dim3(256), 0, s, p->d_blk, p->d_mdiag,
                       (const double2*)fprod_dev, p->d_mf);
  hipLaunchKernelGGL(k_kp_pack, chunk_grid(p, ncols), dim3(256), 0, s, p->d_voff, p->d_Js, (const double2*)x, ld,
                     w.S, 1);
  hipLaunchKernelGGL(k_kp_start, chunk_grid(p, ncols), dim3(256), 0, s, p->d_voff, p->d_Js,
                     (const double2*)nullptr, w.S, w.W, w.ib, w.beta, 0, (double*)nullptr, p->nJ);
  if ((rc = kmatvec(p, w.W, w.ib, w.Y, w.Z, ncols, ntask, zoff_J, f32, s))) return rc;
  if (f32)
    hipLaunchKernelGGL(k_kp_out<float>, chunk_grid(p, ncols), dim3(256), 0, s, p->d_voff, p->d_Js,
                       (const float*)w.Y, (double2*)y, ld);
  else
    hipLaunchKernelGGL(k_kp_out<double>, chunk_grid(p, ncols), dim3(256), 0, s, p->d_voff, p->d_Js,
                       (const double*)w.Y, (double2*)y, ld);
  HIPCHK(hipGetLastError());
  return RM_OK;
}

// nsteps split-operator Lanczos steps (tdse.py:365-414, 417-486) for ncols
// wavepackets: a fixed Krylov dimension, or the reference's adaptive rule
// (krylov_fixed == 0); see the header for the arguments.
static int kplan_propagate(rm_kplan* p, void* psi, int64_t ld, int32_t ncols, int64_t nsteps,
                           const double* fprod, double f_re, double f_im, const void* phase,
                           int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                           int32_t* status, int32_t precision, rm_kplan* obs, const double* obs_fprod,
                           double2* expect, double2* expect_sum, hipStream_t s);

extern "C" int rm_kplan_propagate(rm_kplan* p, void* psi, int64_t ld, int32_t ncols, int64_t nsteps,
                                  const double* fprod, double f_re, double f_im, const void* phase,
                                  int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                                  int32_t* status, int32_t precision, void* stream) {
  return kplan_propagate(p, psi, ld, ncols, nsteps, fprod, f_re, f_im, phase, krylov_fixed, maxorder, tol, kused,
                         status, precision, nullptr, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int rm_kplan_propagate_expect(rm_kplan* p, void* psi, int64_t ld, int32_t ncols, int64_t nsteps,
                                         const double* fprod, double f_re, double f_im, const void* phase,
                                         int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                                         int32_t* status, int32_t precision, rm_kplan* obs,
                                         const double* obs_fprod, void* expect, void* expect_sum, void* stream) {
  if (!p || !obs) return fail(RM_ERR_ARG, "rm_kplan_propagate_expect: NULL plan or observable");
  if (obs == p) return fail(RM_ERR_ARG, "rm_kplan_propagate_expect: the observable must be a plan of its own");
  if (obs->n != p->n || obs->Js != p->Js)
    return fail(RM_ERR_ARG, "rm_kplan_propagate_expect: the observable's basis differs from the plan's");
  if (obs->nslot > 0 && !obs_fprod) return fail(RM_ERR_ARG, "rm_kplan_propagate_expect: NULL obs_fprod");
  if (!expect && !expect_sum) return fail(RM_ERR_ARG, "rm_kplan_propagate_expect: NULL expect and expect_sum");
  return kplan_propagate(p, psi, ld, ncols, nsteps, fprod, f_re, f_im, phase, krylov_fixed, maxorder, tol, kused,
                         status, precision, obs, obs_fprod, (double2*)expect, (double2*)expect_sum,
                         (hipStream_t)stream);
}

static int kplan_propagate(rm_kplan* p, void* psi, int64_t ld, int32_t ncols, int64_t nsteps,
                           const double* fprod, double f_re, double f_im, const void* phase,
                           int32_t krylov_fixed, int32_t maxorder, double tol, int32_t* kused,
                           int32_t* status, int32_t precision, rm_kplan* obs, const double* obs_fprod,
                           double2* expect, double2* expect_sum, hipStream_t s) {
  if (!p) return fail(RM_ERR_ARG, "rm_kplan_propagate: NULL plan");
  if (ncols < 1 || ncols > 65535 || ld < p->n || !psi || nsteps < 0 || (nsteps > 0 && p->nslot > 0 && !fprod))
    return fail(RM_ERR_ARG, "rm_kplan_propagate: bad arguments");
  const bool adaptive = krylov_fixed == 0;
  if (krylov_fixed < 0 || krylov_fixed > KMAX)
    return fail(RM_ERR_ARG, "rm_kplan_propagate: krylov_fixed must be in [0, %d]", KMAX);
  if (adaptive && (maxorder < 1 || maxorder > KMAX || !(tol > 0.0)))
    return fail(RM_ERR_ARG, "rm_kplan_propagate: maxorder must be in [1, %d] and tol > 0", KMAX);
  if (precision != RM_PREC_FP64 && precision != RM_PREC_FP32)
    return fail(RM_ERR_ARG, "rm_kplan_propagate: unknown precision %d", precision);
  const bool f32 = precision == RM_PREC_FP32;
  const int32_t m = adaptive ? maxorder : krylov_fixed;   // Krylov vectors held
  if (status)
    hipLaunchKernelGGL(k_kp_izero, dim3(grid_for(ncols, 256)), dim3(256), 0, s, status, ncols);
  if (nsteps == 0) return RM_OK;
  KWork w;
  int rc = kwork(p, ncols, m, f32, &w, adaptive);
  if (rc) return rc;
  int64_t ntask = 0;
  std::vector<int64_t> zoff_J;
  if ((rc = ktasks(p, ncols, &ntask, zoff_J, s))) return rc;
  // the field products of every step, uploaded once
  const int64_t nf = std::max<int32_t>(p->nslot, 1);
  if ((rc = grow(&p->d_fp, &p->fp_cap, nsteps * nf))) return rc;
  if (p->nslot)
    HIPCHK(hipMemcpyAsync(p->d_fp, fprod, sizeof(double2) * nsteps * p->nslot, hipMemcpyHostToDevice, s));
  const double2* ph = (const double2*)phase;
  const double2 f = make_double2(f_re, f_im);
  const dim3 cg = chunk_grid(p, ncols), cb(256);
  const dim3 fin_grid((unsigned)((ncols + 3) / 4)), fin_block(256);
  const int64_t vs = w.vreal;
  hipLaunchKernelGGL(k_kp_pack, cg, cb, 0, s, p->d_voff, p->d_Js, (const double2*)psi, ld, w.S, 1);
  // observable: its operator (fixed products) formed once; after every step
  // Y = O S in FP64 on its own workspace, then <S_b|Y_b> per wavepacket
  KWork wo;
  int64_t ntask_o = 0;
  std::vector<int64_t> zoff_o;
  double2* ex = expect;
  if (obs) {
    if ((rc = kwork(obs, ncols, 0, false, &wo))) return rc;
    if ((rc = ktasks(obs, ncols, &ntask_o, zoff_o, s))) return rc;
    if (!ex) {
      if ((rc = grow(&obs->d_ex, &obs->ex_cap, nsteps * ncols))) return rc;
      ex = obs->d_ex;
    }
    if ((rc = grow(&obs->d_fp, &obs->fp_cap, std::max<int32_t>(obs->nslot, 1)))) return rc;
    if (obs->nslot)
      HIPCHK(hipMemcpyAsync(obs->d_fp, obs_fprod, sizeof(double2) * obs->nslot, hipMemcpyHostToDevice, s));
    if (obs->nblk)
      hipLaunchKernelGGL(k_kp_mf, dim3(8, (unsigned)obs->nblk), dim3(256), 0, s, obs->d_blk, obs->d_mdiag,
                         obs->d_fp, obs->d_mf);
    hipLaunchKernelGGL(k_kp_fill_d, dim3(grid_for(ncols, 256)), dim3(256), 0, s, wo.ib, ncols, 1.0);
    HIPCHK(hipGetLastError());
  }
  auto observe = [&](int64_t st) -> int {
    if (!obs) return RM_OK;
    int r = kmatvec(obs, w.S, wo.ib, wo.Y, wo.Z, ncols, ntask_o, zoff_o, false, s);
    if (r) return r;
    hipLaunchKernelGGL(k_kp_odot, cg, cb, 0, s, p->d_voff, p->d_Js, (const double*)w.S, (const double*)wo.Y,
                       wo.part, p->nJ);
    hipLaunchKernelGGL(k_kp_ofin, dim3(grid_for(ncols, 256)), dim3(256), 0, s, (const double*)wo.part, p->nJ,
                       ncols, ex + st * ncols);
    HIPCHK(hipGetLastError());
    return RM_OK;
  };
  for (int64_t st = 0; st < nsteps; ++st) {
    bool field = false;
    for (int32_t c = 0; c < p->nslot && !field; ++c)
      field = fprod[2 * (st * p->nslot + c)] != 0.0 || fprod[2 * (st * p->nslot + c) + 1] != 0.0;
    if (!field) {
      if (ph) hipLaunchKernelGGL(k_kp_phase2, cg, cb, 0, s, p->d_voff, p->d_Js, ph, w.S);
      if (kused)
        hipLaunchKernelGGL(k_kp_izero, dim3(grid_for(ncols, 256)), dim3(256), 0, s, kused + st * ncols, ncols);
      if ((rc = observe(st))) return rc;
      continue;
    }
    if (p->nblk)
      hipLaunchKernelGGL(k_kp_mf, dim3(8, (unsigned)p->nblk), dim3(256), 0, s, p->d_blk, p->d_mdiag,
                         p->d_fp + st * p->nslot, p->d_mf);
    hipLaunchKernelGGL(k_kp_start, cg, cb, 0, s, p->d_voff, p->d_Js, ph, w.S, w.W, w.ib, w.beta, m,
                       adaptive ? w.part : (double*)nullptr, p->nJ);
    const bool rule = adaptive && m > 1;   // maxorder 1: V_0 itself, as a fixed dimension
    if (rule)
      hipLaunchKernelGGL(k_kp_fin, fin_grid, fin_block, 0, s, w.part, p->nJ, ncols, 2, 0, m, w.alpha, w.beta, w.ib,
                         w.G, w.cb, w.kdone);
    int32_t kdim = m;
    for (int32_t it = 0; it < m; ++it) {
      double* Vit = w.V + it * vs;
      const double* Vprev = it > 0 ? w.V + (it - 1) * vs : Vit;
      if ((rc = kmatvec(p, w.W, w.ib, w.Y, w.Z, ncols, ntask, zoff_J, f32, s))) return rc;
      if (f32)
        hipLaunchKernelGGL(k_kp_dot<float>, cg, cb, 0, s, p->d_voff, p->d_Js, w.W, w.ib, (const float*)w.Y, Vit,
                           w.part, p->nJ);
      else
        hipLaunchKernelGGL(k_kp_dot<double>, cg, cb, 0, s, p->d_voff, p->d_Js, w.W, w.ib, (const double*)w.Y, Vit,
                           w.part, p->nJ);
      hipLaunchKernelGGL(k_kp_fin, fin_grid, fin_block, 0, s, w.part, p->nJ, ncols, 0, it, m, w.alpha, w.beta, w.ib,
                         w.G, w.cb, w.kdone);
      if (rule && it >= 1) {
        // the stopping rule for dimension it + 1; the host waits for the count of wavepackets going on
        hipLaunchKernelGGL(k_kp_izero, dim3(1), dim3(64), 0, s, w.running, 1);
        hipLaunchKernelGGL(k_kp_conv, dim3((unsigned)ncols), dim3(64), 0, s, w.alpha, w.beta, m, it, f, tol,
                           maxorder, w.G, w.cb, w.kdone, w.coef, status, w.running);
        int32_t h_running = 0;
        HIPCHK(hipMemcpyAsync(&h_running, w.running, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h_running == 0) {
          kdim = it + 1;
          break;
        }
      }
      if (it + 1 == m) break;
      if (f32)
        hipLaunchKernelGGL(k_kp_upd<float>, cg, cb, 0, s, p->d_voff, p->d_Js, (const float*)w.Y, Vit, Vprev,
                           w.alpha, w.beta, it, m, w.W, w.part, p->nJ);
      else
        hipLaunchKernelGGL(k_kp_upd<double>, cg, cb, 0, s, p->d_voff, p->d_Js, (const double*)w.Y, Vit, Vprev,
                           w.alpha, w.beta, it, m, w.W, w.part, p->nJ);
      hipLaunchKernelGGL(k_kp_fin, fin_grid, fin_block, 0, s, w.part, p->nJ, ncols, 1, it, m, w.alpha, w.beta, w.ib,
                         w.G, w.cb, w.kdone);
    }
    if (rule) {
      hipLaunchKernelGGL(k_kp_comb, cg, cb, 0, s, p->d_voff, p->d_Js, ph, w.V, vs, w.coef, kdim, m, w.S, w.kdone);
      if (kused)
        hipLaunchKernelGGL(k_kp_record, dim3(grid_for(ncols, 256)), dim3(256), 0, s, w.kdone, ncols,
                           kused + st * ncols);
    } else {
      hipLaunchKernelGGL(k_kp_expm, dim3((unsigned)ncols), dim3(64), 0, s, w.alpha, w.beta, m, m, f, w.coef);
      hipLaunchKernelGGL(k_kp_comb, cg, cb, 0, s, p->d_voff, p->d_Js, ph, w.V, vs, w.coef, m, m, w.S,
                         (const int32_t*)nullptr);
      if (kused)
        hipLaunchKernelGGL(k_kp_fill_i, dim3(grid_for(ncols, 256)), dim3(256), 0, s, kused + st * ncols, ncols, m);
      if (adaptive && status)   // maxorder 1 is reached at once
        hipLaunchKernelGGL(k_kp_fill_i, dim3(grid_for(ncols, 256)), dim3(256), 0, s, status, ncols,
                           RM_NOT_CONVERGED);
    }
    HIPCHK(hipGetLastError());
    if ((rc = observe(st))) return rc;
  }
  hipLaunchKernelGGL(k_kp_pack, cg, cb, 0, s, p->d_voff, p->d_Js, (const double2*)psi, ld, w.S, 0);
  HIPCHK(hipGetLastError());
  if (obs && expect_sum) return launch_expect_sum(nsteps, ncols, ex, expect_sum, s);
  return RM_OK;
}

extern "C" int rm_kplan_set_kernel(rm_kplan* p, int32_t kernel) {
  if (!p) return fail(RM_ERR_ARG, "rm_kplan_set_kernel: NULL plan");
  if (kernel != RM_KKERNEL_FUSED && kernel != RM_KKERNEL_GEMM && kernel != RM_KKERNEL_MFMA)
    return fail(RM_ERR_ARG, "rm_kplan_set_kernel: unknown kernel %d", kernel);
  p->kernel = kernel;
  return RM_OK;
}
