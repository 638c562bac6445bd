// Nodal vector kernels, CSR SpMV for the coarse AMG levels, and the preconditioned CG
// for the demagnetising potential (``make_pcg_solve`` + ``make_jax_amgcl_vcycle`` /
// ``make_jax_amg_vcycle`` semantics, src/poisson_solve.py:422-478, src/amg_utils.py:491-623).
//
// All PCG scalars live in device memory and the reductions need no fences: a producing
// kernel writes one partial per workgroup, and the consuming kernel (every one of its
// workgroups, in the same fixed order) sums them in its prologue before forming alpha or
// beta. p.Ap comes out of the Poisson apply itself (per-block element energies). The host
// never waits inside an iteration except to sum the r.r partials for the stopping test, and
// that copy overlaps the V-cycle of the same iteration.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include <functional>

#include "mm_device.h"
#include "mm_solver.h"

namespace mm {

// ------------------------------------------------------------------------ vector kernels
__global__ void k_fill(double* __restrict__ x, double v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_fill_f(float* __restrict__ x, float v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) d[i] = s[i];
}

// plain copy under its own name: the known-size reference of the HBM byte counters
__global__ void k_copy_calib(double* __restrict__ d, const double* __restrict__ s, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) d[i] = s[i];
}

// per-workgroup partials of a.b (no cross-workgroup traffic)
__global__ __launch_bounds__(kBlock) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b, long n,
                                                     double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return double2{a[i], b[i]}; }, [&](long, double2 u) { v[0] += u.x * u.y; });
  block_partial<1>(v, part);
}

static inline int vgrid(long n) {
  long g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  return (int)(g < 1 ? 1 : g);
}

int vfill(double* x, double v, long n, hipStream_t s) {
  if (n > 0) k_fill<<<vgrid(n), kBlock, 0, s>>>(x, v, n);
  return 0;
}

int vfill_f(float* x, float v, long n, hipStream_t s) {
  if (n > 0) k_fill_f<<<vgrid(n), kBlock, 0, s>>>(x, v, n);
  return 0;
}
int launch_copy_calib(double* d, const double* src, long n, hipStream_t s) {
  if (n > 0) k_copy_calib<<<vgrid(n), kBlock, 0, s>>>(d, src, n);
  return 0;
}

int vcopy(double* d, const double* src, long n, hipStream_t s) {
  if (n > 0) k_copy<<<vgrid(n), kBlock, 0, s>>>(d, src, n);
  return 0;
}

// ----------------------------------------------------------------------------- PCG kernels
// r = b - A x0 (A x0 in q); partials of r.r and b.b (2 per workgroup) for the host.
// t / tf (the preconditioner's first apply output) are cleared on the way.
__global__ __launch_bounds__(kBlock) void k_pcg_init(const double* __restrict__ b, const double* __restrict__ q,
                                                     double* __restrict__ r, double* __restrict__ t,
                                                     float* __restrict__ tf, long n, double* __restrict__ part_rr) {
  double v[2] = {0.0, 0.0};
  grid_chunks(
      n, [&](long i) { return double2{b[i], q[i]}; },
      [&](long i, double2 u) {
        const double ri = u.x - u.y;
        r[i] = ri;
        if (t) t[i] = 0.0;
        if (tf) tf[i] = 0.0f;
        v[0] += ri * ri;
        v[1] += u.x * u.x;
      });
  block_partial<2>(v, part_rr);
}

// alpha = rz / (p.q + 1e-30) with p.q summed from n_pq partials; x += alpha p; r -= alpha q;
// partials of r.r; t / tf cleared
__global__ __launch_bounds__(kBlock) void k_pcg_xr(const double* __restrict__ sc, int rz_slot,
                                                   const double* __restrict__ part_pq, int n_pq,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   double* __restrict__ t, float* __restrict__ tf, long n,
                                                   double* __restrict__ part_rr) {
  const double alpha = sc[rz_slot] / (sum_partials_all<kBlock>(part_pq, n_pq) + 1e-30);
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return double4{x[i], p[i], r[i], q[i]}; },
      [&](long i, double4 u) {
        if (t) t[i] = 0.0;
        if (tf) tf[i] = 0.0f;
        x[i] = u.x + alpha * u.y;
        const double ri = u.z - alpha * u.w;
        r[i] = ri;
        v[0] += ri * ri;
      });
  block_partial<1>(v, part_rr);
}

// rz_new = sum of n_rz partials, stored to sc[new_slot] by workgroup 0;
// beta = rz_new / (sc[old_slot] + 1e-30); p = z + beta p (first: p = z); q = 0 for the next apply
__global__ __launch_bounds__(kBlock) void k_pcg_p(double* __restrict__ sc, const double* __restrict__ part_rz,
                                                  int n_rz, int new_slot, int old_slot, int first,
                                                  const double* __restrict__ z, double* __restrict__ p,
                                                  double* __restrict__ q, long n) {
  const double rz = sum_partials_all<kBlock>(part_rz, n_rz);
  const double beta = first ? 0.0 : rz / (sc[old_slot] + 1e-30);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[new_slot] = rz;
  // the branch on first stays outside the loops: inside, the compiler waits for each load
  if (first)
    grid_chunks(
        n, [&](long i) { return z[i]; },
        [&](long i, double u) {
          p[i] = u;
          q[i] = 0.0;
        });
  else
    grid_chunks(
        n, [&](long i) { return double2{z[i], p[i]}; },
        [&](long i, double2 u) {
          p[i] = u.x + beta * u.y;
          q[i] = 0.0;
        });
}

// x = M .* b (and t = 0 when t != nullptr)
__global__ void k_scale(const double* __restrict__ M, const double* __restrict__ b, double* __restrict__ x,
                        double* __restrict__ t, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[i] = M[i] * b[i];
    if (t) t[i] = 0.0;
  }
}

// z = x + M .* (b - t), and partials of r.z with r == b (level-0 post-smoothing fused with the dot)
__global__ __launch_bounds__(kBlock) void k_post_dot(const double* __restrict__ x, const double* __restrict__ M,
                                                     const double* __restrict__ b, const double* __restrict__ t,
                                                     double* __restrict__ z, long n, double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return double4{x[i], M[i], b[i], t[i]}; },
      [&](long i, double4 u) {
        const double zi = u.x + u.y * (u.z - u.w);
        z[i] = zi;
        v[0] += u.z * zi;
      });
  block_partial<1>(v, part);
}

// x += M .* (b - t)
__global__ void k_post(double* __restrict__ x, const double* __restrict__ M, const double* __restrict__ b,
                       const double* __restrict__ t, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] += M[i] * (b[i] - t[i]);
}

// Entries per lane whose loads are in flight together: 8 when several lanes share a row (a
// level-1 row of ~44 entries at 8 lanes takes one round of index/value loads and one of gathers
// instead of two each), 4 at one row per lane (the prolongation rows hold ~3 entries).
template <int LANES>
constexpr int spmv_unroll() {
  return LANES >= 2 ? 8 : 4;
}

// Products of up to UNR strided entries of one row: the index and value loads are issued
// together, then the gathers of x together, so a lane waits for two round trips per UNR
// entries. No branch per entry: a padding entry re-reads the row's last entry (j1 - 1 >= jb is a
// valid entry inside the loop) and contributes an exact zero through the select. MODE as in
// k_spmv.
template <int MODE, int LANES, typename TA, typename TV, typename TX, typename TX2>
__device__ __forceinline__ TA row_products(int j0, int j1, int lane, const int* __restrict__ Ai,
                                           const TV* __restrict__ Ax, const TX* __restrict__ x,
                                           const TX2* __restrict__ x2) {
  constexpr int UNR = spmv_unroll<LANES>();
  TA s = 0;
  for (int jb = j0 + lane; jb < j1; jb += LANES * UNR) {
    int cu[UNR];
    TA au[UNR];
    bool ok[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int j = jb + u * LANES;
      ok[u] = j < j1;
      const int jj = ok[u] ? j : j1 - 1;
      cu[u] = Ai[jj];
      au[u] = (TA)Ax[jj];
    }
    TA xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int c = cu[u];
      xv[u] = (MODE == 2) ? (TA)((TA)x[c] - (TA)x2[c]) : (TA)x[c];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) s += ok[u] ? au[u] * xv[u] : (TA)0;
  }
  return s;
}

// Accumulation type of a matrix value type: FP16 values are multiplied and summed in FP32.
template <typename TV>
using acc_t = typename std::conditional<std::is_same<TV, __half>::value, float, TV>::type;

// CSR SpMV with LANES lanes per row; matrix values stored as TV, products and sums in TA.
// MODE 0: y = A x; 1: y += A x (and z = 0 on the same rows when z != nullptr); 2: y = A (x - x2)
// and, when z != nullptr, z = zw .* y on the same rows (a restriction that also forms the next
// level's pre-smoothed guess M b). The row's own entries of the epilogue are loaded first, so
// their latency overlaps the products.
template <int MODE, int LANES, typename TA, typename TX, typename TX2, typename TY, typename TV = TA>
__global__ __launch_bounds__(kBlock) void k_spmv(int row0, int nrows, const int* __restrict__ Ap,
                                                 const int* __restrict__ Ai, const TV* __restrict__ Ax,
                                                 const TX* __restrict__ x, const TX2* __restrict__ x2,
                                                 TY* __restrict__ y, TY* __restrict__ z, const TY* __restrict__ zw,
                                                 TA vs) {
  const int lane = threadIdx.x & (LANES - 1);
  const int row = row0 + (int)(((long)blockIdx.x * kBlock + threadIdx.x) / LANES);
  const bool live = row < nrows;
  TY y_old = 0, w = 0;
  if (live && lane == 0) {
    if (MODE == 1) y_old = y[row];
    if (MODE == 2 && z) w = zw[row];
  }
  TA s = 0;
  if (live) s = row_products<MODE, LANES, TA>(Ap[row], Ap[row + 1], lane, Ai, Ax, x, x2);
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LANES);
  if constexpr (std::is_same<TV, __half>::value) s *= vs;  // undo the FP16 copy's scale
  if (live && lane == 0) {
    if (MODE == 1) {
      y[row] = y_old + (TY)s;
      if (z) z[row] = 0;
    } else {
      const TY yv = (TY)s;
      y[row] = yv;
      if (MODE == 2 && z) z[row] = w * yv;
    }
  }
}

// Post-smoothing step as one SpMV: xs = x + M .* (b - A x), written to a separate vector.
template <int LANES, typename T, typename TV = T>
__global__ __launch_bounds__(kBlock) void k_spmv_post(int row0, int nrows, const int* __restrict__ Ap,
                                                      const int* __restrict__ Ai, const TV* __restrict__ Ax,
                                                      const T* __restrict__ x, const T* __restrict__ M,
                                                      const T* __restrict__ b, T* __restrict__ xs, T vs) {
  const int lane = threadIdx.x & (LANES - 1);
  const int row = row0 + (int)(((long)blockIdx.x * kBlock + threadIdx.x) / LANES);
  const bool live = row < nrows;
  T xr = 0, mr = 0, br = 0;
  if (live && lane == 0) {
    xr = x[row];
    mr = M[row];
    br = b[row];
  }
  T s = 0;
  if (live) s = row_products<0, LANES, T>(Ap[row], Ap[row + 1], lane, Ai, Ax, x, (const T*)nullptr);
#pragma unroll
  for (int o = LANES / 2; o > 0; o >>= 1) s += __shfl_down(s, o, LANES);
  if constexpr (std::is_same<TV, __half>::value) s *= vs;  // undo the FP16 copy's scale
  if (live && lane == 0) xs[row] = xr + mr * (br - s);
}

// y = A b for the symmetric coarsest-level inverse stored as its lower-triangle 64 x 64 tiles
// (each tile read once, half the bytes of the full matrix). k_symv_tiles: one workgroup per
// tile (I, J); it writes the row products A_IJ b_J to part[I][J] and, off the diagonal, the
// column products A_IJ^T b_I to part[J][I]. k_symv_reduce sums part[I][0..T-1] in order, so
// the result is deterministic.
template <typename T>
__global__ __launch_bounds__(256) void k_symv_tiles(int n, int T_, const T* __restrict__ tiles,
                                                    const T* __restrict__ b, double* __restrict__ part) {
  __shared__ double sbI[64], sbJ[64];
  __shared__ double sQ[4][64];
  const int t = blockIdx.x;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const int J = t - I * (I + 1) / 2;
  const int tid = threadIdx.x;
  if (tid < 64) {
    const int i = I * 64 + tid, j = J * 64 + tid;
    sbI[tid] = i < n ? (double)b[i] : 0.0;
    sbJ[tid] = j < n ? (double)b[j] : 0.0;
  }
  __syncthreads();
  // thread (R, Q) holds the 4 x 4 sub-block rows 4R..4R+3, columns 4Q..4Q+3 (R = tid / 16,
  // Q = tid % 16): row sums reduce over Q (lane bits 0..3), column sums over R (lane bits 4, 5,
  // then the 4 waves in LDS)
  const int R = tid >> 4, Q = tid & 15;
  const T* a = tiles + (size_t)t * 4096 + (4 * R) * 64 + 4 * Q;
  double av[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 4; ++k) av[i][k] = (double)a[i * 64 + k];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double u = av[i][0] * sbJ[4 * Q] + av[i][1] * sbJ[4 * Q + 1] + av[i][2] * sbJ[4 * Q + 2] + av[i][3] * sbJ[4 * Q + 3];
    u += __shfl_xor(u, 1, 64);
    u += __shfl_xor(u, 2, 64);
    u += __shfl_xor(u, 4, 64);
    u += __shfl_xor(u, 8, 64);
    if (Q == 0) part[((size_t)I * T_ + J) * 64 + 4 * R + i] = u;
  }
  if (I == J) return;  // uniform per workgroup
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double v = av[0][k] * sbI[4 * R] + av[1][k] * sbI[4 * R + 1] + av[2][k] * sbI[4 * R + 2] + av[3][k] * sbI[4 * R + 3];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (lane < 16) sQ[wv][4 * Q + k] = v;
  }
  __syncthreads();
  if (tid < 64) part[((size_t)J * T_ + I) * 64 + tid] = (sQ[0][tid] + sQ[1][tid]) + (sQ[2][tid] + sQ[3][tid]);
}

// y[64 I + r] = sum over J of part[I][J][r]: one workgroup per tile row I; thread (r, g) sums the
// J = g mod 4 terms, the four group sums are combined in a fixed order
template <typename T>
__global__ __launch_bounds__(256) void k_symv_reduce(int n, int T_, const double* __restrict__ part,
                                                     T* __restrict__ y) {
  __shared__ double sh[4][64];
  const int I = blockIdx.x, r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const double* p = part + (size_t)I * T_ * 64 + r;
  double s = 0.0;
  // J = g, g + 4, ... in order, 16 loads in flight per batch
  for (int J0 = g; J0 < T_; J0 += 64) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int J = J0 + 4 * u;
      v[u] = J < T_ ? p[(size_t)J * 64] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  sh[g][r] = s;
  __syncthreads();
  const int i = I * 64 + threadIdx.x;
  if (threadIdx.x < 64 && i < n) y[i] = (T)((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]));
}

// y = dense_inv b on level L (FP64 or FP32 tiles)
template <typename T>
static void symv(const Level& L, const T* tiles, const T* b, T* y, hipStream_t s) {
  const int T_ = L.tiles;
  k_symv_tiles<T><<<T_ * (T_ + 1) / 2, 256, 0, s>>>(L.n, T_, tiles, b, L.symv_part);
  k_symv_reduce<T><<<T_, 256, 0, s>>>(L.n, T_, L.symv_part, y);
}

template <int MODE, typename TV, typename TX, typename TX2, typename TY>
static void spmv_lanes(int lanes, int row0, int row1, const int* Ap, const int* Ai, const TV* Ax, const TX* x,
                       const TX2* x2, TY* y, TY* z, const TY* zw, hipStream_t s, acc_t<TV> vs) {
  using TA = acc_t<TV>;
  const int grid = (int)(((long)(row1 - row0) * lanes + kBlock - 1) / kBlock);
  switch (lanes) {
    case 1: k_spmv<MODE, 1, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
    case 2: k_spmv<MODE, 2, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
    case 4: k_spmv<MODE, 4, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
    case 8: k_spmv<MODE, 8, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
    case 16: k_spmv<MODE, 16, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
    case 32: k_spmv<MODE, 32, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
    default: k_spmv<MODE, 64, TA, TX, TX2, TY, TV><<<grid, kBlock, 0, s>>>(row0, row1, Ap, Ai, Ax, x, x2, y, z, zw, vs); break;
  }
}

// Row ranges of one launch: every row of the level (runs null or empty), or the rows of the
// segments a segmented solve still iterates (set_block_lists)
static inline void for_runs(const Runs* runs, int nrows, const std::function<void(int, int)>& f) {
  if (runs == nullptr || runs->empty()) {
    if (nrows > 0) f(0, nrows);
    return;
  }
  for (const auto& r : *runs)
    if (r.second > r.first) f(r.first, r.second);
}

// y = A x (mode 0), y += A x and z = 0 (mode 1), y = A (x - x2) and z = zw .* y (mode 2; z optional)
template <typename TV, typename TX, typename TX2, typename TY>
static void spmv(int mode, int lanes, int nrows, const int* Ap, const int* Ai, const TV* Ax, const TX* x,
                 const TX2* x2, TY* y, hipStream_t s, TY* z = nullptr, const Runs* runs = nullptr,
                 const TY* zw = nullptr, acc_t<TV> vs = 1) {
  for_runs(runs, nrows, [&](int r0, int r1) {
    if (mode == 0)
      spmv_lanes<0>(lanes, r0, r1, Ap, Ai, Ax, x, x2, y, (TY*)nullptr, (const TY*)nullptr, s, vs);
    else if (mode == 1)
      spmv_lanes<1>(lanes, r0, r1, Ap, Ai, Ax, x, x2, y, z, (const TY*)nullptr, s, vs);
    else
      spmv_lanes<2>(lanes, r0, r1, Ap, Ai, Ax, x, x2, y, zw ? z : (TY*)nullptr, zw, s, vs);
  });
}

template <typename T, typename TV>
static void spmv_post(int lanes, int nrows, const int* Ap, const int* Ai, const TV* Ax, const T* x, const T* M,
                      const T* b, T* xs, hipStream_t s, const Runs* runs = nullptr, T vs = 1) {
  for_runs(runs, nrows, [&](int r0, int r1) {
    const int grid = (int)(((long)(r1 - r0) * lanes + kBlock - 1) / kBlock);
    switch (lanes) {
      case 1: k_spmv_post<1, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
      case 2: k_spmv_post<2, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
      case 4: k_spmv_post<4, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
      case 8: k_spmv_post<8, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
      case 16: k_spmv_post<16, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
      case 32: k_spmv_post<32, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
      default: k_spmv_post<64, T, TV><<<grid, kBlock, 0, s>>>(r0, r1, Ap, Ai, Ax, x, M, b, xs, vs); break;
    }
  });
}

// ------------------------------------------------------------ FP32 V-cycle helpers
template <typename TB>
__global__ void k_scale_f(const float* __restrict__ M, const TB* __restrict__ b, float* __restrict__ x,
                          float* __restrict__ t, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[i] = M[i] * (float)b[i];
    if (t) t[i] = 0.0f;
  }
}

__global__ void k_post_f(float* __restrict__ x, const float* __restrict__ M, const float* __restrict__ b,
                         const float* __restrict__ t, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] += M[i] * (b[i] - t[i]);
}

// z = x + M (r - t) in FP64 from FP32 x, M, t and FP64 r; partials of r.z (FP64)
__global__ __launch_bounds__(kBlock) void k_post_dot_mixed(const float* __restrict__ x, const float* __restrict__ M,
                                                           const double* __restrict__ r, const float* __restrict__ t,
                                                           double* __restrict__ z, long n, double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return double4{(double)x[i], (double)M[i], r[i], (double)t[i]}; },
      [&](long i, double4 u) {
        const double zi = u.x + u.y * (u.z - u.w);
        z[i] = zi;
        v[0] += u.z * zi;
      });
  block_partial<1>(v, part);
}

// --------------------------------------------------------------------- apply timing
int resolve_apply_timing(Ctx& c) {
  if (c.ev_used == 0) return 0;
  MM_HIP(hipEventSynchronize(c.ev_pool[2 * c.ev_used - 1]));
  for (int k = 0; k < c.ev_used; ++k) {
    float ms = 0.f;
    MM_HIP(hipEventElapsedTime(&ms, c.ev_pool[2 * k], c.ev_pool[2 * k + 1]));
    c.apply_ms += ms;
  }
  c.apply_timed += c.ev_used;
  c.ev_used = 0;
  return 0;
}

// Every kApplySample-th apply inside the PCG is bracketed by an event pair: enough samples for
// the average launch time, without an event pair on every launch of the timed sweep.
constexpr long long kApplySample = 8;

int timed_apply(Ctx& c, const double* x, double* y, hipStream_t s, double* dot_part) {
  if (!c.time_apply || (c.apply_seen++ % kApplySample) != 0) return launch_apply(c, c.geom_mode, x, y, s, dot_part);
  if (c.ev_pool.empty()) {
    c.ev_pool.resize(512);
    for (auto& e : c.ev_pool) MM_HIP(hipEventCreate(&e));
  }
  if (2 * c.ev_used >= (int)c.ev_pool.size()) MM_TRY(resolve_apply_timing(c));
  MM_HIP(hipEventRecord(c.ev_pool[2 * c.ev_used], s));
  MM_TRY(launch_apply(c, c.geom_mode, x, y, s, dot_part));
  MM_HIP(hipEventRecord(c.ev_pool[2 * c.ev_used + 1], s));
  ++c.ev_used;
  return 0;
}

static int timed_apply_scaled(Ctx& c, const double* r, const double* w, double* xo, double* y, hipStream_t s) {
  if (!c.time_apply || (c.apply_seen++ % kApplySample) != 0) return launch_apply_scaled(c, r, w, xo, y, s);
  if (c.ev_pool.empty()) {
    c.ev_pool.resize(512);
    for (auto& e : c.ev_pool) MM_HIP(hipEventCreate(&e));
  }
  if (2 * c.ev_used >= (int)c.ev_pool.size()) MM_TRY(resolve_apply_timing(c));
  MM_HIP(hipEventRecord(c.ev_pool[2 * c.ev_used], s));
  MM_TRY(launch_apply_scaled(c, r, w, xo, y, s));
  MM_HIP(hipEventRecord(c.ev_pool[2 * c.ev_used + 1], s));
  ++c.ev_used;
  return 0;
}

// ---------------------------------------------------------------- segmented PCG kernels
// Several independent systems (segments: contiguous node ranges, e.g. the grains of a grain set)
// share one context and are solved together: each kernel launch covers the segments that are
// still iterating, and each segment keeps its own alpha, beta and stopping test. Workgroup w
// works on chunk wg_chunk[w] (kSegChunk nodes) of segment wg_seg[w] only, so its partial sums
// belong to one segment; a consumer sums the partials of its own segment in a fixed order.
// The element kernels run over the blocks of the iterating segments (set_block_lists); as
// segments converge the launches shrink to the others.
constexpr int kSegChunk = 4 * kBlock;

struct SegView {
  const int* node_off;  // nseg + 1
  const int* blk_off;   // nseg + 1 block offsets of the apply's partition (x.Ax partials)
  const int* wg_off;    // nseg + 1 (workgroups of each segment in this launch)
  const int* wg_seg;    // per workgroup
  const int* wg_chunk;  // per workgroup
  const int* active;    // per segment: still iterating
  double* rz;           // 2 x nseg
};

__device__ __forceinline__ void seg_range(const SegView& S, int& g, long& lo, long& hi) {
  const int w = blockIdx.x;
  g = S.wg_seg[w];
  lo = S.node_off[g] + (long)S.wg_chunk[w] * kSegChunk;
  hi = min(lo + (long)kSegChunk, (long)S.node_off[g + 1]);
}

// r = b - q, q = 0, t = 0 (the FP64 or the FP32 level-0 work vector); partials of r.r and b.b
// (2 per workgroup) for the host
__global__ __launch_bounds__(kBlock) void k_seg_init(SegView S, const double* __restrict__ b, double* __restrict__ q,
                                                     double* __restrict__ r, double* __restrict__ t,
                                                     float* __restrict__ tf, double* __restrict__ part_rr) {
  int g;
  long lo, hi;
  seg_range(S, g, lo, hi);
  double v[2] = {0.0, 0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double ri = b[i] - q[i];
    q[i] = 0.0;
    if (t) t[i] = 0.0;
    if (tf) tf[i] = 0.0f;
    r[i] = ri;
    v[0] += ri * ri;
    v[1] += b[i] * b[i];
  }
  block_partial<2>(v, part_rr);
}

// z = x + M .* (b - t) (level-0 post-smoothing), partials of r.z with r == b
__global__ __launch_bounds__(kBlock) void k_seg_post_dot(SegView S, const double* __restrict__ x,
                                                         const double* __restrict__ M, const double* __restrict__ b,
                                                         const double* __restrict__ t, double* __restrict__ z,
                                                         double* __restrict__ part) {
  int g;
  long lo, hi;
  seg_range(S, g, lo, hi);
  double v[1] = {0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double zi = x[i] + M[i] * (b[i] - t[i]);
    z[i] = zi;
    v[0] += b[i] * zi;
  }
  block_partial<1>(v, part);
}

// mixed-precision form of k_seg_post_dot: z = x + M .* (b - t) from the FP32 level-0 vectors
__global__ __launch_bounds__(kBlock) void k_seg_post_dot_mixed(SegView S, const float* __restrict__ x,
                                                               const float* __restrict__ M,
                                                               const double* __restrict__ b,
                                                               const float* __restrict__ t, double* __restrict__ z,
                                                               double* __restrict__ part) {
  int g;
  long lo, hi;
  seg_range(S, g, lo, hi);
  double v[1] = {0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double zi = (double)x[i] + (double)M[i] * (b[i] - (double)t[i]);
    z[i] = zi;
    v[0] += b[i] * zi;
  }
  block_partial<1>(v, part);
}

// rz_g = the segment's r.z; beta_g = rz_g / rz[old]; p = z + beta p on active segments (first:
// p = z); q = 0 everywhere; the segment's first workgroup stores rz_g to rz[new]
__global__ __launch_bounds__(kBlock) void k_seg_p(SegView S, int first, int slot_new, int slot_old,
                                                  const double* __restrict__ part_rz, const double* __restrict__ z,
                                                  double* __restrict__ p, double* __restrict__ q) {
  int g;
  long lo, hi;
  seg_range(S, g, lo, hi);
  const double rz = sum_partials_all<kBlock>(part_rz + S.wg_off[g], S.wg_off[g + 1] - S.wg_off[g]);
  const double beta = first ? 0.0 : rz / (S.rz[2 * g + slot_old] + 1e-30);
  const bool act = S.active[g] != 0;
  if ((int)blockIdx.x == S.wg_off[g] && threadIdx.x == 0) S.rz[2 * g + slot_new] = rz;
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    if (act) p[i] = first ? z[i] : z[i] + beta * p[i];
    q[i] = 0.0;
  }
}

// alpha_g = rz[slot] / p.Ap (the segment's block partials); x += alpha p, r -= alpha q on active
// segments; t = 0; partials of r.r for the host
__global__ __launch_bounds__(kBlock) void k_seg_xr(SegView S, int slot, const double* __restrict__ dot_part,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   double* __restrict__ t, float* __restrict__ tf,
                                                   double* __restrict__ part_rr) {
  int g;
  long lo, hi;
  seg_range(S, g, lo, hi);
  const double pq = sum_partials_all<kBlock>(dot_part + S.blk_off[g], S.blk_off[g + 1] - S.blk_off[g]);
  const double alpha = S.rz[2 * g + slot] / (pq + 1e-30);
  const bool act = S.active[g] != 0;
  double v[1] = {0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    if (t) t[i] = 0.0;
    if (tf) tf[i] = 0.0f;
    double ri = r[i];
    if (act) {
      x[i] += alpha * p[i];
      ri -= alpha * q[i];
      r[i] = ri;
    }
    v[0] += ri * ri;
  }
  block_partial<1>(v, part_rr);
}

static SegView seg_view(const Ctx& c) {
  SegView S;
  S.node_off = c.d_seg_node_off;
  S.blk_off = c.apply_clusters ? c.d_seg_blk_cl : c.d_seg_blk_all;
  S.wg_off = c.wg_pcg.off;
  S.wg_seg = c.wg_pcg.seg;
  S.wg_chunk = c.wg_pcg.chunk;
  S.active = c.d_seg_active;
  S.rz = c.d_seg_rz;
  return S;
}

void launch_seg_post_dot(const Ctx& c, const double* x, const double* M, const double* b, const double* t, double* z,
                         hipStream_t s) {
  if (c.wg_pcg.n > 0)
    k_seg_post_dot<<<c.wg_pcg.n, kBlock, 0, s>>>(seg_view(c), x, M, b, t, z, c.seg_part_rz);
}

static void launch_seg_post_dot_mixed(const Ctx& c, const float* x, const float* M, const double* b, const float* t,
                                      double* z, hipStream_t s) {
  if (c.wg_pcg.n > 0)
    k_seg_post_dot_mixed<<<c.wg_pcg.n, kBlock, 0, s>>>(seg_view(c), x, M, b, t, z, c.seg_part_rz);
}

// device tables of the segmented decomposition (once per context)
int ensure_seg(Ctx& c) {
  if (c.seg_nwg > 0) return 0;
  std::vector<int> wg_off(c.nseg + 1, 0);
  for (int g = 0; g < c.nseg; ++g) {
    const int n = c.h_seg_node_off[g + 1] - c.h_seg_node_off[g];
    wg_off[g + 1] = wg_off[g] + (n + kSegChunk - 1) / kSegChunk;
  }
  c.seg_nwg = wg_off[c.nseg];
  c.h_seg_wg_off = wg_off;
  auto up = [&](int** d, const std::vector<int>& h) -> int {
    MM_HIP(hipMalloc((void**)d, h.size() * sizeof(int)));
    c.allocs.push_back(*d);
    // on the context's stream and completed here (the stream does not wait for null-stream copies)
    MM_HIP(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
    MM_HIP(hipStreamSynchronize(c.stream));
    return 0;
  };
  MM_TRY(up(&c.d_seg_node_off, c.h_seg_node_off));
  MM_TRY(up(&c.d_seg_blk_all, c.h_seg_blk_all));
  MM_TRY(up(&c.d_seg_blk_cl, c.h_seg_blk_cl));
  // the subset tables, one buffer: active mask, the two workgroup tables, the three block lists
  // (each list holds at most every block of its partition)
  {
    const int G = c.nseg, W = c.seg_nwg;
    const int sizes[10] = {G, W, W, G + 1, W, W, G + 1, std::max(c.part_cl.nblk, 1), std::max(c.part_all.nblk, 1),
                           std::max(c.part_mag.nblk, 1)};
    int off[11] = {0};
    for (int q = 0; q < 10; ++q) off[q + 1] = off[q] + sizes[q];
    c.seg_tab_n = off[10];
    c.seg_master.assign(c.seg_tab_n, 0);
    std::fill(c.seg_master.begin(), c.seg_master.begin() + G, 1);  // every segment active
    MM_TRY(up(&c.seg_tab, c.seg_master));
    int* t = c.seg_tab;
    c.d_seg_active = t + off[0];
    c.wg_pcg.seg = t + off[1];
    c.wg_pcg.chunk = t + off[2];
    c.wg_pcg.off = t + off[3];
    c.wg_min.seg = t + off[4];
    c.wg_min.chunk = t + off[5];
    c.wg_min.off = t + off[6];
    c.d_blist_cl = t + off[7];
    c.d_blist_all = t + off[8];
    c.d_blist_mag = t + off[9];
    c.wg_pcg.key.clear();
    c.wg_min.key.clear();
    MM_HIP(hipHostMalloc((void**)&c.seg_stage, (size_t)kSegStageSlots * c.seg_tab_n * sizeof(int),
                         hipHostMallocDefault));
    for (int k = 0; k < kSegStageSlots; ++k) {
      MM_HIP(hipEventCreateWithFlags(&c.seg_stage_ev[k], hipEventDisableTiming));
      c.seg_stage_rec[k] = false;
    }
    c.seg_dirty_lo = c.seg_dirty_hi = -1;
  }
  c.blist_key.clear();
  MM_HIP(hipMalloc((void**)&c.d_seg_rz, 2 * (size_t)c.nseg * sizeof(double)));
  c.allocs.push_back(c.d_seg_rz);
  MM_HIP(hipMalloc((void**)&c.seg_part_rz, (size_t)c.seg_nwg * sizeof(double)));
  c.allocs.push_back(c.seg_part_rz);
  MM_HIP(hipHostMalloc((void**)&c.hseg, 2 * (size_t)c.seg_nwg * sizeof(double),
                       hipHostMallocMapped | hipHostMallocCoherent));
  MM_HIP(hipHostGetDevicePointer((void**)&c.hseg_dev, c.hseg, 0));
  return 0;
}

int seg_flush(Ctx& c) {
  if (c.seg_dirty_lo < 0) return 0;
  const int k = c.seg_stage_next;
  c.seg_stage_next = (k + 1) % kSegStageSlots;
  if (c.seg_stage_rec[k]) MM_HIP(hipEventSynchronize(c.seg_stage_ev[k]));  // the slot's last copy is done
  int* slot = c.seg_stage + (size_t)k * c.seg_tab_n;
  const long lo = c.seg_dirty_lo, n = c.seg_dirty_hi - c.seg_dirty_lo;
  std::copy(c.seg_master.begin() + lo, c.seg_master.begin() + lo + n, slot + lo);
  MM_HIP(hipMemcpyAsync(c.seg_tab + lo, slot + lo, n * sizeof(int), hipMemcpyHostToDevice, c.stream));
  MM_HIP(hipEventRecord(c.seg_stage_ev[k], c.stream));
  c.seg_stage_rec[k] = true;
  c.seg_dirty_lo = c.seg_dirty_hi = -1;
  return 0;
}

int seg_upload(Ctx& c, int* dev, const int* src, int n) {
  if (n <= 0) return 0;
  const long lo = dev - c.seg_tab;
  MM_CHECK(c.seg_tab && lo >= 0 && lo + n <= c.seg_tab_n, "segment table upload out of range");
  std::copy(src, src + n, c.seg_master.begin() + lo);
  c.seg_dirty_lo = c.seg_dirty_lo < 0 ? lo : std::min(c.seg_dirty_lo, lo);
  c.seg_dirty_hi = std::max(c.seg_dirty_hi, lo + n);
  return c.seg_defer ? 0 : seg_flush(c);
}

static std::vector<int> full_mask(const Ctx& c, const int* mask) {
  std::vector<int> k(c.nseg, 1);
  if (mask)
    for (int g = 0; g < c.nseg; ++g) k[g] = mask[g] != 0;
  return k;
}

int build_wg_tables(Ctx& c, const int* mask, WgTables& T) {
  MM_TRY(ensure_seg(c));
  std::vector<int> key = full_mask(c, mask);
  if (key == T.key) return 0;
  std::vector<int> seg, chunk, off(c.nseg + 1, 0);
  seg.reserve(c.seg_nwg);
  chunk.reserve(c.seg_nwg);
  for (int g = 0; g < c.nseg; ++g) {
    const int k = key[g] ? c.h_seg_wg_off[g + 1] - c.h_seg_wg_off[g] : 0;
    for (int j = 0; j < k; ++j) {
      seg.push_back(g);
      chunk.push_back(j);
    }
    off[g + 1] = off[g] + k;
  }
  T.n = off[c.nseg];
  // stream-ordered after the kernels that read the previous tables
  {
    SegBatch batch(c);
    MM_TRY(seg_upload(c, T.seg, seg.data(), T.n));
    MM_TRY(seg_upload(c, T.chunk, chunk.data(), T.n));
    MM_TRY(seg_upload(c, T.off, off.data(), (int)off.size()));
  }
  if (!c.seg_defer) MM_TRY(seg_flush(c));
  T.h_off = std::move(off);
  T.key = std::move(key);
  return 0;
}

int set_block_lists(Ctx& c, const int* mask) {
  MM_TRY(ensure_seg(c));
  std::vector<int> key = full_mask(c, mask);
  if (key == c.blist_key) return 0;
  bool all = true;
  for (int v : key) all &= v != 0;
  auto fill = [&](Part& p, const std::vector<int>& off, std::vector<int>& h, int* d) -> int {
    if (all) {
      p.blist = nullptr;
      p.nlist = 0;
      return 0;
    }
    h.clear();
    for (int g = 0; g < c.nseg; ++g)
      if (key[g])
        for (int b = off[g]; b < off[g + 1]; ++b) h.push_back(b);
    MM_TRY(seg_upload(c, d, h.data(), (int)h.size()));
    p.blist = d;
    p.nlist = (int)h.size();
    return 0;
  };
  {
    SegBatch batch(c);
    MM_TRY(fill(c.part_cl, c.h_seg_blk_cl, c.h_blist_cl, c.d_blist_cl));
    MM_TRY(fill(c.part_all, c.h_seg_blk_all, c.h_blist_all, c.d_blist_all));
    MM_TRY(fill(c.part_mag, c.h_seg_blk_mag, c.h_blist_mag, c.d_blist_mag));
  }
  if (!c.seg_defer) MM_TRY(seg_flush(c));
  // AMG rows of the same segments: runs of consecutive segments (at most kMaxRuns launches per
  // SpMV, else the span from the first to the last); every row where a level's unknowns are not
  // ordered by segment
  constexpr int kMaxRuns = 8;
  for (Level& L : c.lv) {
    L.runs.clear();
    if (all || (int)L.seg_off.size() != c.nseg + 1) continue;
    for (int g = 0; g < c.nseg;) {
      if (!key[g]) {
        ++g;
        continue;
      }
      int h = g;
      while (h + 1 < c.nseg && key[h + 1]) ++h;
      L.runs.emplace_back(L.seg_off[g], L.seg_off[h + 1]);
      g = h + 1;
    }
    if ((int)L.runs.size() > kMaxRuns) L.runs = {{L.runs.front().first, L.runs.back().second}};
    if (L.runs.empty()) L.runs.emplace_back(0, 0);  // no segment: launch nothing
  }
  c.blist_key = std::move(key);
  return 0;
}

// PCG for every segment at once (make_pcg_solve per system, src/poisson_solve.py:422-478):
// stop segment g when r_g.r_g <= tol_g^2 b_g.b_g or after maxiter iterations. Only segments
// with want[g] != 0 (all when want is null) take part; the others keep U and report r2 = b.b = 0.
static int pcg_seg(Ctx& c, const double* b, double* U, const double* tol, const int* want, int maxiter, int* it_out,
                   double* r2_out, double* bb_out) {
  hipStream_t s = c.stream;
  const long N = c.N;
  const int G = c.nseg;
  c.h_seg_active = full_mask(c, want);
  {
    SegBatch batch(c);  // mask, workgroup tables and block lists in one copy
    MM_TRY(seg_upload(c, c.d_seg_active, c.h_seg_active.data(), G));
    MM_TRY(build_wg_tables(c, c.h_seg_active.data(), c.wg_pcg));
    MM_TRY(set_block_lists(c, c.h_seg_active.data()));
  }
  MM_TRY(seg_flush(c));
  std::vector<double> r2(G, 0.0), bb(G, 0.0), tol2(G, 0.0);
  for (int g = 0; g < G; ++g) it_out[g] = 0;
  int W = c.wg_pcg.n;
  if (W == 0) {
    for (int g = 0; g < G; ++g) r2_out[g] = bb_out[g] = 0.0;
    return set_block_lists(c, nullptr);
  }
  const SegView S = seg_view(c);
  vfill(c.pq, 0.0, N, s);
  MM_TRY(timed_apply(c, U, c.pq, s));
  c.n_apply++;
  float* tf = c.precision >= 1 ? c.ptf : nullptr;
  double* t64 = c.precision >= 1 ? nullptr : c.pt;
  k_seg_init<<<W, kBlock, 0, s>>>(S, b, c.pq, c.pr, t64, tf, c.hseg_dev);
  c.seg_running = true;
  int rc = precond(c, c.pr, c.pz, c.part_rz, s);
  c.seg_running = false;
  MM_TRY(rc);
  k_seg_p<<<W, kBlock, 0, s>>>(S, 1, 0, 0, c.seg_part_rz, c.pz, c.pp, c.pq);
  MM_HIP(hipStreamSynchronize(s));
  {
    const std::vector<int>& wo = c.wg_pcg.h_off;
    for (int g = 0; g < G; ++g) {
      for (int w = wo[g]; w < wo[g + 1]; ++w) {
        r2[g] += c.hseg[2 * w];
        bb[g] += c.hseg[2 * w + 1];
      }
      tol2[g] = tol[g] * tol[g] * bb[g];
    }
  }
  int live = 0;
  for (int g = 0; g < G; ++g) {
    c.h_seg_active[g] = c.h_seg_active[g] && r2[g] > tol2[g] && maxiter > 0;
    live += c.h_seg_active[g];
  }
  auto shrink = [&]() -> int {
    {
      SegBatch batch(c);  // one copy for the mask, the workgroup tables and the block lists
      MM_TRY(seg_upload(c, c.d_seg_active, c.h_seg_active.data(), G));
      MM_TRY(build_wg_tables(c, c.h_seg_active.data(), c.wg_pcg));
      MM_TRY(set_block_lists(c, c.h_seg_active.data()));
    }
    MM_TRY(seg_flush(c));
    W = c.wg_pcg.n;
    return 0;
  };
  MM_TRY(shrink());
  int it = 0, cur = 0;
  while (live > 0 && it < maxiter) {
    MM_TRY(timed_apply(c, c.pp, c.pq, s, c.part_pq));  // q was cleared by k_seg_p
    c.n_apply++;
    k_seg_xr<<<W, kBlock, 0, s>>>(S, cur, c.part_pq, U, c.pr, c.pp, c.pq, t64, tf, c.hseg_dev);
    MM_HIP(hipEventRecord(c.ev, s));
    c.seg_running = true;
    rc = precond(c, c.pr, c.pz, c.part_rz, s);
    c.seg_running = false;
    MM_TRY(rc);
    k_seg_p<<<W, kBlock, 0, s>>>(S, 0, 1 - cur, cur, c.seg_part_rz, c.pz, c.pp, c.pq);
    cur = 1 - cur;
    ++it;
    MM_HIP(hipEventSynchronize(c.ev));
    bool changed = false;
    const std::vector<int>& wo = c.wg_pcg.h_off;
    for (int g = 0; g < G; ++g) {
      if (!c.h_seg_active[g]) continue;
      r2[g] = 0.0;
      for (int w = wo[g]; w < wo[g + 1]; ++w) r2[g] += c.hseg[w];
      it_out[g] = it;
      if (r2[g] <= tol2[g] || it >= maxiter) {
        c.h_seg_active[g] = 0;
        --live;
        changed = true;
      }
    }
    if (changed && live > 0) MM_TRY(shrink());
  }
  MM_HIP(hipStreamSynchronize(s));
  MM_TRY(resolve_apply_timing(c));
  MM_TRY(set_block_lists(c, nullptr));
  c.n_pcg_it += it;
  for (int g = 0; g < G; ++g) {
    r2_out[g] = r2[g];
    bb_out[g] = bb[g];
  }
  return 0;
}

int solve_U_seg(Ctx& c, const double* m, double* U, const double* tol, int maxiter, int* it, double* rel_res) {
  return solve_U_seg_dev(c, m, U, tol, nullptr, maxiter, it, rel_res);
}

int solve_U_seg_dev(Ctx& c, const double* m, double* U, const double* tol, const int* want, int maxiter, int* it,
                    double* rel_res) {
  MM_CHECK(c.lv.size() > 1 && c.lv[0].M != nullptr && c.geom_mode == kGeomCoords,
           "segmented solves need the AMG V-cycle and the on-the-fly apply");
  MM_TRY(ensure_seg(c));
  hipStream_t s = c.stream;
  vfill(c.pb, 0.0, c.N, s);
  MM_TRY(set_block_lists(c, want));
  MM_TRY(launch_rhs(c, m, c.pb, s));
  c.n_rhs++;
  std::vector<double> r2(c.nseg), bb(c.nseg);
  MM_TRY(pcg_seg(c, c.pb, U, tol, want, maxiter, it, r2.data(), bb.data()));
  for (int g = 0; g < c.nseg; ++g) rel_res[g] = std::sqrt(r2[g] / (bb[g] + 1e-30));
  return 0;
}

// ----------------------------------------------------------------------------- V-cycle
// Coarse part of the V-cycle on levels l >= 1; the correction is returned in lv[l].xs.
// Pre-smoothing from a zero guess, x = M b, comes with b from the restriction that formed b;
// the residual SpMV t = A x follows and the post-smoothing x + M (b - A x) is one SpMV into xs.
static void vcycle_coarse(Ctx& c, int l, hipStream_t s) {
  Level& L = c.lv[l];
  const int last = (int)c.lv.size() - 1;
  if (l == last) {
    symv<double>(L, L.dense_inv, L.b, L.xs, s);
    return;
  }
  Level& Nx = c.lv[l + 1];
  const bool nx_smooth = l + 1 < last;  // the next level forms x = M b with its b
  spmv(0, L.lanesA, L.n, L.Ap, L.Ai, L.Ax, L.x, (const double*)nullptr, L.t, s, (double*)nullptr, &L.runs);  // t = A x
  spmv(2, L.lanesR, Nx.n, L.Rp, L.Ri, L.Rx, L.b, L.t, Nx.b, s, nx_smooth ? Nx.x : (double*)nullptr, &Nx.runs,
       nx_smooth ? (const double*)Nx.M : (const double*)nullptr);  // b' = R (b - t), x' = M' b'
  vcycle_coarse(c, l + 1, s);
  spmv(1, L.lanesP, L.n, L.Pp, L.Pi, L.Px, Nx.xs, (const double*)nullptr, L.x, s, (double*)nullptr,
       &L.runs);                                                                      // x += P e
  spmv_post(L.lanesA, L.n, L.Ap, L.Ai, L.Ax, L.x, L.M, L.b, L.xs, s, &L.runs);      // xs = x + M (b - A x)
}

// Matrix values of a level as the FP32 V-cycle reads them: FP32 copies, or FP16 copies (the FP32
// PCG's V-cycle, when every level holds them).
template <typename TV>
struct LevelVals;
template <>
struct LevelVals<float> {
  static const float* A(const Level& L) { return L.Axf; }
  static const float* P(const Level& L) { return L.Pxf; }
  static const float* R(const Level& L) { return L.Rxf; }
  static float As(const Level&) { return 1.0f; }
  static float Ps(const Level&) { return 1.0f; }
  static float Rs(const Level&) { return 1.0f; }
};
template <>
struct LevelVals<__half> {
  static const __half* A(const Level& L) { return L.Axh; }
  static const __half* P(const Level& L) { return L.Pxh; }
  static const __half* R(const Level& L) { return L.Rxh; }
  static float As(const Level& L) { return L.Axh_s; }
  static float Ps(const Level& L) { return L.Pxh_s; }
  static float Rs(const Level& L) { return L.Rxh_s; }
};

// FP32 coarse part of the V-cycle (levels >= 1), same structure as vcycle_coarse; matrix values
// read as TV.
template <typename TV = float>
static void vcycle_coarse_f(Ctx& c, int l, hipStream_t s) {
  using V = LevelVals<TV>;
  Level& L = c.lv[l];
  const int last = (int)c.lv.size() - 1;
  if (l == last) {
    symv<float>(L, L.dense_invf_t, L.bf, L.xsf, s);
    return;
  }
  Level& Nx = c.lv[l + 1];
  const bool nx_smooth = l + 1 < last;
  spmv(0, L.lanesA, L.n, L.Ap, L.Ai, V::A(L), L.xf, (const float*)nullptr, L.tf, s, (float*)nullptr, &L.runs,
       (const float*)nullptr, V::As(L));
  spmv(2, L.lanesR, Nx.n, L.Rp, L.Ri, V::R(L), L.bf, L.tf, Nx.bf, s, nx_smooth ? Nx.xf : (float*)nullptr, &Nx.runs,
       nx_smooth ? (const float*)Nx.Mf : (const float*)nullptr, V::Rs(L));
  vcycle_coarse_f<TV>(c, l + 1, s);
  spmv(1, L.lanesP, L.n, L.Pp, L.Pi, V::P(L), Nx.xsf, (const float*)nullptr, L.xf, s, (float*)nullptr, &L.runs,
       (const float*)nullptr, V::Ps(L));
  spmv_post(L.lanesA, L.n, L.Ap, L.Ai, V::A(L), L.xf, L.Mf, L.bf, L.xsf, s, &L.runs, V::As(L));
}

// Mixed-precision preconditioner: the whole V-cycle in FP32 (FP32 matrix-free fine applies,
// FP32 coarse matrices and dense inverse); input r and output z in FP64, r.z in FP64.
// The caller has cleared ptf.
static int precond_f32(Ctx& c, const double* r, double* z, double* rz_part, hipStream_t s) {
  const long N = c.N;
  Level& L0 = c.lv[0];
  Level& L1 = c.lv[1];
  MM_TRY(launch_apply_f32_scaled(c, r, L0.Mf, c.px0f, c.ptf, s));                 // x = M0 r, t = A x
  const bool l1_smooth = c.lv.size() > 2;  // level 1 is not the coarsest: x1 = M1 b1 with b1
  spmv(2, L0.lanesR, L1.n, L0.Rp, L0.Ri, L0.Rxf, r, c.ptf, L1.bf, s, l1_smooth ? L1.xf : (float*)nullptr, &L1.runs,
       l1_smooth ? (const float*)L1.Mf : (const float*)nullptr);  // b1 = R0 (r - t)
  vcycle_coarse_f(c, 1, s);
  spmv(1, L0.lanesP, (int)N, L0.Pp, L0.Pi, L0.Pxf, L1.xsf, (const float*)nullptr, c.px0f, s, c.ptf, &L0.runs);
  MM_TRY(launch_apply_f32(c, c.px0f, c.ptf, s));                                  // t = A x
  if (c.seg_running)
    launch_seg_post_dot_mixed(c, c.px0f, L0.Mf, r, c.ptf, z, s);
  else
    k_post_dot_mixed<<<red_grid(N), kBlock, 0, s>>>(c.px0f, L0.Mf, r, c.ptf, z, N, rz_part);
  c.n_vcycle += 1;
  return 0;
}

// z = V-cycle(r) with the matrix-free level 0; also writes red_grid(N) partials of r.z.
int precond(Ctx& c, const double* r, double* z, double* rz_part, hipStream_t s) {
  const long N = c.N;
  Level& L0 = c.lv[0];
  if (c.precision >= 1 && c.lv.size() > 1 && L0.Mf != nullptr) return precond_f32(c, r, z, rz_part, s);
  if (c.lv.size() <= 1) {
    if (L0.dense_inv)  // single-level hierarchy: exact dense solve
      symv<double>(L0, L0.dense_inv, r, z, s);
    else  // Jacobi on diag(A)
      k_scale<<<vgrid(N), kBlock, 0, s>>>(L0.M, r, z, nullptr, N);
    k_dot_part<<<red_grid(N), kBlock, 0, s>>>(r, z, N, rz_part);
    return 0;
  }
  Level& L1 = c.lv[1];
  // the caller has cleared pt
  MM_TRY(timed_apply_scaled(c, r, L0.M, c.px0, c.pt, s));                          // x = M0 r, t = A x
  const bool l1_smooth = c.lv.size() > 2;  // level 1 is not the coarsest: x1 = M1 b1 with b1
  spmv(2, L0.lanesR, L1.n, L0.Rp, L0.Ri, L0.Rx, r, c.pt, L1.b, s, l1_smooth ? L1.x : (double*)nullptr, &L1.runs,
       l1_smooth ? (const double*)L1.M : (const double*)nullptr);  // b1 = R0 (r - t), x1 = M1 b1
  vcycle_coarse(c, 1, s);
  spmv(1, L0.lanesP, (int)N, L0.Pp, L0.Pi, L0.Px, L1.xs, (const double*)nullptr, c.px0, s, c.pt,
       &L0.runs);  // x += P0 e1, t = 0
  MM_TRY(timed_apply(c, c.px0, c.pt, s));                                         // t = A x
  if (c.seg_running)
    launch_seg_post_dot(c, c.px0, L0.M, r, c.pt, z, s);
  else
    k_post_dot<<<red_grid(N), kBlock, 0, s>>>(c.px0, L0.M, r, c.pt, z, N, rz_part);
  c.n_apply += 2;
  c.n_vcycle += 1;
  return 0;
}

int precond_apply(Ctx& c, const double* r, double* z, double* rz_part, hipStream_t s) {
  MM_CHECK(c.lv.size() >= 1 && c.lv[0].M != nullptr, "preconditioner not set up");
  vfill(c.pt, 0.0, c.N, s);
  if (c.ptf) MM_HIP(hipMemsetAsync(c.ptf, 0, c.N * sizeof(float), s));
  return precond(c, r, z, rz_part, s);
}

// ------------------------------------------------- FP32 PCG with FP64 iterative refinement
// precision 2: the PCG runs on FP32 vectors (correction, residual, direction, A p, V-cycle
// output) with the FP32 matrix-free apply and the FP32 V-cycle, while its dot products and the
// alpha / beta recurrences stay FP64. Whenever the FP32 recursion has reduced the residual by
// the factor ref_eta since the last true residual (or reaches the tolerance), the correction is
// added to U in FP64 and the true residual r = b - A U is recomputed with the FP64 apply and
// replaces the recursive one; the search direction and the CG scalars carry on (residual
// replacement), so the Krylov iteration is not restarted. The stopping test uses the true
// residual only: r.r <= tol^2 b.b, as in the FP64 PCG, so the solution's accuracy is its.

// rf = (float) r, xc = 0, tf = 0
__global__ void k_ref_start(const double* __restrict__ r, float* __restrict__ rf, float* __restrict__ xc,
                            float* __restrict__ tf, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    rf[i] = (float)r[i];
    xc[i] = 0.0f;
    tf[i] = 0.0f;
  }
}

// z = x + M (r - t) in FP32 from the V-cycle's level-0 vectors, partials of r.z in FP64
__global__ __launch_bounds__(kBlock) void k_post_dot_ff(const float* __restrict__ x, const float* __restrict__ M,
                                                        const float* __restrict__ r, const float* __restrict__ t,
                                                        float* __restrict__ z, long n, double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return float4{x[i], M[i], r[i], t[i]}; },
      [&](long i, float4 u) {
        const float zi = u.x + u.y * (u.z - u.w);
        z[i] = zi;
        v[0] += (double)u.z * (double)zi;
      });
  block_partial<1>(v, part);
}

// alpha = rz / (p.q + 1e-30) in FP64; xc += alpha p, r -= alpha q (FP32 vectors, FP64
// arithmetic); partials of r.r; tf cleared for the next V-cycle
__global__ __launch_bounds__(kBlock) void k_ref_xr(const double* __restrict__ sc, int rz_slot,
                                                   const double* __restrict__ part_pq, int n_pq,
                                                   float* __restrict__ xc, float* __restrict__ r,
                                                   const float* __restrict__ p, const float* __restrict__ q,
                                                   float* __restrict__ tf, long n, double* __restrict__ part_rr) {
  const double alpha = sc[rz_slot] / (sum_partials_all<kBlock>(part_pq, n_pq) + 1e-30);
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return float4{xc[i], p[i], r[i], q[i]}; },
      [&](long i, float4 a) {
        tf[i] = 0.0f;
        xc[i] = (float)((double)a.x + alpha * (double)a.y);
        const float rf = (float)((double)a.z - alpha * (double)a.w);
        r[i] = rf;
        v[0] += (double)rf * (double)rf;
      });
  block_partial<1>(v, part_rr);
}

// rz_new from partials (kept in sc[new_slot]); beta = rz_new / sc[old_slot]; p = z + beta p
// (first: p = z); q = 0
__global__ __launch_bounds__(kBlock) void k_ref_p(double* __restrict__ sc, const double* __restrict__ part_rz,
                                                  int n_rz, int new_slot, int old_slot, int first,
                                                  const float* __restrict__ z, float* __restrict__ p,
                                                  float* __restrict__ q, long n) {
  const double rz = sum_partials_all<kBlock>(part_rz, n_rz);
  const double beta = first ? 0.0 : rz / (sc[old_slot] + 1e-30);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[new_slot] = rz;
  if (first)  // outside the loops, as in k_pcg_p
    grid_chunks(
        n, [&](long i) { return z[i]; },
        [&](long i, float u) {
          p[i] = u;
          q[i] = 0.0f;
        });
  else
    grid_chunks(
        n, [&](long i) { return float2{z[i], p[i]}; },
        [&](long i, float2 u) {
          p[i] = (float)((double)u.x + beta * (double)u.y);
          q[i] = 0.0f;
        });
}

// U += xc; q = 0 for the FP64 apply of the true residual
__global__ void k_ref_update(double* __restrict__ U, const float* __restrict__ xc, double* __restrict__ q, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    U[i] += (double)xc[i];
    q[i] = 0.0;
  }
}

// The FP32 V-cycle on an FP32 residual: z (FP32) and FP64 partials of r.z. The caller has
// cleared ptf. The transfer and coarse-level matrices are read from their FP16 copies when every
// level holds them (products and sums in FP32): this V-cycle is the preconditioner of a PCG whose
// stopping test is the FP64 true residual.
template <typename TV>
static int precond_ff_v(Ctx& c, const float* r, float* z, double* rz_part, hipStream_t s) {
  using V = LevelVals<TV>;
  const long N = c.N;
  Level& L0 = c.lv[0];
  Level& L1 = c.lv[1];
  MM_TRY(launch_apply_f32_scaled_f(c, r, L0.Mf, c.px0f, c.ptf, s));               // x = M0 r, t = A x
  const bool l1_smooth = c.lv.size() > 2;
  spmv(2, L0.lanesR, L1.n, L0.Rp, L0.Ri, V::R(L0), r, c.ptf, L1.bf, s, l1_smooth ? L1.xf : (float*)nullptr,
       &L1.runs, l1_smooth ? (const float*)L1.Mf : (const float*)nullptr, V::Rs(L0));  // b1 = R0 (r - t)
  vcycle_coarse_f<TV>(c, 1, s);
  spmv(1, L0.lanesP, (int)N, L0.Pp, L0.Pi, V::P(L0), L1.xsf, (const float*)nullptr, c.px0f, s, c.ptf, &L0.runs,
       (const float*)nullptr, V::Ps(L0));
  MM_TRY(launch_apply_f32(c, c.px0f, c.ptf, s));                                  // t = A x
  k_post_dot_ff<<<red_grid(N), kBlock, 0, s>>>(c.px0f, L0.Mf, r, c.ptf, z, N, rz_part);
  c.n_vcycle += 1;
  return 0;
}

static int precond_ff(Ctx& c, const float* r, float* z, double* rz_part, hipStream_t s) {
  return c.lv_half ? precond_ff_v<__half>(c, r, z, rz_part, s) : precond_ff_v<float>(c, r, z, rz_part, s);
}

static int timed_apply_f32_dot(Ctx& c, const float* x, float* y, double* dot_part, hipStream_t s) {
  if (!c.time_apply || (c.apply_seen++ % kApplySample) != 0) return launch_apply_f32_dot(c, x, y, dot_part, s);
  if (c.ev_pool.empty()) {
    c.ev_pool.resize(512);
    for (auto& e : c.ev_pool) MM_HIP(hipEventCreate(&e));
  }
  if (2 * c.ev_used >= (int)c.ev_pool.size()) MM_TRY(resolve_apply_timing(c));
  MM_HIP(hipEventRecord(c.ev_pool[2 * c.ev_used], s));
  MM_TRY(launch_apply_f32_dot(c, x, y, dot_part, s));
  MM_HIP(hipEventRecord(c.ev_pool[2 * c.ev_used + 1], s));
  ++c.ev_used;
  return 0;
}

// Inner reduction per refinement cycle: about what an FP32 residual can carry
// The recursion hands over to the true-residual test once its r.r is below this share of tol^2 b.b
// (the replaced residuals carry little drift, so the true one is then almost always below tol^2 b.b
// too; when it is not, the iteration goes on)
constexpr double kRefFinal = 0.81;

static int pcg_ref(Ctx& c, const double* b, double* U, double tol, int maxiter, int* it_out, double* r2_out,
                   double* bb_out, bool q_ready) {
  hipStream_t s = c.stream;
  const long N = c.N;
  double* sc = c.dscal;
  const int gv = red_grid(N);
  const int n_pq = apply_dot_count(c);
  if (!q_ready) {
    vfill(c.pq, 0.0, N, s);
    MM_TRY(launch_apply(c, kGeomCoords, U, c.pq, s));
    c.n_apply++;
  }
  k_pcg_init<<<gv, kBlock, 0, s>>>(b, c.pq, c.pr, nullptr, nullptr, N, c.hpart_dev);  // true FP64 residual
  MM_HIP(hipStreamSynchronize(s));
  double r2 = 0.0, bb = 0.0;
  for (int i = 0; i < gv; ++i) {
    r2 += c.hpart[2 * i];
    bb += c.hpart[2 * i + 1];
  }
  const double tol2 = tol * tol * bb;
  int it = 0, cycles = 0;
  if (it < maxiter && r2 > tol2) {
    k_ref_start<<<vgrid(N), kBlock, 0, s>>>(c.pr, c.rrf, c.rxf, c.ptf, N);  // r32 = r, xc = 0
    MM_TRY(precond_ff(c, c.rrf, c.rzf, c.part_rz, s));
    k_ref_p<<<gv, kBlock, 0, s>>>(sc, c.part_rz, gv, 0, 0, 1, c.rzf, c.rpf, c.rqf, N);
  }
  int cur = 0;
  double target = std::max(kRefFinal * tol2, c.ref_eta * c.ref_eta * r2);
  while (it < maxiter && r2 > tol2) {
    MM_TRY(timed_apply_f32_dot(c, c.rpf, c.rqf, c.part_pq, s));  // q was cleared by k_ref_p
    c.n_apply++;
    k_ref_xr<<<gv, kBlock, 0, s>>>(sc, cur, c.part_pq, n_pq, c.rxf, c.rrf, c.rpf, c.rqf, c.ptf, N, c.hpart_dev);
    MM_HIP(hipEventRecord(c.ev, s));
    // the next direction is formed while the host waits for the recursive r.r
    MM_TRY(precond_ff(c, c.rrf, c.rzf, c.part_rz, s));
    k_ref_p<<<gv, kBlock, 0, s>>>(sc, c.part_rz, gv, 1 - cur, cur, 0, c.rzf, c.rpf, c.rqf, N);
    cur = 1 - cur;
    ++it;
    MM_HIP(hipEventSynchronize(c.ev));
    double r2in = 0.0;
    for (int i = 0; i < gv; ++i) r2in += c.hpart[i];
    if (r2in > target && it < maxiter) continue;
    // residual replacement: U += xc in FP64, the true residual b - A U, loaded into the FP32
    // residual (xc restarts at zero); the direction p just formed and the scalars are kept
    k_ref_update<<<vgrid(N), kBlock, 0, s>>>(U, c.rxf, c.pq, N);
    MM_TRY(launch_apply(c, kGeomCoords, U, c.pq, s));
    c.n_apply++;
    k_pcg_init<<<gv, kBlock, 0, s>>>(b, c.pq, c.pr, nullptr, nullptr, N, c.hpart_dev);
    k_ref_start<<<vgrid(N), kBlock, 0, s>>>(c.pr, c.rrf, c.rxf, c.ptf, N);
    MM_HIP(hipStreamSynchronize(s));
    r2 = 0.0;
    for (int i = 0; i < gv; ++i) r2 += c.hpart[2 * i];
    target = std::max(kRefFinal * tol2, c.ref_eta * c.ref_eta * r2);
    ++cycles;
  }
  MM_HIP(hipStreamSynchronize(s));
  MM_TRY(resolve_apply_timing(c));
  c.n_pcg_it += it;
  c.ref_cycles = cycles;
  *it_out = it;
  *r2_out = r2;
  *bb_out = bb;
  return 0;
}

// Device-resident PCG: U (in/out, device) solves A U = b (device), stop when
// it >= maxiter or r.r <= tol^2 b.b (make_pcg_solve, src/poisson_solve.py:422-478).
// Returns iterations, r.r and b.b; r.r is summed on the host from the per-workgroup
// partials (written by the kernels into host-coherent memory) in a fixed order.
int pcg_dev(Ctx& c, const double* b, double* U, double tol, int maxiter, int* it_out, double* r2_out,
            double* bb_out, bool q_ready) {
  hipStream_t s = c.stream;
  const long N = c.N;
  MM_CHECK(c.lv.size() >= 1 && c.lv[0].M != nullptr, "preconditioner not set up (call mm_amg_upload or mm_precond_jacobi)");
  if (c.precision >= 2) {
    MM_CHECK(c.lv.size() > 1 && c.lv[0].Mf != nullptr && c.rxf != nullptr && c.geom_mode == kGeomCoords &&
                 c.apply_clusters,
             "the FP32 PCG needs the AMG V-cycle, the on-the-fly cluster apply and mm_set_precision(ctx, 2)");
    return pcg_ref(c, b, U, tol, maxiter, it_out, r2_out, bb_out, q_ready);
  }
  double* sc = c.dscal;  // slots 0/1: r.z ping-pong
  const int gv = red_grid(N);
  // p.Ap from the apply's per-block element energies (on-the-fly geometry), else a dot pass
  const bool fused = c.geom_mode == kGeomCoords;
  const int n_pq = fused ? apply_dot_count(c) : gv;
  if (!q_ready) {  // else the caller has left A U in pq
    vfill(c.pq, 0.0, N, s);
    MM_TRY(timed_apply(c, U, c.pq, s));
    c.n_apply++;
  }
  float* tf = c.precision >= 1 ? c.ptf : nullptr;
  double* t64 = c.precision >= 1 ? nullptr : c.pt;
  k_pcg_init<<<gv, kBlock, 0, s>>>(b, c.pq, c.pr, t64, tf, N, c.hpart_dev);
  MM_TRY(precond(c, c.pr, c.pz, c.part_rz, s));
  k_pcg_p<<<gv, kBlock, 0, s>>>(sc, c.part_rz, gv, 0, 0, 1, c.pz, c.pp, c.pq, N);
  MM_HIP(hipStreamSynchronize(s));
  double r2 = 0.0, bb = 0.0;
  for (int i = 0; i < gv; ++i) {
    r2 += c.hpart[2 * i];
    bb += c.hpart[2 * i + 1];
  }
  const double tol2 = tol * tol * bb;
  int it = 0;
  int cur = 0;  // slot holding the current r.z
  while (it < maxiter && r2 > tol2) {
    if (fused) {
      MM_TRY(timed_apply(c, c.pp, c.pq, s, c.part_pq));  // q was cleared by k_pcg_p
    } else {
      MM_TRY(timed_apply(c, c.pp, c.pq, s));
      k_dot_part<<<gv, kBlock, 0, s>>>(c.pp, c.pq, N, c.part_pq);
    }
    c.n_apply++;
    k_pcg_xr<<<gv, kBlock, 0, s>>>(sc, cur, c.part_pq, n_pq, U, c.pr, c.pp, c.pq, t64, tf, N, c.hpart_dev);
    MM_HIP(hipEventRecord(c.ev, s));
    // preconditioner and direction update run while the host waits for r.r
    MM_TRY(precond(c, c.pr, c.pz, c.part_rz, s));
    k_pcg_p<<<gv, kBlock, 0, s>>>(sc, c.part_rz, gv, 1 - cur, cur, 0, c.pz, c.pp, c.pq, N);
    cur = 1 - cur;
    ++it;
    MM_HIP(hipEventSynchronize(c.ev));
    r2 = 0.0;
    for (int i = 0; i < gv; ++i) r2 += c.hpart[i];
  }
  MM_HIP(hipStreamSynchronize(s));
  MM_TRY(resolve_apply_timing(c));
  c.n_pcg_it += it;
  *it_out = it;
  *r2_out = r2;
  *bb_out = bb;
  return 0;
}

// solve_U (src/poisson_solve.py:748-771): b = D m, PCG from the initial guess in U.
// ------------------------------------------------------------ projected initial guess
// Successive right-hand sides (the line-search trials and field steps of a minimisation) have
// nearby solutions. The candidate initial guess of the next solve is the combination of the
// last k solutions U_i whose right-hand sides b_i best match the new one,
//   x_p = sum c_i U_i,  c = argmin || b - sum c_i b_i ||  (Gram matrix G_ij = b_i . b_j),
// so that, with A U_i = b_i to the solver tolerance, b - A x_p is the least-squares residual
// (the projection technique for successive right-hand sides of Fischer, CMAME 163, 1998, in
// its residual-minimising form). The solve starts from x_p or from the given guess, whichever
// has the smaller true residual; the PCG and its stopping test are unchanged.

// per workgroup: h_i = b_i . b (i < n) and, for the newest slot j, b_j . b_i
__global__ __launch_bounds__(kBlock) void k_guess_dots(long N, const double* __restrict__ BR,
                                                       const double* __restrict__ b, int n, int j,
                                                       double* __restrict__ part) {
  double v[2 * kGuessMax];
#pragma unroll
  for (int k = 0; k < 2 * kGuessMax; ++k) v[k] = 0.0;
  for (long x = blockIdx.x * (long)kBlock + threadIdx.x; x < N; x += (long)gridDim.x * kBlock) {
    const double bx = b[x];
    const double bj = j >= 0 ? BR[(long)j * N + x] : 0.0;
#pragma unroll
    for (int i = 0; i < kGuessMax; ++i) {
      if (i >= n) break;
      const double bi = BR[(long)i * N + x];
      v[i] += bi * bx;
      v[kGuessMax + i] += bi * bj;
    }
  }
  block_partial<2 * kGuessMax>(v, part);
}

struct GuessCoef {
  double c[kGuessMax];
};

// x = sum_i c_i U_i
__global__ void k_guess_combine(long N, const double* __restrict__ UR, GuessCoef cf, int n,
                                double* __restrict__ x) {
  for (long k = blockIdx.x * (long)kBlock + threadIdx.x; k < N; k += (long)gridDim.x * kBlock) {
    double u = 0.0;
    for (int i = 0; i < n; ++i) u += cf.c[i] * UR[(long)i * N + k];
    x[k] = u;
  }
}

// per workgroup: |b - qa|^2 and |b - qb|^2
__global__ __launch_bounds__(kBlock) void k_two_residuals(long N, const double* __restrict__ b,
                                                          const double* __restrict__ qa,
                                                          const double* __restrict__ qb,
                                                          double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (long k = blockIdx.x * (long)kBlock + threadIdx.x; k < N; k += (long)gridDim.x * kBlock) {
    const double ra = b[k] - qa[k], rb = b[k] - qb[k];
    v[0] += ra * ra;
    v[1] += rb * rb;
  }
  block_partial<2>(v, part);
}

// Every workgroup sums the two residual partials (fixed order) and, where x_p has the smaller
// residual, copies x_p into U and A x_p into q.
__global__ __launch_bounds__(kBlock) void k_guess_choose(long N, const double* __restrict__ part, int n_part,
                                                         const double* __restrict__ xp,
                                                         const double* __restrict__ qp, double* __restrict__ U,
                                                         double* __restrict__ q) {
  __shared__ double sh[2][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double ab[2] = {0.0, 0.0};
  strided_sums<2, kBlock>(part, 0, n_part, ab);
  const double a = wave_red<false>(ab[0]);
  const double b = wave_red<false>(ab[1]);
  if (lane == 0) {
    sh[0][wv] = a;
    sh[1][wv] = b;
  }
  __syncthreads();
  double rf = sh[0][0], rp = sh[1][0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) {
    rf += sh[0][w];
    rp += sh[1][w];
  }
  if (!(rp < rf)) return;
  for (long k = blockIdx.x * (long)kBlock + tid; k < N; k += (long)gridDim.x * kBlock) {
    U[k] = xp[k];
    q[k] = qp[k];
  }
}

// Eigen-decomposition of a small symmetric matrix (cyclic Jacobi): a = V diag(w) V^T.
static void sym_eig(int n, std::vector<double> a, std::vector<double>& w, std::vector<double>& V) {
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; ++p) {
      diag += a[p * n + p] * a[p * n + p];
      for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    }
    if (off <= 1e-34 * diag) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (apq == 0.0) continue;
        const double th = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = cs * akp - sn * akq;
          a[k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = cs * apk - sn * aqk;
          a[q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = cs * vkp - sn * vkq;
          V[k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = a[i * n + i];
}

int set_solve_guess(Ctx& c, int k) {
  MM_CHECK(k >= 0 && k <= kGuessMax, "projection depth must be in [0, 8]");
  if (k > 0 && c.guess_U == nullptr) {
    for (double** v : {&c.guess_U, &c.guess_b}) {
      MM_HIP(hipMalloc((void**)v, (size_t)kGuessMax * c.N * sizeof(double)));
      c.allocs.push_back(*v);
    }
    MM_HIP(hipMalloc((void**)&c.guess_x, (size_t)c.N * sizeof(double)));
    c.allocs.push_back(c.guess_x);
    MM_HIP(hipHostMalloc((void**)&c.hguess, (size_t)kRedGridMax * 2 * kGuessMax * sizeof(double),
                         hipHostMallocMapped | hipHostMallocCoherent));
    MM_HIP(hipHostGetDevicePointer((void**)&c.hguess_dev, c.hguess, 0));
  }
  c.guess_k = k;
  c.guess_n = 0;
  c.guess_head = 0;
  c.guess_new = -1;
  c.guess_G.assign(kGuessMax * kGuessMax, 0.0);
  return 0;
}

int solve_U_guess(Ctx& c, const double* m, double* U, double tol, int maxiter, int* it, double* rel_res) {
  if (c.guess_k == 0) return solve_U_dev(c, m, U, tol, maxiter, it, rel_res);
  hipStream_t s = c.stream;
  const long N = c.N;
  const int gr = red_grid(N);
  vfill(c.pb, 0.0, N, s);
  MM_TRY(launch_rhs(c, m, c.pb, s));
  c.n_rhs++;
  const int n = c.guess_n;
  bool q_ready = false;
  if (n > 0) {
    k_guess_dots<<<gr, kBlock, 0, s>>>(N, c.guess_b, c.pb, n, c.guess_new, c.hguess_dev);
    MM_HIP(hipStreamSynchronize(s));
    double sums[2 * kGuessMax] = {};
    for (int w = 0; w < gr; ++w)
      for (int k = 0; k < 2 * kGuessMax; ++k) sums[k] += c.hguess[w * 2 * kGuessMax + k];
    std::vector<double>& G = c.guess_G;
    if (c.guess_new >= 0) {
      const int j = c.guess_new;
      for (int i = 0; i < n; ++i) G[j * kGuessMax + i] = G[i * kGuessMax + j] = sums[kGuessMax + i];
      c.guess_new = -1;
    }
    // least squares on the well-conditioned eigenspace of G
    std::vector<double> a(n * n), w, V;
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < n; ++q) a[i * n + q] = G[i * kGuessMax + q];
    sym_eig(n, a, w, V);
    double wmax = 0.0;
    for (double x : w) wmax = std::max(wmax, x);
    if (wmax > 0.0) {
      GuessCoef cf{};
      for (int e = 0; e < n; ++e) {
        if (!(w[e] > 1e-14 * wmax)) continue;
        double vh = 0.0;
        for (int i = 0; i < n; ++i) vh += V[i * n + e] * sums[i];
        for (int i = 0; i < n; ++i) cf.c[i] += V[i * n + e] * vh / w[e];
      }
      k_guess_combine<<<vgrid(N), kBlock, 0, s>>>(N, c.guess_U, cf, n, c.guess_x);
      // true residuals of the given guess (A U into pq) and of x_p (A x_p into pz)
      vfill(c.pq, 0.0, N, s);
      MM_TRY(timed_apply(c, U, c.pq, s));
      vfill(c.pz, 0.0, N, s);
      MM_TRY(timed_apply(c, c.guess_x, c.pz, s));
      c.n_apply += 2;
      k_two_residuals<<<gr, kBlock, 0, s>>>(N, c.pb, c.pq, c.pz, c.part_rr);
      k_guess_choose<<<vgrid(N), kBlock, 0, s>>>(N, c.part_rr, gr, c.guess_x, c.pz, U, c.pq);
      q_ready = true;
    }
  }
  double r2 = 0, bb = 0;
  MM_TRY(pcg_dev(c, c.pb, U, tol, maxiter, it, &r2, &bb, q_ready));
  *rel_res = std::sqrt(r2 / (bb + 1e-30));
  // add (U, b) to the ring; the new Gram entries are formed with the next solve's dots
  const int j = c.guess_head;
  MM_TRY(vcopy(c.guess_U + (long)j * N, U, N, s));
  MM_TRY(vcopy(c.guess_b + (long)j * N, c.pb, N, s));
  c.guess_head = (j + 1) % c.guess_k;
  c.guess_n = std::min(c.guess_n + 1, c.guess_k);
  c.guess_new = j;
  for (int i = 0; i < kGuessMax; ++i) c.guess_G[j * kGuessMax + i] = c.guess_G[i * kGuessMax + j] = 0.0;
  return 0;
}

int solve_U_dev(Ctx& c, const double* m, double* U, double tol, int maxiter, int* it, double* rel_res) {
  hipStream_t s = c.stream;
  vfill(c.pb, 0.0, c.N, s);
  MM_TRY(launch_rhs(c, m, c.pb, s));
  c.n_rhs++;
  double r2 = 0, bb = 0;
  MM_TRY(pcg_dev(c, c.pb, U, tol, maxiter, it, &r2, &bb));
  *rel_res = std::sqrt(r2 / (bb + 1e-30));
  return 0;
}

}  // namespace mm
