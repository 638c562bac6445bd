// Device helpers shared by the kernel translation units: block lookup, tet geometry,
// LDS/global scatter, and the fence-free per-workgroup reductions and their consumers.
#pragma once

#include "mm_kernels.h"

namespace mm {

// Workgroup -> element block. Workgroups are dispatched round-robin over the 8 XCDs
// (id mod 8), so XCD j gets the contiguous position range [j*chunk, (j+1)*chunk) of the
// block sequence (every block, or the subset blist). bi.b is the block id, -1 for workgroups
// past the end (they write no per-block partial).
__device__ __forceinline__ BlkInfo blk_info(const Part& p) {
  BlkInfo bi;
  const int n = p.blist ? p.nlist : p.nblk;
  const int chunk = (n + 7) >> 3;
  const int i = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
  bi.valid = i < n;
  const int b = bi.valid ? (p.blist ? p.blist[i] : i) : -1;
  bi.b = b;
  if (bi.valid) {
    bi.e0 = p.eptr[b];
    bi.e1 = p.eptr[b + 1];
    bi.n0 = p.nptr[b];
    bi.nb = p.nptr[b + 1] - bi.n0;
  } else {
    bi.e0 = bi.e1 = bi.n0 = bi.nb = 0;
  }
  return bi;
}

// Shape-function gradients and volume of one tet from LDS-staged corner coordinates.
// grad phi_b (b = 1..3) are the rows of J^-1, J = [x1-x0, x2-x0, x3-x0]:
// (c2 x c3)/det, (c3 x c1)/det, (c1 x c2)/det; grad phi_0 = -(their sum); Ve = |det|/6.
template <typename T>
__device__ __forceinline__ void tet_grads(const T (*X)[kMaxBlockNodes], int l0, int l1, int l2, int l3,
                                          T g[4][3], T& vol) {
  const T ox = X[0][l0], oy = X[1][l0], oz = X[2][l0];
  const T c1x = X[0][l1] - ox, c1y = X[1][l1] - oy, c1z = X[2][l1] - oz;
  const T c2x = X[0][l2] - ox, c2y = X[1][l2] - oy, c2z = X[2][l2] - oz;
  const T c3x = X[0][l3] - ox, c3y = X[1][l3] - oy, c3z = X[2][l3] - oz;
  const T n1x = c2y * c3z - c2z * c3y, n1y = c2z * c3x - c2x * c3z, n1z = c2x * c3y - c2y * c3x;
  const T n2x = c3y * c1z - c3z * c1y, n2y = c3z * c1x - c3x * c1z, n2z = c3x * c1y - c3y * c1x;
  const T n3x = c1y * c2z - c1z * c2y, n3y = c1z * c2x - c1x * c2z, n3z = c1x * c2y - c1y * c2x;
  const T det = c1x * n1x + c1y * n1y + c1z * n1z;
  const T inv = (T)1 / det;
  g[1][0] = n1x * inv; g[1][1] = n1y * inv; g[1][2] = n1z * inv;
  g[2][0] = n2x * inv; g[2][1] = n2y * inv; g[2][2] = n2z * inv;
  g[3][0] = n3x * inv; g[3][1] = n3y * inv; g[3][2] = n3z * inv;
  g[0][0] = -(g[1][0] + g[2][0] + g[3][0]);
  g[0][1] = -(g[1][1] + g[2][1] + g[3][1]);
  g[0][2] = -(g[1][2] + g[2][2] + g[3][2]);
  vol = fabs(det) * ((T)1 / (T)6);
}

// 1/a from the hardware estimate (v_rcp_f64) refined by two Newton-Raphson steps
// r <- r + r (1 - a r); each step squares the relative error, so the result is within
// about one ulp of 1/a for normal a (the element determinants here are O(1) in nm^3).
__device__ __forceinline__ double rcp_nr(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = fma(-a, r, 1.0);
  r = fma(r, e, r);
  e = fma(-a, r, 1.0);
  return fma(r, e, r);
}

// Scaled corner normals of one tet: n_b = det * grad phi_b (b = 1..3) are the cross products
// (c2 x c3), (c3 x c1), (c1 x c2) of the edge vectors c_b = x_b - x_0; returns det.
// With them the element stiffness is Ve grad phi_a . grad phi_b = n_a . n_b / (6 |det|).
template <typename T>
__device__ __forceinline__ T tet_normals(const T (*X)[kMaxBlockNodes], int l0, int l1, int l2, int l3, T n[3][3]) {
  const T ox = X[0][l0], oy = X[1][l0], oz = X[2][l0];
  const T c1x = X[0][l1] - ox, c1y = X[1][l1] - oy, c1z = X[2][l1] - oz;
  const T c2x = X[0][l2] - ox, c2y = X[1][l2] - oy, c2z = X[2][l2] - oz;
  const T c3x = X[0][l3] - ox, c3y = X[1][l3] - oy, c3z = X[2][l3] - oz;
  n[0][0] = c2y * c3z - c2z * c3y; n[0][1] = c2z * c3x - c2x * c3z; n[0][2] = c2x * c3y - c2y * c3x;
  n[1][0] = c3y * c1z - c3z * c1y; n[1][1] = c3z * c1x - c3x * c1z; n[1][2] = c3x * c1y - c3y * c1x;
  n[2][0] = c1y * c2z - c1z * c2y; n[2][1] = c1z * c2x - c1x * c2z; n[2][2] = c1x * c2y - c1y * c2x;
  return c1x * n[0][0] + c1y * n[0][1] + c1z * n[0][2];
}

// Laplacian stiffness action of one tet on corner values x_0..x_3 (Ve grad phi_a . grad u):
// w = sum_b (x_b - x_0) n_b = det * grad u, y_a = n_a . w / (6 |det|) for a = 1..3 and
// y_0 = -(y_1 + y_2 + y_3) (the gradients sum to zero). Returns the element energy
// sum_a x_a y_a = |w|^2 / (6 |det|).
template <typename T>
__device__ __forceinline__ T tet_laplace(const T n[3][3], T det, T d1, T d2, T d3, T y[4]) {
  const T wx = d1 * n[0][0] + d2 * n[1][0] + d3 * n[2][0];
  const T wy = d1 * n[0][1] + d2 * n[1][1] + d3 * n[2][1];
  const T wz = d1 * n[0][2] + d2 * n[1][2] + d3 * n[2][2];
  T s;
  if constexpr (sizeof(T) == 8)
    s = rcp_nr(6.0 * fabs(det));
  else
    s = __builtin_amdgcn_rcpf(6.0f * fabsf(det));
  y[1] = (n[0][0] * wx + n[0][1] * wy + n[0][2] * wz) * s;
  y[2] = (n[1][0] * wx + n[1][1] * wy + n[1][2] * wz) * s;
  y[3] = (n[2][0] * wx + n[2][1] * wy + n[2][2] * wz) * s;
  y[0] = -(y[1] + y[2] + y[3]);
  return (wx * wx + wy * wy + wz * wz) * s;
}

__device__ __forceinline__ void unpack(uint32_t lc, int& l0, int& l1, int& l2, int& l3) {
  l0 = lc & 255u;
  l1 = (lc >> 8) & 255u;
  l2 = (lc >> 16) & 255u;
  l3 = lc >> 24;
}

// Block-partial of node `f` into a global vector: plain store when no other block
// touches the node, hardware FP64 atomic add otherwise (the vector is zeroed first).
__device__ __forceinline__ void node_out(double* __restrict__ y, uint32_t f, double v) {
  const uint32_t id = f & kIdMask;
  if (f & kPrivate)
    y[id] = v;
  else
    unsafeAtomicAdd(&y[id], v);
}

template <bool MAX>
__device__ __forceinline__ double wave_red(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_down(v, o, 64);
    v = MAX ? fmax(v, u) : v + u;
  }
  return v;
}

// Compensated (Neumaier) accumulation: s + c carries the sum with the rounding error of every
// addition kept in c, so sums of large terms that cancel (the energy contributions cross zero
// during a sweep) keep their low-order digits.
struct CSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ void add(const CSum& o) {
    add(o.s);
    c += o.c;
  }
};

// Wave-level compensated reduction (lane 0 holds the total).
__device__ __forceinline__ CSum wave_red_cs(CSum v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    CSum u;
    u.s = __shfl_down(v.s, o, 64);
    u.c = __shfl_down(v.c, o, 64);
    v.add(u);
  }
  return v;
}

// Block compensated sum written as (s, c) to partial[2 * idx + {0, 1}].
__device__ void block_partial_cs(CSum v, double* __restrict__ partial, int idx) {
  __shared__ double sh[2][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  v = wave_red_cs(v);
  if (lane == 0) {
    sh[0][wv] = v.s;
    sh[1][wv] = v.c;
  }
  __syncthreads();
  if (tid == 0) {
    CSum t;
    for (int w = 0; w < kBlock / 64; ++w) {
      CSum u;
      u.s = sh[0][w];
      u.c = sh[1][w];
      t.add(u);
    }
    partial[2 * idx] = t.s;
    partial[2 * idx + 1] = t.c;
  }
}

// Block sum (or maximum) of K values written to partial[idx * K + k], idx = blockIdx.x by
// default (no cross-block traffic); sum_partials then reduces them in a fixed order.
template <int K, int NT = kBlock, bool MAX = false>
__device__ void block_partial(double (&v)[K], double* __restrict__ partial, int idx = -1) {
  __shared__ double sh[K][NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (idx < 0) idx = blockIdx.x;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_red<MAX>(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][wv] = v[k];
  __syncthreads();
  if (tid == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = sh[k][0];
      for (int w = 1; w < NT / 64; ++w) s = MAX ? fmax(s, sh[k][w]) : s + sh[k][w];
      partial[idx * K + k] = s;
    }
}

// Per-thread strided sums s[k] += part[i * K + k] for i = i0 + tid, i0 + tid + NT, ... < i1, added
// in that order. A plain strided loop waits for each load before its add; here the loads of up
// to 16 / K consecutive terms are in flight together, so the few thousand partials of an element
// kernel cost one or two memory round trips per thread instead of one per term.
template <int K, int NT>
__device__ __forceinline__ void strided_sums(const double* __restrict__ part, int i0, int i1, double (&s)[K]) {
  constexpr int B = 16 / K > 0 ? 16 / K : 1;
  for (int base = i0 + (int)threadIdx.x; base < i1; base += B * NT) {
    double v[B][K];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int i = base + u * NT;
#pragma unroll
      for (int k = 0; k < K; ++k) v[u][k] = i < i1 ? part[(long)i * K + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) s[k] += v[u][k];
  }
}

// Compensated form of strided_sums over (s, c) pairs part[2 i], part[2 i + 1], added in the
// order i = i0 + tid, i0 + tid + NT, ... (a zero pair leaves the sum unchanged).
template <int NT>
__device__ __forceinline__ void strided_csums(const double* __restrict__ part, int i0, int i1, CSum& v) {
  constexpr int B = 8;
  for (int base = i0 + (int)threadIdx.x; base < i1; base += B * NT) {
    CSum u[B];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      const int i = base + q * NT;
      u[q].s = i < i1 ? part[2L * i] : 0.0;
      u[q].c = i < i1 ? part[2L * i + 1] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < B; ++q) v.add(u[q]);
  }
}

// Sum of n partials written by an earlier kernel on the same stream (the kernel boundary
// makes them visible). Every workgroup that calls it computes the same total in the same
// fixed order, so the result is deterministic and identical across the grid; no fence or
// counter atomic is involved. Valid in every thread.
template <int NT>
__device__ double sum_partials_all(const double* __restrict__ part, int n) {
  __shared__ double sh[NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double sv[1] = {0.0};
  strided_sums<1, NT>(part, 0, n, sv);
  double s = wave_red<false>(sv[0]);
  if (lane == 0) sh[wv] = s;
  __syncthreads();
  double t = sh[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) t += sh[w];
  return t;
}

// K sums of n partials part[i * K + k] (K per workgroup of the producer), as sum_partials_all:
// same fixed order in every workgroup, valid in every thread.
template <int K, int NT>
__device__ void sum_partials_k(const double* __restrict__ part, int n, double (&out)[K]) {
  __shared__ double sh[K][NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double s[K];
#pragma unroll
  for (int k = 0; k < K; ++k) s[k] = 0.0;
  strided_sums<K, NT>(part, 0, n, s);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    s[k] = wave_red<false>(s[k]);
    if (lane == 0) sh[k][wv] = s[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = sh[k][0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) t += sh[k][w];
    out[k] = t;
  }
}

// Maximum of n partials written by an earlier kernel, valid in every thread.
template <int NT>
__device__ double max_partials_all(const double* __restrict__ part, int n) {
  __shared__ double sh[NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double m = 0.0;
  constexpr int B = 16;
  for (int base = tid; base < n; base += B * NT) {
    double v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int i = base + u * NT;
      v[u] = i < n ? part[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u) m = fmax(m, v[u]);
  }
  m = wave_red<true>(m);
  if (lane == 0) sh[wv] = m;
  __syncthreads();
  double t = sh[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) t = fmax(t, sh[w]);
  return t;
}

// Block sum of v written by thread 0 to *out (no cross-block traffic).
template <int NT>
__device__ void block_sum_to(double v, double* __restrict__ out) {
  __shared__ double sh[NT / 64];
  const int tid = threadIdx.x;
  v = wave_red<false>(v);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    double s = sh[0];
    for (int w = 1; w < NT / 64; ++w) s += sh[w];
    *out = s;
  }
}

// Grid-stride loop over [0, n) that gives each thread U items per round (kChunk unless given): the
// loads of a round (load(i) returns the item's operands by value) are all issued before the first
// body(i, operands) runs, so a wave keeps U loads per stream in flight instead of waiting once per
// item. Each
// thread still visits its items i, i + S, i + 2 S, ... (S = grid threads) in increasing order, so
// a per-thread accumulation in body adds in the same order as a plain grid-stride loop.
constexpr int kChunk = 4;
template <int U = kChunk, typename L, typename B>
__device__ __forceinline__ void grid_chunks(long n, L&& load, B&& body) {
  const long S = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + (U - 1) * S < n; i += U * S) {
    decltype(load(i)) v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = load(i + u * S);
#pragma unroll
    for (int u = 0; u < U; ++u) body(i + u * S, v[u]);
  }
  for (; i < n; i += S) body(i, load(i));
}

// Grid size for grid-stride vector kernels with reductions: at most 256 workgroups (one per CU),
// so a consumer sums at most 256 partials per value and each thread loops over more items.
constexpr int kRedGridMax = 256;
inline int red_grid(long n) {
  long g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > kRedGridMax) g = kRedGridMax;
  return (int)g;
}

}  // namespace mm
