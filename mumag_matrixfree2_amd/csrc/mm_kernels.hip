// Element kernels of the MI355X micromagnetics library: Poisson apply, demag right-hand
// side, energy + gradient, local (exchange + anisotropy) Hessian action, plus the
// deterministic reductions and nodal vector kernels they share.
//
// Every element kernel works on one element block per workgroup:
//   1. the block's distinct nodes (sorted list, flags in the top bits) are staged in LDS
//      (node values and, for on-the-fly geometry, the coordinates);
//   2. each thread processes elements of the block, gathering corner values from LDS and
//      accumulating corner contributions with LDS FP64 atomics (ds_add_f64);
//   3. the block writes each node once: a plain store for nodes no other block touches,
//      a hardware FP64 atomic add (global_atomic_add_f64) for shared nodes.
// Workgroup ids are remapped so that each XCD (ids i mod 8) runs a contiguous, spatially
// compact range of blocks and keeps its node data in its own L2.
#include <algorithm>
#include <type_traits>

#include "mm_device.h"

namespace mm {

template <typename TG>
__device__ __forceinline__ void stored_grads(const TG* __restrict__ grad, const TG* __restrict__ Ve, int Es, int e,
                                             double g[4][3], double& vol) {
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int k = 0; k < 3; ++k) g[b + 1][k] = (double)grad[(3 * b + k) * Es + e];
#pragma unroll
  for (int k = 0; k < 3; ++k) g[0][k] = -(g[1][k] + g[2][k] + g[3][k]);
  vol = (double)Ve[e];
}

// ------------------------------------------------------------------ Poisson apply (K1)
// y = A x with A the assembled Poisson matrix: element stiffness Ve grad phi_a . grad phi_b,
// rows and columns of Dirichlet nodes removed (x masked to 0 at the gather, the boundary
// row written as y = x by its owner), reg added on the diagonal of interior rows.
// With SCALED the input is x = w .* r (the level-0 pre-smoothing step of the V-cycle from a
// zero guess): it is formed during the staging and each node's owner also writes it to xo.
// With DOT (plain apply only) the workgroup also writes its share of x . (A x) to
// dot_part[logical block id]: the element energies |w|^2 / (6 |det|) plus, per owned node, reg x^2
// (interior) or x^2 (Dirichlet row), so x . A x needs no separate pass over x and y.
// NT threads per workgroup (64, 128 or 256): the block's node flags stay in registers
// between the staging and the write-back (kMaxBlockNodes / NT per thread).
template <int GEOM, bool SCALED, bool DOT, int NT>
__global__ __launch_bounds__(NT) void k_apply(Part p, Geo geo, const double* __restrict__ x,
                                              const double* __restrict__ w, double* __restrict__ xo,
                                              double* __restrict__ y, double reg, double* __restrict__ dot_part) {
  __shared__ double sx[kMaxBlockNodes];
  __shared__ double sy[kMaxBlockNodes];
  __shared__ double sX[GEOM == kGeomCoords ? 3 : 1][kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  constexpr int kNode = kMaxBlockNodes / NT;  // node slots per thread
  constexpr int kPre = 1024 / NT;             // element words loaded before the staging barrier
  uint32_t lcv[kPre], fv[kNode];
#pragma unroll
  for (int r = 0; r < kPre; ++r) {
    const int e = bi.e0 + tid + r * NT;
    lcv[r] = e < bi.e1 ? p.lconn[e] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kNode; ++r) {
    const int i = tid + r * NT;
    fv[r] = i < bi.nb ? p.nodes[bi.n0 + i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kNode; ++r) {
    const int i = tid + r * NT;
    if (i < bi.nb) {
      const uint32_t f = fv[r];
      const uint32_t id = f & kIdMask;
      double xv = x[id];
      if constexpr (SCALED) {
        xv *= w[id];
        if (f & kOwner) xo[id] = xv;
      }
      sx[i] = (f & kBoundary) ? 0.0 : xv;
      sy[i] = 0.0;
      if constexpr (GEOM == kGeomCoords) {
        sX[0][i] = geo.knt[3 * id + 0];
        sX[1][i] = geo.knt[3 * id + 1];
        sX[2][i] = geo.knt[3 * id + 2];
      }
    }
  }
  __syncthreads();
  double acc = 0.0;
  auto element = [&](int e, uint32_t lc) {
    int l0, l1, l2, l3;
    unpack(lc, l0, l1, l2, l3);
    const double x0 = sx[l0];
    const double d1 = sx[l1] - x0, d2 = sx[l2] - x0, d3 = sx[l3] - x0;
    double ya[4];
    if constexpr (GEOM == kGeomCoords) {
      double n[3][3];
      const double det = tet_normals(sX, l0, l1, l2, l3, n);
      const double en = tet_laplace(n, det, d1, d2, d3, ya);
      if constexpr (DOT) acc += en;
    } else {
      double g[4][3], vol;
      if constexpr (GEOM == kGeomStored64)
        stored_grads(geo.grad, geo.Ve, p.Es, e, g, vol);
      else
        stored_grads(geo.grad32, geo.Ve32, p.Es, e, g, vol);
      const double gx = (d1 * g[1][0] + d2 * g[2][0] + d3 * g[3][0]) * vol;
      const double gy = (d1 * g[1][1] + d2 * g[2][1] + d3 * g[3][1]) * vol;
      const double gz = (d1 * g[1][2] + d2 * g[2][2] + d3 * g[3][2]) * vol;
#pragma unroll
      for (int a = 0; a < 4; ++a) ya[a] = g[a][0] * gx + g[a][1] * gy + g[a][2] * gz;
      if constexpr (DOT) acc += d1 * ya[1] + d2 * ya[2] + d3 * ya[3];
    }
    atomicAdd(&sy[l0], ya[0]);
    atomicAdd(&sy[l1], ya[1]);
    atomicAdd(&sy[l2], ya[2]);
    atomicAdd(&sy[l3], ya[3]);
  };
#pragma unroll
  for (int r = 0; r < kPre; ++r) {
    const int e = bi.e0 + tid + r * NT;
    if (e < bi.e1) element(e, lcv[r]);
  }
  for (int e = bi.e0 + tid + kPre * NT; e < bi.e1; e += NT) element(e, p.lconn[e]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kNode; ++r) {
    const int i = tid + r * NT;
    if (i >= bi.nb) continue;
    const uint32_t f = fv[r];
    if (f & kBoundary) {
      if (f & kOwner) {
        const uint32_t id = f & kIdMask;
        const double xb = SCALED ? x[id] * w[id] : x[id];
        y[id] = xb;
        if constexpr (DOT) acc += xb * xb;
      }
      continue;
    }
    double v = sy[i];
    if (f & kOwner) {
      v += reg * sx[i];
      if constexpr (DOT) acc += reg * sx[i] * sx[i];
    }
    node_out(y, f, v);
  }
  if constexpr (DOT)
    if (bi.b >= 0) block_sum_to<NT>(acc, dot_part + bi.b);
}

// ------------------------------------------------ Poisson apply, FP32 (preconditioner)
// Same operator in FP32: node values, block-relative coordinates (origin = the block's first
// node, so edge vectors keep full FP32 precision) and element arithmetic in FP32; x and y are
// FP32. The LDS accumulator stays FP64: on gfx950 ds_add_f32 runs at a fraction of the rate of
// ds_add_f64 (measured: 53 us vs 16 us per apply on the 2.06M-tet cube). Used inside the
// mixed-precision V-cycle only.
template <bool SCALED, int NT>
__global__ __launch_bounds__(NT) void k_apply_f32(Part p, Geo geo, const float* __restrict__ x,
                                                  const double* __restrict__ r, const float* __restrict__ w,
                                                  float* __restrict__ xo, float* __restrict__ y, float reg) {
  __shared__ float sx[kMaxBlockNodes];
  __shared__ double sy[kMaxBlockNodes];
  __shared__ float sX[3][kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  constexpr int kNode = kMaxBlockNodes / NT;
  constexpr int kPre = 1024 / NT;
  uint32_t lcv[kPre], fv[kNode];
#pragma unroll
  for (int q = 0; q < kPre; ++q) {
    const int e = bi.e0 + tid + q * NT;
    lcv[q] = e < bi.e1 ? p.lconn[e] : 0u;
  }
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    fv[q] = i < bi.nb ? p.nodes[bi.n0 + i] : 0u;
  }
  const uint32_t id0 = p.nodes[bi.n0] & kIdMask;
  const double ox = geo.knt[3 * id0], oy = geo.knt[3 * id0 + 1], oz = geo.knt[3 * id0 + 2];
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    if (i < bi.nb) {
      const uint32_t f = fv[q];
      const uint32_t id = f & kIdMask;
      float xv;
      if constexpr (SCALED) {
        xv = (float)(w[id] * r[id]);
        if (f & kOwner) xo[id] = xv;
      } else {
        xv = x[id];
      }
      sx[i] = (f & kBoundary) ? 0.0f : xv;
      sy[i] = 0.0;
      sX[0][i] = (float)(geo.knt[3 * id + 0] - ox);
      sX[1][i] = (float)(geo.knt[3 * id + 1] - oy);
      sX[2][i] = (float)(geo.knt[3 * id + 2] - oz);
    }
  }
  __syncthreads();
  auto element = [&](uint32_t lc) {
    int l0, l1, l2, l3;
    unpack(lc, l0, l1, l2, l3);
    float n[3][3], ya[4];
    const float det = tet_normals(sX, l0, l1, l2, l3, n);
    const float x0 = sx[l0];
    (void)tet_laplace(n, det, sx[l1] - x0, sx[l2] - x0, sx[l3] - x0, ya);
    atomicAdd(&sy[l0], (double)ya[0]);
    atomicAdd(&sy[l1], (double)ya[1]);
    atomicAdd(&sy[l2], (double)ya[2]);
    atomicAdd(&sy[l3], (double)ya[3]);
  };
#pragma unroll
  for (int q = 0; q < kPre; ++q) {
    const int e = bi.e0 + tid + q * NT;
    if (e < bi.e1) element(lcv[q]);
  }
  for (int e = bi.e0 + tid + kPre * NT; e < bi.e1; e += NT) element(p.lconn[e]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    if (i >= bi.nb) continue;
    const uint32_t f = fv[q];
    const uint32_t id = f & kIdMask;
    if (f & kBoundary) {
      if (f & kOwner) y[id] = SCALED ? (float)(w[id] * r[id]) : x[id];
      continue;
    }
    float v = (float)sy[i];
    if (f & kOwner) v += reg * sx[i];
    if (f & kPrivate)
      y[id] = v;
    else
      unsafeAtomicAdd(&y[id], v);
  }
}

// ------------------------------------------- Poisson apply over edge-star clusters (K1)
// One cluster per thread: up to 6 tets sharing the edge (slot 0, slot 1), tet j with corners
// (slot 0, slot 1, slot 2+j, slot 2+(j+1)%6). The 8 slots' coordinates and values are read
// from LDS once, the edge vector c1 = x1 - x0 and the cross products c1 x (x_s - x0) are
// shared by the tets, and the slot sums leave through 8 LDS atomics. Adds the tets' energies
// sum_a x_a y_a to acc_e when DOT.
// Slots of a cluster's tets (bit q: slot q): tet j has corners 0, 1, 2 + j and 2 + (j + 1) % 6,
// so ring slot 2 + k belongs to tets k and k - 1 (mod 6); no tets, no slots.
__device__ __forceinline__ uint32_t cluster_slots(uint32_t wm) {
  const uint32_t t = wm & 63u;
  const uint32_t ring = (t | (t << 1) | (t >> 5)) & 63u;
  return t ? 3u | (ring << 2) : 0u;
}

template <typename T, bool DOT, int NB>
__device__ __forceinline__ void cluster_apply(uint32_t w0, uint32_t w1, uint32_t wm, const T (*sX)[NB],
                                              const T* sx, double* sy, double& acc_e) {
  int li[8];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    li[q] = (w0 >> (8 * q)) & 255u;
    li[4 + q] = (w1 >> (8 * q)) & 255u;
  }
  if constexpr (NB > 256) {
#pragma unroll
    for (int q = 0; q < 8; ++q) li[q] |= (int)((wm >> (8 + 3 * q)) & 7u) << 8;
  }
  const uint32_t used = cluster_slots(wm);
  T X[8][3], v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (used & (1u << q)) {
      X[q][0] = sX[0][li[q]];
      X[q][1] = sX[1][li[q]];
      X[q][2] = sX[2][li[q]];
      v[q] = sx[li[q]];
    } else {
      X[q][0] = X[q][1] = X[q][2] = v[q] = (T)0;
    }
  }
  // edge vectors from slot 0, value differences, shared cross products m_s = c1 x c_s
  T c[8][3], dv[8], m[8][3];
#pragma unroll
  for (int q = 1; q < 8; ++q) {
    c[q][0] = X[q][0] - X[0][0];
    c[q][1] = X[q][1] - X[0][1];
    c[q][2] = X[q][2] - X[0][2];
    dv[q] = v[q] - v[0];
  }
#pragma unroll
  for (int q = 2; q < 8; ++q) {
    m[q][0] = c[1][1] * c[q][2] - c[1][2] * c[q][1];
    m[q][1] = c[1][2] * c[q][0] - c[1][0] * c[q][2];
    m[q][2] = c[1][0] * c[q][1] - c[1][1] * c[q][0];
  }
  T acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = (T)0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    if (wm & (1u << j)) {
      const int s2 = 2 + j, s3 = 2 + (j + 1) % 6;
      T n[3][3];
      // corners (0, 1, s2, s3): n_1 = c_s2 x c_s3, n_2 = c_s3 x c_1 = -m_s3, n_3 = c_1 x c_s2 = m_s2
      n[0][0] = c[s2][1] * c[s3][2] - c[s2][2] * c[s3][1];
      n[0][1] = c[s2][2] * c[s3][0] - c[s2][0] * c[s3][2];
      n[0][2] = c[s2][0] * c[s3][1] - c[s2][1] * c[s3][0];
      n[1][0] = -m[s3][0];
      n[1][1] = -m[s3][1];
      n[1][2] = -m[s3][2];
      n[2][0] = m[s2][0];
      n[2][1] = m[s2][1];
      n[2][2] = m[s2][2];
      const T det = c[1][0] * n[0][0] + c[1][1] * n[0][1] + c[1][2] * n[0][2];
      T ya[4];
      const T en = tet_laplace(n, det, dv[1], dv[s2], dv[s3], ya);
      if constexpr (DOT) acc_e += (double)en;
      acc[0] += ya[0];
      acc[1] += ya[1];
      acc[s2] += ya[2];
      acc[s3] += ya[3];
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (used & (1u << q)) atomicAdd(&sy[li[q]], (double)acc[q]);
}

// cluster_apply walking the ring: ring slot r (slot 2 + r) is loaded when tet r - 1 needs it
// and its LDS atomic issues as soon as its two tets (r - 1, r) are done, so only two ring slots
// (and slot 0's, which waits for tet 5) are live at a time. Same arithmetic and the same
// summation order per slot as cluster_apply. k_apply_cl uses it in both precisions: FP64 110-116
// VGPRs instead of 140-142, so four 1024-node workgroups fit a CU (their LDS allows four) without
// spills; FP32 66-75 instead of 79-88, so three 512-thread 2048-node workgroups fit a CU (two
// before) and the FP32 apply on the 24.6M-tet polycrystal runs 65 instead of 72 us back-to-back.
template <typename T, bool DOT, int NB>
__device__ __forceinline__ void cluster_apply_ring(uint32_t w0, uint32_t w1, uint32_t wm, const T (*sX)[NB],
                                                   const T* sx, double* sy, double& acc_e) {
  int li[8];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    li[q] = (w0 >> (8 * q)) & 255u;
    li[4 + q] = (w1 >> (8 * q)) & 255u;
  }
  if constexpr (NB > 256) {
#pragma unroll
    for (int q = 0; q < 8; ++q) li[q] |= (int)((wm >> (8 + 3 * q)) & 7u) << 8;
  }
  const uint32_t t = wm & 63u;
  const uint32_t ring = (t | (t << 1) | (t >> 5)) & 63u;
  const T x0 = sX[0][li[0]], y0 = sX[1][li[0]], z0 = sX[2][li[0]], v0 = sx[li[0]];
  const T c1[3] = {sX[0][li[1]] - x0, sX[1][li[1]] - y0, sX[2][li[1]] - z0};
  const T d1 = sx[li[1]] - v0;
  auto load = [&](int r, T c[3], T& d, T m[3]) {
    if (ring & (1u << r)) {
      const int l = li[2 + r];
      c[0] = sX[0][l] - x0;
      c[1] = sX[1][l] - y0;
      c[2] = sX[2][l] - z0;
      d = sx[l] - v0;
    } else {
      c[0] = c[1] = c[2] = d = (T)0;
    }
    m[0] = c1[1] * c[2] - c1[2] * c[1];
    m[1] = c1[2] * c[0] - c1[0] * c[2];
    m[2] = c1[0] * c[1] - c1[1] * c[0];
  };
  T a0 = (T)0, a1 = (T)0;
  T cF[3], dF, mF[3], aF = (T)0;
  load(0, cF, dF, mF);
  T cA[3] = {cF[0], cF[1], cF[2]}, dA = dF, mA[3] = {mF[0], mF[1], mF[2]}, aA = (T)0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    T cB[3], dB, mB[3], aB = (T)0;
    if (j == 5) {
      cB[0] = cF[0], cB[1] = cF[1], cB[2] = cF[2], dB = dF, mB[0] = mF[0], mB[1] = mF[1], mB[2] = mF[2];
    } else {
      load(j + 1, cB, dB, mB);
    }
    if (t & (1u << j)) {
      T n[3][3];
      // corners (0, 1, s2, s3): n_1 = c_s2 x c_s3, n_2 = c_s3 x c_1 = -m_s3, n_3 = c_1 x c_s2 = m_s2
      n[0][0] = cA[1] * cB[2] - cA[2] * cB[1];
      n[0][1] = cA[2] * cB[0] - cA[0] * cB[2];
      n[0][2] = cA[0] * cB[1] - cA[1] * cB[0];
      n[1][0] = -mB[0];
      n[1][1] = -mB[1];
      n[1][2] = -mB[2];
      n[2][0] = mA[0];
      n[2][1] = mA[1];
      n[2][2] = mA[2];
      const T det = c1[0] * n[0][0] + c1[1] * n[0][1] + c1[2] * n[0][2];
      T ya[4];
      const T en = tet_laplace(n, det, d1, dA, dB, ya);
      if constexpr (DOT) acc_e += (double)en;
      a0 += ya[0];
      a1 += ya[1];
      aA += ya[2];
      aB += ya[3];
    }
    if (j == 0)
      aF = aA;  // ring slot 0 also belongs to tet 5
    else if (ring & (1u << j))
      atomicAdd(&sy[li[2 + j]], (double)aA);
    cA[0] = cB[0], cA[1] = cB[1], cA[2] = cB[2], dA = dB, mA[0] = mB[0], mA[1] = mB[1], mA[2] = mB[2], aA = aB;
  }
  if (ring & 1u) atomicAdd(&sy[li[2]], (double)(aF + aA));
  atomicAdd(&sy[li[0]], (double)a0);
  atomicAdd(&sy[li[1]], (double)a1);
}

// Same operator as k_apply, one workgroup per block of clusters. T = double: the FP64
// operator (x, y FP64, DOT as in k_apply); T = float: the FP32 operator of the
// mixed-precision V-cycle (block-relative FP32 coordinates, FP64 LDS accumulation).
// SCALED: input w .* r from the FP64 residual r.
// NB: distinct nodes per block (256, 512 or 1024; indices above 255 carry their high bits in
// cmask), NB / NT staged nodes per thread.
// REC (FP32 only): the staging streams the block's entry records (block-relative FP32
// coordinates with id | flags, in block order) instead of loading the node ids and then
// gathering the FP64 coordinates; the staged values are the same, so are the results. (FP64
// records, 32 bytes per entry, were measured slower than the gathers: 151 vs 117 us on the
// 24.6M-tet polycrystal, where the gathered coordinates of shared nodes hit the caches.)
template <typename T>
struct same_type {  // a non-deduced parameter type: the template argument is given or defaulted
  using type = T;
};

template <typename T, bool SCALED, bool DOT, int NT, int NB, bool REC, typename TR = double>
__global__ __launch_bounds__(NT) void k_apply_cl(Part p, Geo geo, const T* __restrict__ x,
                                                 const typename same_type<TR>::type* __restrict__ r,
                                                 const T* __restrict__ w,
                                                 T* __restrict__ xo, T* __restrict__ y, T reg,
                                                 double* __restrict__ dot_part) {
  static_assert(NB % NT == 0 && NB <= kMaxClusterBlockNodes, "block nodes per thread");
  static_assert(!REC || sizeof(T) == 4, "entry records carry FP32 coordinates");
  __shared__ T sx[NB];
  __shared__ double sy[NB];
  __shared__ T sX[3][NB];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  constexpr int kNode = NB / NT;
  uint32_t fv[kNode];
  T rc[kNode][3], xg[kNode];
  const int c0 = bi.e0 + tid;
  uint32_t cw0 = 0u, cw1 = 0u, cwm = 0u;
  if (c0 < bi.e1) {
    cw0 = p.lconn[c0];
    cw1 = p.lconn2[c0];
    cwm = p.cmask[c0];
  }
  // Every lane reads a valid entry (the block's last one past its end; a block holds at least
  // one node), so each round of loads issues without branches and waits once.
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int e = bi.n0 + min(tid + q * NT, bi.nb - 1);
    if constexpr (REC) {
      const float4 rv = p.rec32[e];
      rc[q][0] = rv.x;
      rc[q][1] = rv.y;
      rc[q][2] = rv.z;
      fv[q] = __float_as_uint(rv.w);
    } else {
      fv[q] = p.nodes[e];
    }
  }
  double ox = 0.0, oy = 0.0, oz = 0.0;
  if constexpr (sizeof(T) == 4 && !REC) {
    const uint32_t id0 = p.nodes[bi.n0] & kIdMask;
    ox = geo.knt[3 * id0];
    oy = geo.knt[3 * id0 + 1];
    oz = geo.knt[3 * id0 + 2];
  }
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const uint32_t id = fv[q] & kIdMask;
    if constexpr (SCALED)
      xg[q] = (T)(w[id] * r[id]);
    else
      xg[q] = x[id];
    if constexpr (!REC) {
      rc[q][0] = (T)(geo.knt[3 * id + 0] - ox);
      rc[q][1] = (T)(geo.knt[3 * id + 1] - oy);
      rc[q][2] = (T)(geo.knt[3 * id + 2] - oz);
    }
  }
  // slots past the block's end receive copies of its last entry, which no cluster reads
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    const uint32_t f = fv[q];
    if constexpr (SCALED)
      if ((f & kOwner) && i < bi.nb) xo[f & kIdMask] = xg[q];
    sx[i] = (f & kBoundary) ? (T)0 : xg[q];
    sy[i] = 0.0;
    sX[0][i] = rc[q][0];
    sX[1][i] = rc[q][1];
    sX[2][i] = rc[q][2];
  }
  __syncthreads();
  double acc_e = 0.0;
  if (c0 < bi.e1) {
    cluster_apply_ring<T, DOT, NB>(cw0, cw1, cwm, sX, sx, sy, acc_e);
  }
  for (int cc = c0 + NT; cc < bi.e1; cc += NT)
    cluster_apply_ring<T, DOT, NB>(p.lconn[cc], p.lconn2[cc], p.cmask[cc], sX, sx, sy, acc_e);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    if (i >= bi.nb) continue;
    const uint32_t f = fv[q];
    const uint32_t id = f & kIdMask;
    if (f & kBoundary) {
      if (f & kOwner) {
        const T xb = SCALED ? (T)(w[id] * r[id]) : x[id];
        y[id] = xb;
        if constexpr (DOT) acc_e += (double)xb * (double)xb;
      }
      continue;
    }
    double v = sy[i];
    if (f & kOwner) {
      v += (double)reg * (double)sx[i];
      if constexpr (DOT) acc_e += (double)reg * (double)sx[i] * (double)sx[i];
    }
    if (f & kPrivate)
      y[id] = (T)v;
    else if (p.seq)
      y[id] = y[id] + (T)v;  // the node's only writer in this launch (colour-sequenced blocks)
    else
      unsafeAtomicAdd(&y[id], (T)v);
  }
  if constexpr (DOT) {
    // block sum through sy, which every thread is done with after the barrier: no extra LDS, so
    // the DOT kernel fits as many workgroups per CU as the plain one
    acc_e = wave_red<false>(acc_e);
    __syncthreads();
    if ((tid & 63) == 0) sy[tid >> 6] = acc_e;
    __syncthreads();
    if (tid == 0 && bi.b >= 0) {
      double s = sy[0];
      for (int w = 1; w < NT / 64; ++w) s += sy[w];
      dot_part[bi.b] = s;
    }
  }
}

// ------------------------------- multi-RHS Poisson apply over edge-star clusters (K vectors)
// y_r = A x_r for K right-hand sides stored node-major (x[K id + r]): one staging of the block's
// node list and coordinates, and one set of edge vectors and cross products per cluster, serve
// all K value vectors; the tet normals are re-formed from them for each vector (registers, no
// memory traffic). Per vector: 8 LDS value reads and 8 LDS atomics per cluster, and K values per
// node in the write-back. 256-node blocks (LDS: (2K + 3) x 2 KiB). With DOT, dot_part[K b + r]
// receives block b's share of x_r . A x_r.
template <int K, bool DOT, int NT>
__global__ __launch_bounds__(NT) void k_apply_cl_multi(Part p, Geo geo, const double* __restrict__ x,
                                                       double* __restrict__ y, double reg,
                                                       double* __restrict__ dot_part) {
  constexpr int NB = kMaxBlockNodes;
  __shared__ double sx[K][NB];
  __shared__ double sy[K][NB];
  __shared__ double sX[3][NB];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  constexpr int kNode = NB / NT;
  uint32_t fv[kNode];
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    fv[q] = i < bi.nb ? p.nodes[bi.n0 + i] : 0u;
  }
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    if (i < bi.nb) {
      const uint32_t f = fv[q];
      const uint32_t id = f & kIdMask;
#pragma unroll
      for (int r = 0; r < K; ++r) {
        sx[r][i] = (f & kBoundary) ? 0.0 : x[(long)K * id + r];
        sy[r][i] = 0.0;
      }
      sX[0][i] = geo.knt[3 * id + 0];
      sX[1][i] = geo.knt[3 * id + 1];
      sX[2][i] = geo.knt[3 * id + 2];
    }
  }
  __syncthreads();
  double acc_e[K];
#pragma unroll
  for (int r = 0; r < K; ++r) acc_e[r] = 0.0;
  for (int cc = bi.e0 + tid; cc < bi.e1; cc += NT) {
    const uint32_t w0 = p.lconn[cc], w1 = p.lconn2[cc], wm = p.cmask[cc];
    const uint32_t used_slots = cluster_slots(wm);
    int li[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      li[q] = (w0 >> (8 * q)) & 255u;
      li[4 + q] = (w1 >> (8 * q)) & 255u;
    }
    double c[8][3], mm[8][3];
    {
      double X0[3] = {sX[0][li[0]], sX[1][li[0]], sX[2][li[0]]};
#pragma unroll
      for (int q = 1; q < 8; ++q) {
        const bool used = used_slots & (1u << q);
#pragma unroll
        for (int k = 0; k < 3; ++k) c[q][k] = used ? sX[k][li[q]] - X0[k] : 0.0;
      }
    }
#pragma unroll
    for (int q = 2; q < 8; ++q) {
      mm[q][0] = c[1][1] * c[q][2] - c[1][2] * c[q][1];
      mm[q][1] = c[1][2] * c[q][0] - c[1][0] * c[q][2];
      mm[q][2] = c[1][0] * c[q][1] - c[1][1] * c[q][0];
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
      double v[8], acc[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        v[q] = (used_slots & (1u << q)) ? sx[r][li[q]] : 0.0;
        acc[q] = 0.0;
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        if (wm & (1u << j)) {
          const int s2 = 2 + j, s3 = 2 + (j + 1) % 6;
          double n[3][3];
          n[0][0] = c[s2][1] * c[s3][2] - c[s2][2] * c[s3][1];
          n[0][1] = c[s2][2] * c[s3][0] - c[s2][0] * c[s3][2];
          n[0][2] = c[s2][0] * c[s3][1] - c[s2][1] * c[s3][0];
          n[1][0] = -mm[s3][0];
          n[1][1] = -mm[s3][1];
          n[1][2] = -mm[s3][2];
          n[2][0] = mm[s2][0];
          n[2][1] = mm[s2][1];
          n[2][2] = mm[s2][2];
          const double det = c[1][0] * n[0][0] + c[1][1] * n[0][1] + c[1][2] * n[0][2];
          double ya[4];
          const double en = tet_laplace(n, det, v[1] - v[0], v[s2] - v[0], v[s3] - v[0], ya);
          if constexpr (DOT) acc_e[r] += en;
          acc[0] += ya[0];
          acc[1] += ya[1];
          acc[s2] += ya[2];
          acc[s3] += ya[3];
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (used_slots & (1u << q)) atomicAdd(&sy[r][li[q]], acc[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kNode; ++q) {
    const int i = tid + q * NT;
    if (i >= bi.nb) continue;
    const uint32_t f = fv[q];
    const uint32_t id = f & kIdMask;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      if (f & kBoundary) {
        if (f & kOwner) {
          const double xb = x[(long)K * id + r];
          y[(long)K * id + r] = xb;
          if constexpr (DOT) acc_e[r] += xb * xb;
        }
        continue;
      }
      double val = sy[r][i];
      if (f & kOwner) {
        val += reg * sx[r][i];
        if constexpr (DOT) acc_e[r] += reg * sx[r][i] * sx[r][i];
      }
      if (f & kPrivate)
        y[(long)K * id + r] = val;
      else
        unsafeAtomicAdd(&y[(long)K * id + r], val);
    }
  }
  if constexpr (DOT) {
#pragma unroll
    for (int r = 0; r < K; ++r) {
      __syncthreads();
      if (bi.b >= 0) block_sum_to<NT>(acc_e[r], dot_part + (long)K * bi.b + r);
    }
  }
}

// Persistent form of k_apply_cl (FP64): a grid of a few workgroups per CU walks the blocks of
// its XCD's range, and while it computes the clusters of one block it already has the node
// flags, values and coordinates of its next block in flight (register double buffering), so
// the global-load latency of the staging overlaps the LDS work of the previous block. One
// node per thread (NT == kMaxBlockNodes). DOT partials go to dot_part[block id].
template <bool SCALED, bool DOT>
__global__ __launch_bounds__(kMaxBlockNodes) void k_apply_clp(Part p, Geo geo, const double* __restrict__ x,
                                                              const double* __restrict__ r,
                                                              const double* __restrict__ w, double* __restrict__ xo,
                                                              double* __restrict__ y, double reg,
                                                              double* __restrict__ dot_part) {
  constexpr int NT = kMaxBlockNodes;
  __shared__ double sx[kMaxBlockNodes];
  __shared__ double sy[kMaxBlockNodes];
  __shared__ double sX[3][kMaxBlockNodes];
  __shared__ double sred[NT / 64];
  const int tid = threadIdx.x;
  const int nxcd = (int)(gridDim.x >> 3);          // workgroups per XCD (grid is a multiple of 8)
  // b walks positions of the block sequence (every block, or the subset blist)
  const int nseq = p.blist ? p.nlist : p.nblk;
  const int chunk = (nseq + 7) >> 3;
  const int b_end = min(nseq, ((int)(blockIdx.x & 7) + 1) * chunk);
  int b = (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
  // staged registers of the current block. The loads below are unconditional (indices are
  // clamped into the block, out-of-range lanes discard their values): a load under a branch
  // makes the compiler wait for every outstanding load at the join.
  uint32_t f = 0u;
  double xv = 0.0, X0 = 0.0, X1 = 0.0, X2 = 0.0;
  int e0 = 0, e1 = 0, n0 = 0, nb = 0;
  auto fetch_meta = [&](int pos) {
    const int bb = p.blist ? p.blist[pos] : pos;
    e0 = p.eptr[bb];
    e1 = p.eptr[bb + 1];
    n0 = p.nptr[bb];
    nb = p.nptr[bb + 1] - n0;
  };
  auto fetch_node = [&]() { f = p.nodes[n0 + min(tid, nb - 1)]; };
  auto fetch_vals = [&]() {
    const uint32_t id = f & kIdMask;
    xv = SCALED ? r[id] * w[id] : x[id];
    X0 = geo.knt[3 * id];
    X1 = geo.knt[3 * id + 1];
    X2 = geo.knt[3 * id + 2];
  };
  if (b < b_end) {
    fetch_meta(b);
    fetch_node();
    fetch_vals();
  }
  for (; b < b_end; b += nxcd) {
    // stage the current block
    const int ce0 = e0, ce1 = e1, cnb = nb;
    const uint32_t cf = tid < cnb ? f : 0u;
    if (tid < cnb) {
      if (SCALED && (cf & kOwner)) xo[cf & kIdMask] = xv;
      sx[tid] = (cf & kBoundary) ? 0.0 : xv;
      sy[tid] = 0.0;
      sX[0][tid] = X0;
      sX[1][tid] = X1;
      sX[2][tid] = X2;
    }
    const double cxv = xv;
    // this thread's first cluster words are loaded before any prefetch is issued: vector loads
    // complete in order, so a later wait for them would also wait for the prefetch
    const int c0 = ce0 + tid;
    const int cl = min(c0, ce1 - 1);
    const uint32_t cw0 = p.lconn[cl], cw1 = p.lconn2[cl], cwm = c0 < ce1 ? p.cmask[cl] : 0u;
    __syncthreads();
    // next block (this one again after the last): metadata and node flags in flight during the
    // cluster work, its node values and coordinates during the write-back
    const int bn = b + nxcd < b_end ? b + nxcd : b;
    fetch_meta(bn);
    fetch_node();
    double acc_e = 0.0;
    cluster_apply<double, DOT, kMaxBlockNodes>(cw0, cw1, cwm, sX, sx, sy, acc_e);  // cwm == 0: no tets, no slots
    for (int cc = c0 + NT; cc < ce1; cc += NT)
      cluster_apply<double, DOT, kMaxBlockNodes>(p.lconn[cc], p.lconn2[cc], p.cmask[cc], sX, sx, sy, acc_e);
    fetch_vals();
    __syncthreads();
    if (tid < cnb) {
      const uint32_t id = cf & kIdMask;
      if (cf & kBoundary) {
        if (cf & kOwner) {
          y[id] = cxv;
          if constexpr (DOT) acc_e += cxv * cxv;
        }
      } else {
        double v = sy[tid];
        if (cf & kOwner) {
          v += reg * sx[tid];
          if constexpr (DOT) acc_e += reg * sx[tid] * sx[tid];
        }
        node_out(y, cf, v);
      }
    }
    if constexpr (DOT) {
      const double t = wave_red<false>(acc_e);
      if ((tid & 63) == 0) sred[tid >> 6] = t;
      __syncthreads();
      if (tid == 0) {
        double s = sred[0];
        for (int q = 1; q < NT / 64; ++q) s += sred[q];
        dot_part[p.blist ? p.blist[b] : b] = s;
      }
    }
    __syncthreads();  // LDS is restaged for the next block
  }
}

// ------------------------------------------- element loop of the per-tet magnetic kernels
// Thread tid takes elements e0 + tid, e0 + tid + kBlock, ... of its block. The words (local
// corners, material) of its first kElemPre elements are loaded before the node staging, so
// their latency overlaps it instead of opening each loop iteration; the loop over them is
// unrolled (register arrays indexed by constants) and further elements, if any, follow in a
// plain loop. One staged node per thread: a block has at most kMaxBlockNodes <= kBlock nodes.
static_assert(kMaxBlockNodes <= kBlock, "one staged node per thread");
constexpr int kElemPre = 4;
struct ElemWords {
  uint32_t lc[kElemPre];
  uint32_t mt[kElemPre];
};

__device__ __forceinline__ ElemWords elem_prefetch(const Part& p, const BlkInfo& bi, int tid) {
  ElemWords w;
#pragma unroll
  for (int q = 0; q < kElemPre; ++q) {
    // clamped into the block's element range (a launched partition has elements), so the loads
    // need no branch; for_elems skips the clamped slots
    const int e = max(bi.e0, min(bi.e0 + tid + q * kBlock, bi.e1 - 1));
    w.lc[q] = p.lconn[e];
    w.mt[q] = p.mat[e];
  }
  return w;
}

template <typename F>
__device__ __forceinline__ void for_elems(const Part& p, const BlkInfo& bi, int tid, const ElemWords& w, F&& f) {
#pragma unroll
  for (int q = 0; q < kElemPre; ++q)
    if (bi.e0 + tid + q * kBlock < bi.e1) f(w.lc[q], (int)w.mt[q]);
  for (int e = bi.e0 + tid + kElemPre * kBlock; e < bi.e1; e += kBlock) f(p.lconn[e], (int)p.mat[e]);
}

// ------------------------------------------------------------- demag right-hand side (K2)
// b_a = sum_e (Js_e Ve / 4) grad phi_a . (m_0 + m_1 + m_2 + m_3) over magnetic elements.
__global__ __launch_bounds__(kBlock) void k_rhs(Part p, Geo geo, const double* __restrict__ m,
                                                double* __restrict__ b) {
  __shared__ double sm[3][kMaxBlockNodes];
  __shared__ double sX[3][kMaxBlockNodes];
  __shared__ double sb[kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  const bool has_node = tid < bi.nb;
  const uint32_t fl = has_node ? p.nodes[bi.n0 + tid] : 0u;  // first: the staging gathers wait for it only
  const ElemWords ew = elem_prefetch(p, bi, tid);
  if (has_node) {
    const uint32_t id = fl & kIdMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sm[k][tid] = m[3 * id + k];
      sX[k][tid] = geo.knt[3 * id + k];
    }
    sb[tid] = 0.0;
  }
  __syncthreads();
  for_elems(p, bi, tid, ew, [&](uint32_t lc, int mi) {
    const double Js = geo.mt->Js[mi];
    if (Js == 0.0) return;
    int l0, l1, l2, l3;
    unpack(lc, l0, l1, l2, l3);
    double g[4][3], vol;
    tet_grads(sX, l0, l1, l2, l3, g, vol);
    const double f = Js * vol * 0.25;
    const double sx_ = (sm[0][l0] + sm[0][l1] + sm[0][l2] + sm[0][l3]) * f;
    const double sy_ = (sm[1][l0] + sm[1][l1] + sm[1][l2] + sm[1][l3]) * f;
    const double sz_ = (sm[2][l0] + sm[2][l1] + sm[2][l2] + sm[2][l3]) * f;
    atomicAdd(&sb[l0], g[0][0] * sx_ + g[0][1] * sy_ + g[0][2] * sz_);
    atomicAdd(&sb[l1], g[1][0] * sx_ + g[1][1] * sy_ + g[1][2] * sz_);
    atomicAdd(&sb[l2], g[2][0] * sx_ + g[2][1] * sy_ + g[2][2] * sz_);
    atomicAdd(&sb[l3], g[3][0] * sx_ + g[3][1] * sy_ + g[3][2] * sz_);
  });
  __syncthreads();
  if (has_node) node_out(b, fl, sb[tid]);
}

// ------------------------------------------------------------ energy + gradient (K7)
// Per magnetic element: exchange 2A Ve grad phi_a . grad m, uniaxial anisotropy
// -(2 K1 Ve/20) (sum_b m_b.k + m_a.k) k, demag (2 Js Ve/4) grad U on all four corners.
// Owners add the Zeeman term -2 M_i B. Energy E = 1/2 sum m.(g_quad) + sum m.g_z with
// compensated summation from the element terms up to the total (per-block (sum, error)
// partials, reduced by k_sum_partials_cs). Output g and E are divided by V_mag.
// T = float (precision form 3): the element energy densities and gradient terms in FP32
// on block-relative FP32 coordinates and FP32 copies of m and U; LDS accumulation, the Zeeman
// term (from the FP64 m), the compensated energy sum and the output stay FP64.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy(Part p, Geo geo, const double* __restrict__ m,
                                                   const double* __restrict__ U, double Bx, double By, double Bz,
                                                   double inv_V, double* __restrict__ g, double* __restrict__ partial) {
  __shared__ T sm[3][kMaxBlockNodes];
  __shared__ T sX[3][kMaxBlockNodes];
  __shared__ T su[kMaxBlockNodes];
  __shared__ double sg[3][kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  const bool has_node = tid < bi.nb;
  const uint32_t fl = has_node ? p.nodes[bi.n0 + tid] : 0u;  // first: the staging gathers wait for it only
  const ElemWords ew = elem_prefetch(p, bi, tid);
  // a block lies inside one segment: one 1 / V_mag for all its nodes
  const double iv = (geo.inv_V_node && bi.nb > 0) ? geo.inv_V_node[p.nodes[bi.n0] & kIdMask] : inv_V;
  CSum acc;
  double Mi = 0.0;  // M_nodal of the thread's node (Zeeman term of its owner)
  double mt[3] = {0.0, 0.0, 0.0};  // the thread's node's m in FP64
  double o[3] = {0.0, 0.0, 0.0};   // FP32: coordinates relative to the block's first node
  if constexpr (sizeof(T) == 4)
    if (bi.nb > 0) {
      const uint32_t id0 = p.nodes[bi.n0] & kIdMask;
      o[0] = geo.knt[3 * id0];
      o[1] = geo.knt[3 * id0 + 1];
      o[2] = geo.knt[3 * id0 + 2];
    }
  if (has_node) {
    const uint32_t id = fl & kIdMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mt[k] = m[3 * id + k];
      sm[k][tid] = (T)mt[k];
      sX[k][tid] = (T)(geo.knt[3 * id + k] - o[k]);
      sg[k][tid] = 0.0;
    }
    su[tid] = (T)U[id];
    Mi = geo.M_nodal[id];
  }
  __syncthreads();
  for_elems(p, bi, tid, ew, [&](uint32_t lc, int mi) {
    const T A = (T)geo.mt->A[mi], K1 = (T)geo.mt->K1[mi], Js = (T)geo.mt->Js[mi];
    const T kx = (T)geo.mt->k[mi][0], ky = (T)geo.mt->k[mi][1], kz = (T)geo.mt->k[mi][2];
    int l[4];
    unpack(lc, l[0], l[1], l[2], l[3]);
    T g4[4][3], vol;
    tet_grads<T>(sX, l[0], l[1], l[2], l[3], g4, vol);
    const T a_ve = (T)2 * A * vol, q_ve = (T)2 * K1 * vol / (T)20, j_ve = (T)2 * Js * vol / (T)4;
    // grad m_c (c = x,y,z) and grad U, relative to corner 0
    T gm[3][3], gu[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T m0 = sm[c][l[0]];
      const T d1 = sm[c][l[1]] - m0, d2 = sm[c][l[2]] - m0, d3 = sm[c][l[3]] - m0;
#pragma unroll
      for (int i = 0; i < 3; ++i) gm[c][i] = d1 * g4[1][i] + d2 * g4[2][i] + d3 * g4[3][i];
    }
    {
      const T u0 = su[l[0]];
      const T d1 = su[l[1]] - u0, d2 = su[l[2]] - u0, d3 = su[l[3]] - u0;
#pragma unroll
      for (int i = 0; i < 3; ++i) gu[i] = (d1 * g4[1][i] + d2 * g4[2][i] + d3 * g4[3][i]) * j_ve;
    }
    T v[4], vs = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      v[a] = sm[0][l[a]] * kx + sm[1][l[a]] * ky + sm[2][l[a]] * kz;
      vs += v[a];
    }
    double e_el = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const T fa = -q_ve * (vs + v[a]);
      T c3[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        c3[c] = a_ve * (g4[a][0] * gm[c][0] + g4[a][1] * gm[c][1] + g4[a][2] * gm[c][2]) + gu[c];
      c3[0] += fa * kx;
      c3[1] += fa * ky;
      c3[2] += fa * kz;
      e_el += (double)(sm[0][l[a]] * c3[0] + sm[1][l[a]] * c3[1] + sm[2][l[a]] * c3[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) atomicAdd(&sg[c][l[a]], (double)c3[c]);
    }
    acc.add(0.5 * e_el);
  });
  __syncthreads();
  if (has_node) {
    const uint32_t f = fl;
    const uint32_t id = f & kIdMask;
    const int i = tid;
    double v0 = sg[0][i], v1 = sg[1][i], v2 = sg[2][i];
    if (f & kOwner) {
      const double z0 = -2.0 * Mi * Bx, z1 = -2.0 * Mi * By, z2 = -2.0 * Mi * Bz;
      v0 += z0;
      v1 += z1;
      v2 += z2;
      acc.add(mt[0] * z0 + mt[1] * z1 + mt[2] * z2);
    }
    if (f & kPrivate) {
      g[3 * id + 0] = v0 * iv;
      g[3 * id + 1] = v1 * iv;
      g[3 * id + 2] = v2 * iv;
    } else {
      unsafeAtomicAdd(&g[3 * id + 0], v0 * iv);
      unsafeAtomicAdd(&g[3 * id + 1], v1 * iv);
      unsafeAtomicAdd(&g[3 * id + 2], v2 * iv);
    }
  }
  acc.s *= iv;
  acc.c *= iv;
  if (bi.b >= 0) block_partial_cs(acc, partial, bi.b);  // by block id: a segment's blocks are contiguous
}

// out[k] = sum_b partial[b * K + k], one workgroup, fixed order
template <int K>
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ partial, int n,
                                                         double* __restrict__ out) {
  __shared__ double sh[K][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  strided_sums<K, kBlock>(partial, 0, n, v);
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_red<false>(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][wv] = v[k];
  __syncthreads();
  if (tid == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = sh[k][0];
      for (int w = 1; w < kBlock / 64; ++w) s += sh[k][w];
      out[k] = s;
    }
}

// out[0] = compensated sum of n (s, c) partial pairs, one workgroup, fixed order
__global__ __launch_bounds__(kBlock) void k_sum_partials_cs(const double* __restrict__ partial, int n,
                                                            double* __restrict__ out) {
  __shared__ double sh[2][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  CSum v;
  strided_csums<kBlock>(partial, 0, n, v);
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
    out[0] = t.s + t.c;
  }
}

// Per-segment sums of block partials written by logical block: workgroup g reduces blocks
// blk_off[g] .. blk_off[g + 1] (a segment's blocks are contiguous) in a fixed order into
// out[g * stride + slot + k], K values per block; CS: (s, c) pairs with compensation (K = 1).
template <int K, bool CS>
__global__ __launch_bounds__(kBlock) void k_seg_block_sums(const int* __restrict__ blk_off,
                                                           const double* __restrict__ part, double* __restrict__ out,
                                                           int stride, int slot) {
  const int g = blockIdx.x;
  const int b0 = blk_off[g], b1 = blk_off[g + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if constexpr (CS) {
    __shared__ double sh[2][kBlock / 64];
    CSum v;
    strided_csums<kBlock>(part, b0, b1, v);
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
      out[(long)g * stride + slot] = t.s + t.c;
    }
  } else {
    __shared__ double sh[K][kBlock / 64];
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    strided_sums<K, kBlock>(part, b0, b1, v);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      v[k] = wave_red<false>(v[k]);
      if (lane == 0) sh[k][wv] = v[k];
    }
    __syncthreads();
    if (tid == 0)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        double t = sh[k][0];
        for (int w = 1; w < kBlock / 64; ++w) t += sh[k][w];
        out[(long)g * stride + slot + k] = t;
      }
  }
}

// --------------------------------------------------- local Hessian action C v (K8)
// C v = (K_ex + K_an) v / V_mag: exchange and anisotropy parts of K7 without demag/Zeeman.
// T = float: element arithmetic in FP32 on block-relative FP32 coordinates and FP32 copies of v
// (the mixed-precision forms' Steihaug loop), accumulation in FP64. TV: the type of v and of the
// output (FP32 in the mixed forms' Steihaug loop, whose vectors are FP32).
template <typename T, typename TV = double>
__global__ __launch_bounds__(kBlock) void k_local(Part p, Geo geo, const TV* __restrict__ vin,
                                                  double inv_V, TV* __restrict__ out,
                                                  const double* __restrict__ gate) {
  if (gate != nullptr && *gate == 0.0) return;
  __shared__ T sm[3][kMaxBlockNodes];
  __shared__ T sX[3][kMaxBlockNodes];
  __shared__ double sg[3][kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  const bool has_node = tid < bi.nb;
  const uint32_t fl = has_node ? p.nodes[bi.n0 + tid] : 0u;  // first: the staging gathers wait for it only
  const ElemWords ew = elem_prefetch(p, bi, tid);
  const double iv = (geo.inv_V_node && bi.nb > 0) ? geo.inv_V_node[p.nodes[bi.n0] & kIdMask] : inv_V;
  double o[3] = {0.0, 0.0, 0.0};  // FP32: coordinates relative to the block's first node
  if constexpr (sizeof(T) == 4)
    if (bi.nb > 0) {
      const uint32_t id0 = p.nodes[bi.n0] & kIdMask;
      o[0] = geo.knt[3 * id0];
      o[1] = geo.knt[3 * id0 + 1];
      o[2] = geo.knt[3 * id0 + 2];
    }
  if (has_node) {
    const uint32_t id = fl & kIdMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      sm[k][tid] = (T)vin[3 * id + k];
      sX[k][tid] = (T)(geo.knt[3 * id + k] - o[k]);
      sg[k][tid] = 0.0;
    }
  }
  __syncthreads();
  for_elems(p, bi, tid, ew, [&](uint32_t lc, int mi) {
    const T A = (T)geo.mt->A[mi], K1 = (T)geo.mt->K1[mi];
    const T kx = (T)geo.mt->k[mi][0], ky = (T)geo.mt->k[mi][1], kz = (T)geo.mt->k[mi][2];
    int l[4];
    unpack(lc, l[0], l[1], l[2], l[3]);
    T g4[4][3], vol;
    tet_grads<T>(sX, l[0], l[1], l[2], l[3], g4, vol);
    const T a_ve = (T)2 * A * vol, q_ve = (T)2 * K1 * vol / (T)20;
    T gm[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T m0 = sm[c][l[0]];
      const T d1 = sm[c][l[1]] - m0, d2 = sm[c][l[2]] - m0, d3 = sm[c][l[3]] - m0;
#pragma unroll
      for (int i = 0; i < 3; ++i) gm[c][i] = (d1 * g4[1][i] + d2 * g4[2][i] + d3 * g4[3][i]) * a_ve;
    }
    T v[4], vs = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      v[a] = sm[0][l[a]] * kx + sm[1][l[a]] * ky + sm[2][l[a]] * kz;
      vs += v[a];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const T fa = -q_ve * (vs + v[a]);
      atomicAdd(&sg[0][l[a]], (double)(g4[a][0] * gm[0][0] + g4[a][1] * gm[0][1] + g4[a][2] * gm[0][2] + fa * kx));
      atomicAdd(&sg[1][l[a]], (double)(g4[a][0] * gm[1][0] + g4[a][1] * gm[1][1] + g4[a][2] * gm[1][2] + fa * ky));
      atomicAdd(&sg[2][l[a]], (double)(g4[a][0] * gm[2][0] + g4[a][1] * gm[2][1] + g4[a][2] * gm[2][2] + fa * kz));
    }
  });
  __syncthreads();
  if (has_node) {
    const uint32_t id = fl & kIdMask;
    if (fl & kPrivate) {
#pragma unroll
      for (int k = 0; k < 3; ++k) out[3 * id + k] = (TV)(sg[k][tid] * iv);
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&out[3 * id + k], (TV)(sg[k][tid] * iv));
    }
  }
}

// ------------------------------------------------------------------ geometry (set-up)
// Stored-gradient streams for the streamed-geometry variant of the apply, and Ve.
__global__ void k_geometry(int E, int Es, const double* __restrict__ knt, const int* __restrict__ conn,
                           double* __restrict__ grad, double* __restrict__ Ve, float* __restrict__ grad32,
                           float* __restrict__ Ve32) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double c[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) c[a][k] = knt[3 * conn[4 * e + a] + k];
  const double c1x = c[1][0] - c[0][0], c1y = c[1][1] - c[0][1], c1z = c[1][2] - c[0][2];
  const double c2x = c[2][0] - c[0][0], c2y = c[2][1] - c[0][1], c2z = c[2][2] - c[0][2];
  const double c3x = c[3][0] - c[0][0], c3y = c[3][1] - c[0][1], c3z = c[3][2] - c[0][2];
  const double n[3][3] = {{c2y * c3z - c2z * c3y, c2z * c3x - c2x * c3z, c2x * c3y - c2y * c3x},
                          {c3y * c1z - c3z * c1y, c3z * c1x - c3x * c1z, c3x * c1y - c3y * c1x},
                          {c1y * c2z - c1z * c2y, c1z * c2x - c1x * c2z, c1x * c2y - c1y * c2x}};
  const double det = c1x * n[0][0] + c1y * n[0][1] + c1z * n[0][2];
  const double inv = 1.0 / det;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double v = n[b][k] * inv;
      grad[(3 * b + k) * Es + e] = v;
      grad32[(3 * b + k) * Es + e] = (float)v;
    }
  Ve[e] = fabs(det) / 6.0;
  Ve32[e] = (float)(fabs(det) / 6.0);
}

// ------------------------------------------------------------- diag(A) (set-up)
// d_a = sum_e Ve |grad phi_a|^2 + reg on interior rows, 1 on Dirichlet rows
// (``assemble_diag`` semantics on the masked, regularised matrix).
__global__ __launch_bounds__(kBlock) void k_diag(Part p, Geo geo, double reg, double* __restrict__ d) {
  __shared__ double sX[3][kMaxBlockNodes];
  __shared__ double sd[kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  if (!bi.valid) return;
  const int tid = threadIdx.x;
  for (int i = tid; i < bi.nb; i += kBlock) {
    const uint32_t id = p.nodes[bi.n0 + i] & kIdMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) sX[k][i] = geo.knt[3 * id + k];
    sd[i] = 0.0;
  }
  __syncthreads();
  for (int e = bi.e0 + tid; e < bi.e1; e += kBlock) {
    int l[4];
    unpack(p.lconn[e], l[0], l[1], l[2], l[3]);
    double g4[4][3], vol;
    tet_grads(sX, l[0], l[1], l[2], l[3], g4, vol);
#pragma unroll
    for (int a = 0; a < 4; ++a) atomicAdd(&sd[l[a]], vol * (g4[a][0] * g4[a][0] + g4[a][1] * g4[a][1] + g4[a][2] * g4[a][2]));
  }
  __syncthreads();
  for (int i = tid; i < bi.nb; i += kBlock) {
    const uint32_t f = p.nodes[bi.n0 + i];
    if (f & kBoundary) {
      if (f & kOwner) d[f & kIdMask] = 1.0;
      continue;
    }
    node_out(d, f, sd[i] + ((f & kOwner) ? reg : 0.0));
  }
}

// ---------------------------------------------------------- volume-averaged J along h
// out[0] = sum_e Ve Js_e mean_a(m_a).h, out[1] = sum_{Js_e > 0} Ve (``src/hysteresis_loop.py:153-190``).
__global__ __launch_bounds__(kBlock) void k_jpar(Part p, Geo geo, const double* __restrict__ m, double hx, double hy,
                                                 double hz, double* __restrict__ partial) {
  __shared__ double sX[3][kMaxBlockNodes];
  __shared__ double sh[kMaxBlockNodes];
  const BlkInfo bi = blk_info(p);
  const int tid = threadIdx.x;
  for (int i = tid; i < bi.nb; i += kBlock) {
    const uint32_t id = p.nodes[bi.n0 + i] & kIdMask;
#pragma unroll
    for (int k = 0; k < 3; ++k) sX[k][i] = geo.knt[3 * id + k];
    sh[i] = m[3 * id] * hx + m[3 * id + 1] * hy + m[3 * id + 2] * hz;
  }
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  for (int e = bi.e0 + tid; e < bi.e1; e += kBlock) {
    const double Js = geo.mt->Js[p.mat[e]];
    if (!(Js > 0.0)) continue;
    int l[4];
    unpack(p.lconn[e], l[0], l[1], l[2], l[3]);
    double g4[4][3], vol;
    tet_grads(sX, l[0], l[1], l[2], l[3], g4, vol);
    acc[0] += vol * Js * 0.25 * (sh[l[0]] + sh[l[1]] + sh[l[2]] + sh[l[3]]);
    acc[1] += vol;
  }
  if (bi.b >= 0) block_partial<2>(acc, partial, bi.b);
}

// ------------------------------------------------------------------------ launchers
static inline int part_len(const Part& p) { return p.blist ? p.nlist : p.nblk; }
static inline int grid_for(const Part& p) { return 8 * ((part_len(p) + 7) / 8); }

// Calls f(part, grid) for one launch of the cluster apply over p, or, with the colour-sequenced
// scatter on, for one launch per colour (the blocks of that colour as p's block list), in colour
// order on the caller's stream.
template <typename F>
static void for_cluster_launches(const Ctx& c, const Part& p, F&& f) {
  if (!c.apply_seq || p.blist != nullptr || c.h_colour_off.size() < 2) {
    f(p, grid_for(p));
    return;
  }
  for (size_t k = 0; k + 1 < c.h_colour_off.size(); ++k) {
    Part q = p;
    q.blist = c.d_colour_blk + c.h_colour_off[k];
    q.nlist = c.h_colour_off[k + 1] - c.h_colour_off[k];
    q.seq = 1;
    if (q.nlist > 0) f(q, grid_for(q));
  }
}

// Calls f(std::integral_constant<int, NT>) for the context's workgroup size NT (64, 128, 256).
template <typename F>
static void with_nt(int nt, F&& f) {
  switch (nt) {
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 128: f(std::integral_constant<int, 128>{}); break;
    default: f(std::integral_constant<int, 256>{}); break;
  }
}

// workgroup size and block node count of the cluster apply: blocks of 512 or 1024 nodes run
// 256 threads, blocks of 2048 nodes 512 threads; blocks of more than 256 nodes stage either way
// in the FP32 kernels (gathers or entry records)
template <typename F>
static void with_cl_shape(const Ctx& c, F&& f) {
  using nt256 = std::integral_constant<int, 256>;
  using rec = std::true_type;
  using gather = std::false_type;
  if (c.cl_cap == 2048) {
    using nt512 = std::integral_constant<int, 512>;
    if (c.cl_records)
      f(nt512{}, std::integral_constant<int, 2048>{}, rec{});
    else
      f(nt512{}, std::integral_constant<int, 2048>{}, gather{});
  } else if (c.cl_cap == 1024) {
    if (c.cl_records)
      f(nt256{}, std::integral_constant<int, 1024>{}, rec{});
    else
      f(nt256{}, std::integral_constant<int, 1024>{}, gather{});
  } else if (c.cl_cap == 512) {
    if (c.cl_records)
      f(nt256{}, std::integral_constant<int, 512>{}, rec{});
    else
      f(nt256{}, std::integral_constant<int, 512>{}, gather{});
  } else {
    with_nt(c.elem_threads, [&](auto nt) { f(nt, std::integral_constant<int, kMaxBlockNodes>{}, gather{}); });
  }
}

// persistent grid: apply_persist workgroups per CU, a multiple of 8 (one share per XCD)
static int persist_grid(const Ctx& c) {
  const int g = 8 * std::max(1, (c.num_cu * c.apply_persist) / 8);
  return std::min(g, 8 * ((part_len(c.part_cl) + 7) / 8));
}

int apply_dot_count(const Ctx& c) { return (c.apply_clusters ? c.part_cl : c.part_all).nblk; }

int launch_apply(const Ctx& c, int geom, const double* x, double* y, hipStream_t s, double* dot_part) {
  MM_CHECK(dot_part == nullptr || geom == kGeomCoords, "x.Ax partials need the on-the-fly apply");
  if (geom == kGeomCoords && c.apply_clusters && c.apply_persist > 0) {
    const Part& p = c.part_cl;
    if (part_len(p) == 0) return 0;
    if (dot_part)
      k_apply_clp<false, true><<<persist_grid(c), kMaxBlockNodes, 0, s>>>(p, c.geo, x, nullptr, nullptr, nullptr, y,
                                                                        c.reg, dot_part);
    else
      k_apply_clp<false, false><<<persist_grid(c), kMaxBlockNodes, 0, s>>>(p, c.geo, x, nullptr, nullptr, nullptr, y,
                                                                         c.reg, nullptr);
    return 0;
  }
  if (geom == kGeomCoords && c.apply_clusters) {
    if (part_len(c.part_cl) == 0) return 0;
    for_cluster_launches(c, c.part_cl, [&](const Part& p, int grid) {
      with_cl_shape(c, [&](auto nt, auto nb, auto) {  // FP64 blocks always gather
        constexpr int NT = decltype(nt)::value, NB = decltype(nb)::value;
        if (dot_part)
          k_apply_cl<double, false, true, NT, NB, false><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, nullptr,
                                                                            y, c.reg, dot_part);
        else
          k_apply_cl<double, false, false, NT, NB, false><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, nullptr,
                                                                             y, c.reg, nullptr);
      });
    });
    return 0;
  }
  const Part& p = c.part_all;
  if (part_len(p) == 0) return 0;
  const int grid = grid_for(p);
  with_nt(c.elem_threads, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    if (dot_part)
      k_apply<kGeomCoords, false, true, NT><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, y, c.reg, dot_part);
    else if (geom == kGeomCoords)
      k_apply<kGeomCoords, false, false, NT><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, y, c.reg, nullptr);
    else if (geom == kGeomStored64)
      k_apply<kGeomStored64, false, false, NT><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, y, c.reg, nullptr);
    else
      k_apply<kGeomStored32, false, false, NT><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, y, c.reg, nullptr);
  });
  return 0;
}

int launch_apply_scaled(const Ctx& c, const double* r, const double* w, double* xo, double* y, hipStream_t s) {
  const Part& p = c.apply_clusters ? c.part_cl : c.part_all;
  if (part_len(p) == 0) return 0;
  if (c.apply_clusters && c.apply_persist > 0) {
    k_apply_clp<true, false><<<persist_grid(c), kMaxBlockNodes, 0, s>>>(p, c.geo, nullptr, r, w, xo, y, c.reg,
                                                                      nullptr);
    return 0;
  }
  const int grid = grid_for(p);
  if (c.apply_clusters) {
    for_cluster_launches(c, p, [&](const Part& q, int g) {
      with_cl_shape(c, [&](auto nt, auto nb, auto) {  // FP64 blocks always gather
        constexpr int NT = decltype(nt)::value, NB = decltype(nb)::value;
        k_apply_cl<double, true, false, NT, NB, false><<<g, NT, 0, s>>>(q, c.geo, nullptr, r, w, xo, y, c.reg, nullptr);
      });
    });
    return 0;
  }
  with_nt(c.elem_threads, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    k_apply<kGeomCoords, true, false, NT><<<grid, NT, 0, s>>>(p, c.geo, r, w, xo, y, c.reg, nullptr);
  });
  return 0;
}

int launch_apply_f32(const Ctx& c, const float* x, float* y, hipStream_t s) {
  const Part& p = c.apply_clusters ? c.part_cl : c.part_all;
  if (part_len(p) == 0) return 0;
  const int grid = grid_for(p);
  if (c.apply_clusters) {
    for_cluster_launches(c, p, [&](const Part& q, int g) {
      with_cl_shape(c, [&](auto nt, auto nb, auto rec) {
        constexpr int NT = decltype(nt)::value, NB = decltype(nb)::value;
        constexpr bool REC = decltype(rec)::value;
        k_apply_cl<float, false, false, NT, NB, REC><<<g, NT, 0, s>>>(q, c.geo, x, nullptr, nullptr, nullptr, y,
                                                                   (float)c.reg, nullptr);
      });
    });
    return 0;
  }
  with_nt(c.elem_threads, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    k_apply_f32<false, NT><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, nullptr, y, (float)c.reg);
  });
  return 0;
}

int launch_apply_multi(const Ctx& c, int K, const double* x, double* y, double* dot_part, hipStream_t s) {
  MM_CHECK(c.apply_clusters && c.cl_cap == kMaxBlockNodes,
           "the multi-vector apply runs the cluster apply on 256-node blocks");
  MM_CHECK(K == 1 || K == 2 || K == 4 || K == 8, "vectors per multi-vector apply: 1, 2, 4 or 8");
  const Part& p = c.part_cl;
  if (part_len(p) == 0) return 0;
  const int grid = grid_for(p);
  auto go = [&](auto kc) {
    constexpr int KK = decltype(kc)::value;
    if (dot_part)
      k_apply_cl_multi<KK, true, 256><<<grid, 256, 0, s>>>(p, c.geo, x, y, c.reg, dot_part);
    else
      k_apply_cl_multi<KK, false, 256><<<grid, 256, 0, s>>>(p, c.geo, x, y, c.reg, nullptr);
  };
  if (K == 1)
    go(std::integral_constant<int, 1>{});
  else if (K == 2)
    go(std::integral_constant<int, 2>{});
  else if (K == 4)
    go(std::integral_constant<int, 4>{});
  else
    go(std::integral_constant<int, 8>{});
  return 0;
}

// FP32 cluster apply of the FP32 PCG: y = A x with per-block partials of x . A x (FP64 sums of
// the FP32 element energies) in dot_part
int launch_apply_f32_dot(const Ctx& c, const float* x, float* y, double* dot_part, hipStream_t s) {
  MM_CHECK(c.apply_clusters, "the FP32 PCG needs the cluster apply");
  if (part_len(c.part_cl) == 0) return 0;
  for_cluster_launches(c, c.part_cl, [&](const Part& p, int grid) {
    with_cl_shape(c, [&](auto nt, auto nb, auto rec) {
      constexpr int NT = decltype(nt)::value, NB = decltype(nb)::value;
      constexpr bool REC = decltype(rec)::value;
      k_apply_cl<float, false, true, NT, NB, REC><<<grid, NT, 0, s>>>(p, c.geo, x, nullptr, nullptr, nullptr, y,
                                                                 (float)c.reg, dot_part);
    });
  });
  return 0;
}

// y = A (w .* r) with FP32 r (the V-cycle of the FP32 PCG)
int launch_apply_f32_scaled_f(const Ctx& c, const float* r, const float* w, float* xo, float* y, hipStream_t s) {
  MM_CHECK(c.apply_clusters, "the FP32 PCG needs the cluster apply");
  if (part_len(c.part_cl) == 0) return 0;
  for_cluster_launches(c, c.part_cl, [&](const Part& p, int grid) {
    with_cl_shape(c, [&](auto nt, auto nb, auto rec) {
      constexpr int NT = decltype(nt)::value, NB = decltype(nb)::value;
      constexpr bool REC = decltype(rec)::value;
      k_apply_cl<float, true, false, NT, NB, REC, float><<<grid, NT, 0, s>>>(p, c.geo, nullptr, r, w, xo, y,
                                                                        (float)c.reg, nullptr);
    });
  });
  return 0;
}

int launch_apply_f32_scaled(const Ctx& c, const double* r, const float* w, float* xo, float* y, hipStream_t s) {
  const Part& p = c.apply_clusters ? c.part_cl : c.part_all;
  if (part_len(p) == 0) return 0;
  const int grid = grid_for(p);
  if (c.apply_clusters) {
    for_cluster_launches(c, p, [&](const Part& q, int g) {
      with_cl_shape(c, [&](auto nt, auto nb, auto rec) {
        constexpr int NT = decltype(nt)::value, NB = decltype(nb)::value;
        constexpr bool REC = decltype(rec)::value;
        k_apply_cl<float, true, false, NT, NB, REC><<<g, NT, 0, s>>>(q, c.geo, nullptr, r, w, xo, y, (float)c.reg,
                                                                  nullptr);
      });
    });
    return 0;
  }
  with_nt(c.elem_threads, [&](auto nt) {
    constexpr int NT = decltype(nt)::value;
    k_apply_f32<true, NT><<<grid, NT, 0, s>>>(p, c.geo, nullptr, r, w, xo, y, (float)c.reg);
  });
  return 0;
}

int launch_rhs(const Ctx& c, const double* m, double* b, hipStream_t s) {
  const Part& p = c.part_mag;
  if (part_len(p) == 0) return 0;
  k_rhs<<<grid_for(p), kBlock, 0, s>>>(p, c.geo, m, b);
  return 0;
}

int launch_energy_buf(const Ctx& c, const double* m, const double* U, const double B[3], double* g, double* E_out,
                      double* partial, hipStream_t s) {
  const Part& p = c.part_mag;
  if (part_len(p) == 0) return 0;
  const int grid = grid_for(p);
  if (c.precision >= 3)
    k_energy<float><<<grid, kBlock, 0, s>>>(p, c.geo, m, U, B[0], B[1], B[2], 1.0 / c.V_mag, g, partial);
  else
    k_energy<double><<<grid, kBlock, 0, s>>>(p, c.geo, m, U, B[0], B[1], B[2], 1.0 / c.V_mag, g, partial);
  k_sum_partials_cs<<<1, kBlock, 0, s>>>(partial, p.nblk, E_out);  // 2 x nblk partials by block id
  return 0;
}

int launch_energy(const Ctx& c, const double* m, const double* U, const double B[3], double* g, double* E_out,
                  hipStream_t s) {
  return launch_energy_buf(c, m, U, B, g, E_out, c.partial, s);
}

int launch_local(const Ctx& c, const double* v, double* out, const double* gate, hipStream_t s) {
  const Part& p = c.part_mag;
  if (part_len(p) == 0) return 0;
  if (c.precision >= 1)
    k_local<float><<<grid_for(p), kBlock, 0, s>>>(p, c.geo, v, 1.0 / c.V_mag, out, gate);
  else
    k_local<double><<<grid_for(p), kBlock, 0, s>>>(p, c.geo, v, 1.0 / c.V_mag, out, gate);
  return 0;
}

int launch_local(const Ctx& c, const float* v, float* out, const double* gate, hipStream_t s) {
  const Part& p = c.part_mag;
  if (part_len(p) == 0) return 0;
  k_local<float, float><<<grid_for(p), kBlock, 0, s>>>(p, c.geo, v, 1.0 / c.V_mag, out, gate);
  return 0;
}

int launch_seg_block_sums(const Ctx& c, const double* part, int K, bool cs, double* out, int stride, int slot,
                          hipStream_t s) {
  if (c.d_seg_blk_mag == nullptr) return -1;
  if (cs)
    k_seg_block_sums<1, true><<<c.nseg, kBlock, 0, s>>>(c.d_seg_blk_mag, part, out, stride, slot);
  else if (K == 1)
    k_seg_block_sums<1, false><<<c.nseg, kBlock, 0, s>>>(c.d_seg_blk_mag, part, out, stride, slot);
  else
    k_seg_block_sums<2, false><<<c.nseg, kBlock, 0, s>>>(c.d_seg_blk_mag, part, out, stride, slot);
  return 0;
}

int launch_geometry(const Ctx& c, const int* conn_dev, hipStream_t s) {
  const int E = c.E;
  if (E == 0) return 0;
  k_geometry<<<(E + 255) / 256, 256, 0, s>>>(E, c.part_all.Es, c.geo.knt, conn_dev, const_cast<double*>(c.geo.grad),
                                             const_cast<double*>(c.geo.Ve), const_cast<float*>(c.geo.grad32),
                                             const_cast<float*>(c.geo.Ve32));
  return 0;
}

int launch_diag(const Ctx& c, double* d, hipStream_t s) {
  const Part& p = c.part_all;
  if (part_len(p) == 0) return 0;
  k_diag<<<grid_for(p), kBlock, 0, s>>>(p, c.geo, c.reg, d);
  return 0;
}

int launch_jpar(const Ctx& c, const double* m, const double h[3], double* out, hipStream_t s) {
  const Part& p = c.part_mag;
  if (part_len(p) == 0) return 0;
  const int grid = grid_for(p);
  k_jpar<<<grid, kBlock, 0, s>>>(p, c.geo, m, h[0], h[1], h[2], c.partial);
  k_sum_partials<2><<<1, kBlock, 0, s>>>(c.partial, p.nblk, out);  // partials by block id
  return 0;
}

}  // namespace mm
