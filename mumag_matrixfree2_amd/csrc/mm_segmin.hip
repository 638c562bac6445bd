// Field-step minimiser for every segment of a segmented context at once: the independent
// systems (the grains of a grain set) share every kernel launch, while each keeps the scalars,
// line-search state and convergence test of its own run of preconditioned Cohen-HS conjugate
// gradients (src/minimizers.py:950-1073) with the curvilinear Armijo line search
// (src/minimizers.py:265-385), the Steihaug-PCG preconditioner (src/minimizers.py:393-492) and
// the Gill-Murray test (src/minimizers.py:160-172) -- the same steps as the single-system
// driver, per segment.
//
// Work decomposition: vector kernels use the segmented PCG's layout (workgroup w covers
// kSegChunk nodes of segment wg_seg[w]), so each per-workgroup partial belongs to one segment
// and a consumer sums the partials of its own segment in a fixed order. Per-segment scalars
// live in a device table (kSegSlots per segment); the host sets per-segment inputs (iterating,
// restart, trial step, accept) and reads the table once per direction and once per line-search
// round. A segment that has converged keeps its state: every write of a minimiser vector is
// masked to the segments that are still iterating or taking a trial.
#include <chrono>
#include <cmath>
#include <type_traits>
#include <utility>

#include "mm_device.h"
#include "mm_solver.h"

namespace mm {

int alloc_min_state(Ctx& c);

constexpr int kSegChunkM = 4 * kBlock;  // nodes per workgroup (the segmented PCG's chunk)

enum SegSlot {
  T_GN = 0, T_GS = 1, T_PG = 2, T_PCIT = 3, T_E = 4, T_DM = 5, T_YY = 6, T_ZZ = 7, T_YG = 8,
  T_PAP = 9, T_ALPHA = 10, T_DQ = 11, T_DN = 12,
  T_BANK0 = 16,  // Steihaug banks at 16..21 and 24..29
};
constexpr int kSegSlots = 32;
enum SegIn { I_WANT = 0, I_RESTART = 1, I_TRIAL = 2, I_STEP = 3, I_ACCEPT = 4 };
constexpr int kSegIn = 8;
enum PcBankS { SB_RHO = 0, SB_Q = 1, SB_IT = 2, SB_DONE = 3, SB_ACT = 4, SB_TGT = 5 };
__host__ __device__ constexpr int seg_bank(int k) { return T_BANK0 + 8 * (k & 1); }

struct SegM {
  const int* node_off;  // nseg + 1
  const int* wg_off;    // nseg + 1 (workgroups of each segment in this launch)
  const int* wg_seg;    // per workgroup
  const int* wg_chunk;  // per workgroup
  double* tab;          // nseg x kSegSlots
  const double* in;     // nseg x kSegIn
};

__device__ __forceinline__ void segm_range(const SegM& S, int& g, long& lo, long& hi) {
  const int w = blockIdx.x;
  g = S.wg_seg[w];
  lo = S.node_off[g] + (long)S.wg_chunk[w] * kSegChunkM;
  hi = min(lo + (long)kSegChunkM, (long)S.node_off[g + 1]);
}

__device__ __forceinline__ bool first_wg(const SegM& S, int g) {
  return (int)blockIdx.x == S.wg_off[g] && threadIdx.x == 0;
}

// K strided sums of the segment's workgroup partials (part[w * K + k], w in [w0, w1)), fixed
// order, valid in every thread
template <int K>
__device__ void seg_sums(const double* __restrict__ part, int w0, int w1, double (&out)[K]) {
  __shared__ double sh[K][kBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  strided_sums<K, kBlock>(part, w0, w1, v);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = wave_red<false>(v[k]);
    if (lane == 0) sh[k][wv] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = sh[k][0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) t += sh[k][w];
    out[k] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ void tangent3s(const double* m, const double* g, double* out) {
  const double md = m[0] * g[0] + m[1] * g[1] + m[2] * g[2];
  out[0] = g[0] - md * m[0];
  out[1] = g[1] - md * m[1];
  out[2] = g[2] - md * m[2];
}

// tab[g * kSegSlots + slot + k] = sum (or max) of the segment's K-strided partials; one
// workgroup per segment
template <int K, bool MAX>
__global__ __launch_bounds__(kBlock) void k_sg_reduce(SegM S, const double* __restrict__ part, int slot) {
  __shared__ double sh[K][kBlock / 64];
  const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int w0 = S.wg_off[g], w1 = S.wg_off[g + 1];
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int w = w0 + tid; w < w1; w += kBlock)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = MAX ? fmax(v[k], part[w * K + k]) : v[k] + part[w * K + k];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = wave_red<MAX>(v[k]);
    if (lane == 0) sh[k][wv] = v[k];
  }
  __syncthreads();
  if (tid == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = sh[k][0];
      for (int w = 1; w < kBlock / 64; ++w) t = MAX ? fmax(t, sh[k][w]) : t + sh[k][w];
      S.tab[(long)g * kSegSlots + slot + k] = t;
    }
}

// g_tan = T(m, g invMr), g_ext = T(m, g) on iterating segments; partial max |g_tan|
__global__ __launch_bounds__(kBlock) void k_sg_tangents(SegM S, const double* __restrict__ m,
                                                        const double* __restrict__ g,
                                                        const double* __restrict__ invMr, double* __restrict__ g_tan,
                                                        double* __restrict__ g_ext, double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  const bool act = S.in[sg * kSegIn + I_WANT] != 0.0;
  double v[1] = {0.0};
  if (act)
    for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
      double t[3];
      tangent3s(m + 3 * i, g + 3 * i, t);
      g_ext[3 * i] = t[0];
      g_ext[3 * i + 1] = t[1];
      g_ext[3 * i + 2] = t[2];
      const double s = invMr[i];
      const double gs[3] = {g[3 * i] * s, g[3 * i + 1] * s, g[3 * i + 2] * s};
      tangent3s(m + 3 * i, gs, t);
      g_tan[3 * i] = t[0];
      g_tan[3 * i + 1] = t[1];
      g_tan[3 * i + 2] = t[2];
      v[0] = fmax(v[0], fmax(fabs(t[0]), fmax(fabs(t[1]), fabs(t[2]))));
    }
  block_partial<1, kBlock, true>(v, part);
}

// ------------------------------------------------------------------ Steihaug-PCG per segment
// The loop's r, z, p, q and C p are stored as TV: FP64, or FP32 in the mixed-precision forms
// (products in FP64 from the stored values; with TV = double exactly the FP64 loop).
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_sg_pc_init(SegM S, const double* __restrict__ g_ext,
                                                       const double* __restrict__ Mp, double* __restrict__ y,
                                                       TV* __restrict__ r, TV* __restrict__ z,
                                                       TV* __restrict__ p, TV* __restrict__ cv,
                                                       double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  const bool act = S.in[sg * kSegIn + I_WANT] != 0.0;
  double v[1] = {0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
#pragma unroll
    for (int k = 0; k < 3; ++k) cv[3 * i + k] = (TV)0;  // every segment: k_local adds into cv
    if (!act) continue;
    const double s = Mp[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double ri = (double)(TV)g_ext[3 * i + k];
      const double zi = (double)(TV)(ri * s);
      y[3 * i + k] = 0.0;
      r[3 * i + k] = (TV)ri;
      z[3 * i + k] = (TV)zi;
      p[3 * i + k] = (TV)zi;
      if constexpr (sizeof(TV) == 8)
        v[0] += ri * ri * s;  // the FP64 loop's own order
      else
        v[0] += ri * zi;  // r.z of the stored vectors
    }
  }
  block_partial<1>(v, part);
}

__device__ __forceinline__ double sg_pc_active(const double* bank, int max_iter) {
  const double rho = bank[SB_RHO];
  const bool act = (bank[SB_IT] < (double)max_iter) && (rho > bank[SB_TGT]) && (rho > 1e-25) && (rho < 1e20) &&
                   (bank[SB_DONE] == 0.0);
  return act ? 1.0 : 0.0;
}

// bank 0 of every segment (one workgroup per segment): rho, forcing tolerance, Q, iteration,
// active = iterating && the single-system test
__global__ __launch_bounds__(kBlock) void k_sg_pc_state_init(SegM S, const double* __restrict__ part, int pc_auto,
                                                             double eta, double alpha_f, int max_iter) {
  const int g = blockIdx.x;
  double rho[1];
  seg_sums<1>(part, S.wg_off[g], S.wg_off[g + 1], rho);
  if (threadIdx.x != 0) return;
  double* t = S.tab + (long)g * kSegSlots;
  double* bk = t + seg_bank(0);
  const double tol = pc_auto ? fmin(eta, pow(t[T_GN], alpha_f)) : 0.0;
  bk[SB_RHO] = rho[0];
  bk[SB_TGT] = tol * tol * rho[0];
  bk[SB_Q] = 0.0;
  bk[SB_IT] = 0.0;
  bk[SB_DONE] = 0.0;
  bk[SB_ACT] = S.in[g * kSegIn + I_WANT] != 0.0 ? sg_pc_active(bk, max_iter) : 0.0;
}

// q = P p = Cv - (m.Cv) m - (m.g) p + reg p; partials of p.q
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_sg_pc_ap(SegM S, int bank_slot, const double* __restrict__ m,
                                                     const double* __restrict__ g, const TV* __restrict__ p,
                                                     const TV* __restrict__ cv, TV* __restrict__ q, double reg,
                                                     double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  if (S.tab[(long)sg * kSegSlots + bank_slot + SB_ACT] == 0.0) return;
  double v[1] = {0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double* mi = m + 3 * i;
    const double c0 = (double)cv[3 * i], c1 = (double)cv[3 * i + 1], c2 = (double)cv[3 * i + 2];
    const double ci[3] = {c0, c1, c2};
    const double mc = mi[0] * c0 + mi[1] * c1 + mi[2] * c2;
    const double mg = mi[0] * g[3 * i] + mi[1] * g[3 * i + 1] + mi[2] * g[3 * i + 2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double pk = (double)p[3 * i + k];
      const double qk = (double)(TV)(ci[k] - mc * mi[k] - mg * pk + reg * pk);
      q[3 * i + k] = (TV)qk;
      v[0] += pk * qk;
    }
  }
  block_partial<1>(v, part);
}

// p.q of the segment -> alpha, dq, negative curvature / stagnation; y += alpha p; r -= alpha q
// (unless done now); z = Mp r; partials of r.z
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_sg_pc_yr(SegM S, int bank_slot, const double* __restrict__ part_a,
                                                     double nu, double* __restrict__ y, TV* __restrict__ r,
                                                     TV* __restrict__ z, const TV* __restrict__ p,
                                                     const TV* __restrict__ q, const double* __restrict__ Mp,
                                                     double* __restrict__ part_b) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double* t = S.tab + (long)sg * kSegSlots;
  const double* bank = t + bank_slot;
  if (bank[SB_ACT] == 0.0) return;
  double pq[1];
  seg_sums<1>(part_a, S.wg_off[sg], S.wg_off[sg + 1], pq);
  const double pAp = pq[0];
  const double rho = bank[SB_RHO];
  const double alpha = rho / (pAp + 1e-30);
  const double dq = 0.5 * alpha * rho;
  const bool stag = (bank[SB_IT] > 0.0) && (dq <= nu * bank[SB_Q]);
  const bool dn = pAp <= 0.0 || stag;
  if (first_wg(S, sg)) {
    t[T_PAP] = pAp;
    t[T_ALPHA] = alpha;
    t[T_DQ] = dq;
    t[T_DN] = dn ? 1.0 : 0.0;
  }
  double v[1] = {0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double s = Mp[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double rk = (double)r[3 * i + k];
      if (!dn) {
        y[3 * i + k] += alpha * (double)p[3 * i + k];
        rk = (double)(TV)(rk - alpha * (double)q[3 * i + k]);
        r[3 * i + k] = (TV)rk;
      }
      const double zk = (double)(TV)(rk * s);
      z[3 * i + k] = (TV)zk;
      v[0] += rk * zk;
    }
  }
  block_partial<1>(v, part_b);
}

// r.z of the segment; p = z + beta p (unless done now); cv = 0; the segment's first workgroup
// writes the next bank (a copy of this one once the segment's loop has stopped)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_sg_pc_p(SegM S, int bank_slot, int next_slot,
                                                    const double* __restrict__ part_b, const TV* __restrict__ z,
                                                    TV* __restrict__ p, TV* __restrict__ cv, int max_iter) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double* t = S.tab + (long)sg * kSegSlots;
  const double* bank = t + bank_slot;
  double* next = t + next_slot;
  if (bank[SB_ACT] == 0.0) {
    if ((int)blockIdx.x == S.wg_off[sg] && threadIdx.x < 6) next[threadIdx.x] = bank[threadIdx.x];
    return;
  }
  double rz[1];
  seg_sums<1>(part_b, S.wg_off[sg], S.wg_off[sg + 1], rz);
  const double rhon = rz[0];
  const bool upd = t[T_DN] == 0.0;
  const double beta = rhon / (bank[SB_RHO] + 1e-30);
  for (long i = lo + threadIdx.x; i < hi; i += kBlock)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (upd) p[3 * i + k] = (TV)((double)z[3 * i + k] + beta * (double)p[3 * i + k]);
      cv[3 * i + k] = (TV)0;
    }
  if (first_wg(S, sg)) {
    next[SB_RHO] = rhon;
    next[SB_Q] = bank[SB_Q] + t[T_DQ];
    next[SB_IT] = bank[SB_IT] + 1.0;
    next[SB_DONE] = t[T_DN];
    next[SB_TGT] = bank[SB_TGT];
    next[SB_ACT] = sg_pc_active(next, max_iter);
  }
}

// partials of y.y, zf.zf (zf = Mp g_ext), y.g_ext
__global__ __launch_bounds__(kBlock) void k_sg_pc_norms(SegM S, const double* __restrict__ y,
                                                        const double* __restrict__ g, const double* __restrict__ Mp,
                                                        double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double v[3] = {0.0, 0.0, 0.0};
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double s = Mp[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double yk = y[3 * i + k], gk = g[3 * i + k];
      v[0] += yk * yk;
      v[1] += gk * s * gk * s;
      v[2] += yk * gk;
    }
  }
  block_partial<3>(v, part);
}

// clip |y| <= 10 |zf|, fall back to zf when y is no descent direction (iterating segments);
// partial max |y|; the segment's first workgroup stores the norms and the inner iterations
__global__ __launch_bounds__(kBlock) void k_sg_pc_finish(SegM S, int bank_slot, const double* __restrict__ part,
                                                         double* __restrict__ y, const double* __restrict__ g,
                                                         const double* __restrict__ Mp, double* __restrict__ part_max) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double* t = S.tab + (long)sg * kSegSlots;
  const bool act = S.in[sg * kSegIn + I_WANT] != 0.0;
  double tot[3];
  seg_sums<3>(part, S.wg_off[sg], S.wg_off[sg + 1], tot);
  const double yy = tot[0], zz = tot[1], yg = tot[2];
  const double yn = sqrt(yy), zn = sqrt(zz);
  const double scale = (yn > 10.0 * zn) ? (10.0 * zn / (yn + 1e-30)) : 1.0;
  const bool fallback = !(scale * yg > 1e-12);
  double v[1] = {0.0};
  if (act)
    for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
      const double s = Mp[i];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double yk = fallback ? g[3 * i + k] * s : y[3 * i + k] * scale;
        y[3 * i + k] = yk;
        v[0] = fmax(v[0], fabs(yk));
      }
    }
  block_partial<1, kBlock, true>(v, part_max);
  if (first_wg(S, sg)) {
    t[T_YY] = yy;
    t[T_ZZ] = zz;
    t[T_YG] = yg;
    t[T_PCIT] = t[bank_slot + SB_IT];
  }
}

// ----------------------------------------------------------------- direction per segment
// partials of y.(g_ext - g_prev) and d_prev.(g_ext - g_prev)
__global__ __launch_bounds__(kBlock) void k_sg_beta_dots(SegM S, const double* __restrict__ y,
                                                         const double* __restrict__ ge, const double* __restrict__ gp,
                                                         const double* __restrict__ dp, double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double v[2] = {0.0, 0.0};
  for (long i = 3 * lo + threadIdx.x; i < 3 * hi; i += kBlock) {
    const double dg = ge[i] - gp[i];
    v[0] += y[i] * dg;
    v[1] += dp[i] * dg;
  }
  block_partial<2>(v, part);
}

// d = -y + beta T(m, d_prev) (beta_HS >= 0, 0 on restart) on iterating segments; partials of d.g_ext
__global__ __launch_bounds__(kBlock) void k_sg_direction(SegM S, const double* __restrict__ m,
                                                         const double* __restrict__ y, const double* __restrict__ dp,
                                                         const double* __restrict__ ge, double* __restrict__ d,
                                                         const double* __restrict__ part_beta,
                                                         double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  const double* in = S.in + sg * kSegIn;
  double nd[2];
  seg_sums<2>(part_beta, S.wg_off[sg], S.wg_off[sg + 1], nd);
  const double beta = in[I_RESTART] != 0.0 ? 0.0 : fmax(0.0, nd[0] / (nd[1] + 1e-30));
  double v[1] = {0.0};
  if (in[I_WANT] != 0.0)
    for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
      double t[3];
      tangent3s(m + 3 * i, dp + 3 * i, t);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double dk = -y[3 * i + k] + beta * t[k];
        d[3 * i + k] = dk;
        v[0] += dk * ge[3 * i + k];
      }
    }
  block_partial<1>(v, part);
}

// d = -y where d.g_ext > 0; H = m x d; partials of pg = g_raw . d
__global__ __launch_bounds__(kBlock) void k_sg_direction_fix(SegM S, const double* __restrict__ m,
                                                             const double* __restrict__ alt,
                                                             const double* __restrict__ g, double* __restrict__ d,
                                                             double* __restrict__ H, const double* __restrict__ part_dg,
                                                             double* __restrict__ part) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double dg[1];
  seg_sums<1>(part_dg, S.wg_off[sg], S.wg_off[sg + 1], dg);
  const bool flip = dg[0] > 0.0;
  double v[1] = {0.0};
  if (S.in[sg * kSegIn + I_WANT] != 0.0)
    for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
      double dd[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        dd[k] = flip ? -alt[3 * i + k] : d[3 * i + k];
        d[3 * i + k] = dd[k];
        v[0] += g[3 * i + k] * dd[k];
      }
      const double* mi = m + 3 * i;
      H[3 * i + 0] = mi[1] * dd[2] - mi[2] * dd[1];
      H[3 * i + 1] = mi[2] * dd[0] - mi[0] * dd[2];
      H[3 * i + 2] = mi[0] * dd[1] - mi[1] * dd[0];
    }
  block_partial<1>(v, part);
}

// ------------------------------------------------------------------ line search / accept
// m_t = cayley(m, H, s_g) on segments taking a trial (src/minimizers.py:71-79)
__global__ __launch_bounds__(kBlock) void k_sg_cayley(SegM S, const double* __restrict__ m,
                                                      const double* __restrict__ H, double* __restrict__ out) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  const double* in = S.in + sg * kSegIn;
  if (in[I_TRIAL] == 0.0) return;
  const double s = in[I_STEP];
  for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
    const double k0 = 0.5 * s * H[3 * i], k1 = 0.5 * s * H[3 * i + 1], k2 = 0.5 * s * H[3 * i + 2];
    const double m0 = m[3 * i], m1 = m[3 * i + 1], m2 = m[3 * i + 2];
    const double kk = k0 * k0 + k1 * k1 + k2 * k2;
    const double km0 = k1 * m2 - k2 * m1, km1 = k2 * m0 - k0 * m2, km2 = k0 * m1 - k1 * m0;
    const double kd = k0 * m0 + k1 * m1 + k2 * m2;
    const double inv = 1.0 / (1.0 + kk);
    const double a = ((1.0 - kk) * m0 + 2.0 * km0 + 2.0 * kd * k0) * inv;
    const double b = ((1.0 - kk) * m1 + 2.0 * km1 + 2.0 * kd * k1) * inv;
    const double c = ((1.0 - kk) * m2 + 2.0 * km2 + 2.0 * kd * k2) * inv;
    const double n = sqrt(a * a + b * b + c * c);
    out[3 * i] = a / n;
    out[3 * i + 1] = b / n;
    out[3 * i + 2] = c / n;
  }
}

// accepted segments: partial max |m_t - m|, then m = m_t, U = U_t, g_raw = g_t
__global__ __launch_bounds__(kBlock) void k_sg_accept(SegM S, double* __restrict__ m, const double* __restrict__ m_t,
                                                      double* __restrict__ U, const double* __restrict__ U_t,
                                                      double* __restrict__ g, const double* __restrict__ g_t,
                                                      double* __restrict__ part_max) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  double v[1] = {0.0};
  if (S.in[sg * kSegIn + I_ACCEPT] != 0.0)
    for (long i = lo + threadIdx.x; i < hi; i += kBlock) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double a = m_t[3 * i + k];
        v[0] = fmax(v[0], fabs(a - m[3 * i + k]));
        m[3 * i + k] = a;
        g[3 * i + k] = g_t[3 * i + k];
      }
      U[i] = U_t[i];
    }
  block_partial<1, kBlock, true>(v, part_max);
}

// x = 0 on the nodes (3 values each) of the launched segments
__global__ __launch_bounds__(kBlock) void k_sg_zero3(SegM S, double* __restrict__ x) {
  int sg;
  long lo, hi;
  segm_range(S, sg, lo, hi);
  for (long i = 3 * lo + threadIdx.x; i < 3 * hi; i += kBlock) x[i] = 0.0;
}

// m /= |m|
__global__ void k_sg_normalize(double* __restrict__ m, long N) {
  for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < N; i += (long)gridDim.x * kBlock) {
    const double a = m[3 * i], b = m[3 * i + 1], c = m[3 * i + 2];
    const double n = sqrt(a * a + b * b + c * c);
    m[3 * i] = a / n;
    m[3 * i + 1] = b / n;
    m[3 * i + 2] = c / n;
  }
}

// x = -y over 3N (d_prev = -g_ext at the start)
__global__ void k_sg_neg(double* __restrict__ x, const double* __restrict__ y, long n) {
  for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) x[i] = -y[i];
}

// ----------------------------------------------------------------------- host driver
struct SegLS {
  int phase = 0;  // 1 expansion, 2 contraction, 3 finished
  double s = 1.0, s_min = 0.0, s_min_exp = 0.0, d_v = 0.0, E_v = 0.0, pg = 0.0, E0 = 0.0, tau = 0.0, E = 0.0;
  int it_exp = 0, it_con = 0, demag = 0;
  bool moved = false;
};

static int seg_min_alloc(Ctx& c) {
  MM_TRY(ensure_seg(c));
  if (c.sm_tab) return 0;
  const size_t W = (size_t)c.seg_nwg;
  for (double** v : {&c.sm_pa, &c.sm_pb}) {
    MM_HIP(hipMalloc((void**)v, 4 * W * sizeof(double)));
    c.allocs.push_back(*v);
  }
  MM_HIP(hipMalloc((void**)&c.sm_tab, (size_t)c.nseg * kSegSlots * sizeof(double)));
  c.allocs.push_back(c.sm_tab);
  // on the context's stream (non-blocking: a null-stream fill would not be ordered with its kernels)
  MM_HIP(hipMemsetAsync(c.sm_tab, 0, (size_t)c.nseg * kSegSlots * sizeof(double), c.stream));
  MM_HIP(hipMalloc((void**)&c.sm_in, (size_t)c.nseg * kSegIn * sizeof(double)));
  c.allocs.push_back(c.sm_in);
  return 0;
}

int minimize_segmented(Ctx& c, const MinParams& p, const double B[3], MinInfo* info) {
  MM_CHECK(p.method == 0, "the segmented minimiser runs pcohen_hs (method 0)");
  MM_CHECK(p.ls_max_evals >= 1, "the segmented minimiser needs ls_max_evals >= 1");
  MM_TRY(alloc_min_state(c));
  MM_TRY(seg_min_alloc(c));
  const auto t0 = std::chrono::steady_clock::now();
  const int G = c.nseg;
  int W = 0;  // workgroups of the current segment subset
  const long N = c.N, n3 = 3 * N;
  hipStream_t s = c.stream;
  SegM S;
  S.node_off = c.d_seg_node_off;
  S.wg_off = c.wg_min.off;
  S.wg_seg = c.wg_min.seg;
  S.wg_chunk = c.wg_min.chunk;
  S.tab = c.sm_tab;
  S.in = c.sm_in;
  std::vector<double> in((size_t)G * kSegIn, 0.0), tab((size_t)G * kSegSlots, 0.0);
  auto put_in = [&]() -> int {
    MM_HIP(hipMemcpyAsync(c.sm_in, in.data(), in.size() * sizeof(double), hipMemcpyHostToDevice, s));
    return 0;
  };
  auto get_tab = [&]() -> int {
    MM_HIP(hipMemcpyAsync(tab.data(), c.sm_tab, tab.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MM_HIP(hipStreamSynchronize(s));
    return 0;
  };
  auto T = [&](int g, int slot) { return tab[(size_t)g * kSegSlots + slot]; };
  // restrict the vector kernels and the element kernels to the segments in mask
  auto use = [&](const std::vector<int>& mask) -> int {
    {
      SegBatch batch(c);  // workgroup tables and block lists in one copy
      MM_TRY(build_wg_tables(c, mask.data(), c.wg_min));
      MM_TRY(set_block_lists(c, mask.data()));
    }
    MM_TRY(seg_flush(c));
    W = c.wg_min.n;
    return 0;
  };
  // energy and gradient of the current subset's segments; the gradient of the other segments
  // is left as it is (a segment whose line search ended earlier keeps its accepted trial)
  auto energy = [&](const double* m, const double* U, double* g) -> int {
    k_sg_zero3<<<W, kBlock, 0, s>>>(S, g);
    MM_TRY(launch_energy(c, m, U, B, g, c.dscal + 16, s));
    c.n_energy++;
    return launch_seg_block_sums(c, c.partial, 1, true, c.sm_tab, kSegSlots, T_E, s);
  };

  std::vector<int> it(G, 0), evals(G, 1), preco(G, 0), demag(G, 0), conv(G, 0), want(G, 1), trial(G, 0);
  std::vector<double> E(G, 0.0), gnorm(G, 0.0), tolv(G, p.cg_tol), rr(G, 0.0);
  std::vector<int> its(G, 0);

  // solve_and_minimize per segment (src/minimizers.py:3408-3435)
  for (int g = 0; g < G; ++g) in[(size_t)g * kSegIn + I_WANT] = 1.0;
  MM_TRY(put_in());
  MM_TRY(use(want));
  const int vgrid = (int)std::min<long>((n3 + kBlock - 1) / kBlock, 8192);
  k_sg_normalize<<<vgrid, kBlock, 0, s>>>(c.m, N);
  // every segment's trial buffers hold a valid state from the start: the coarsest AMG level
  // multiplies whole vectors by the dense block-diagonal inverse, so a non-finite value in
  // one segment would reach all of them
  MM_TRY(vcopy(c.m_t, c.m, n3, s));
  MM_TRY(vcopy(c.U_t, c.U, N, s));
  MM_TRY(solve_U_seg_dev(c, c.m, c.U, tolv.data(), nullptr, p.cg_maxiter, its.data(), rr.data()));
  for (int g = 0; g < G; ++g) demag[g] = its[g];
  MM_TRY(use(want));
  MM_TRY(energy(c.m, c.U, c.g_raw));
  k_sg_tangents<<<W, kBlock, 0, s>>>(S, c.m, c.g_raw, c.inv_M_rel, c.g_tan, c.g_ext, c.sm_pa);
  k_sg_reduce<1, true><<<G, kBlock, 0, s>>>(S, c.sm_pa, T_GN);
  MM_TRY(get_tab());
  for (int g = 0; g < G; ++g) {
    E[g] = T(g, T_E);
    gnorm[g] = T(g, T_GN);
  }
  // PCGState init: g_prev = g_ext, d_prev = -g_ext
  MM_TRY(vcopy(c.g_prev, c.g_ext, n3, s));
  k_sg_neg<<<vgrid, kBlock, 0, s>>>(c.d_prev, c.g_ext, n3);

  std::vector<int> Lp(G);
  for (int g = 0; g < G; ++g) Lp[g] = p.L > 0 ? p.L : c.h_seg_node_off[g + 1] - c.h_seg_node_off[g];
  for (int g = 0; g < G; ++g) want[g] = p.max_iter > 0;
  int live = 0;
  for (int g = 0; g < G; ++g) live += want[g];

  const double phi_tol = p.phi_tol;
  while (live > 0) {
    for (int g = 0; g < G; ++g) {
      double* q = &in[(size_t)g * kSegIn];
      q[I_WANT] = want[g];
      q[I_RESTART] = (it[g] % Lp[g]) == 0;
      q[I_TRIAL] = 0.0;
      q[I_ACCEPT] = 0.0;
    }
    MM_TRY(put_in());
    MM_TRY(use(want));
    k_sg_tangents<<<W, kBlock, 0, s>>>(S, c.m, c.g_raw, c.inv_M_rel, c.g_tan, c.g_ext, c.sm_pa);
    k_sg_reduce<1, true><<<G, kBlock, 0, s>>>(S, c.sm_pa, T_GN);
    // Steihaug-PCG solve of P y = g_ext per segment, on FP32 vectors in the mixed-precision forms
    auto pc_loop = [&](auto* r3, auto* z3, auto* p3, auto* q3, auto* cv3) -> int {
      using TV = std::remove_pointer_t<decltype(r3)>;
      k_sg_pc_init<TV><<<W, kBlock, 0, s>>>(S, c.g_ext, c.inv_M_prec, c.y, r3, z3, p3, cv3, c.sm_pb);
      k_sg_pc_state_init<<<G, kBlock, 0, s>>>(S, c.sm_pb, p.pc_auto, p.pc_force_eta, p.pc_force_alpha,
                                              p.pc_iters);
      for (int k = 0; k < p.pc_iters; ++k) {
        MM_TRY(launch_local(c, p3, cv3, nullptr, s));
        c.n_local++;
        k_sg_pc_ap<TV><<<W, kBlock, 0, s>>>(S, seg_bank(k), c.m, c.g_raw, p3, cv3, q3, p.pc_reg, c.sm_pa);
        k_sg_pc_yr<TV><<<W, kBlock, 0, s>>>(S, seg_bank(k), c.sm_pa, p.pc_stagnation_nu, c.y, r3, z3, p3, q3,
                                             c.inv_M_prec, c.sm_pb);
        k_sg_pc_p<TV><<<W, kBlock, 0, s>>>(S, seg_bank(k), seg_bank(k + 1), c.sm_pb, z3, p3, cv3, p.pc_iters);
      }
      return 0;
    };
    if (c.precision >= 1 && c.r3f != nullptr)
      MM_TRY(pc_loop(c.r3f, c.z3f, c.p3f, c.q3f, c.cv3f));
    else
      MM_TRY(pc_loop(c.r3, c.z3, c.p3, c.q3, c.cv3));
    k_sg_pc_norms<<<W, kBlock, 0, s>>>(S, c.y, c.g_ext, c.inv_M_prec, c.sm_pa);
    k_sg_pc_finish<<<W, kBlock, 0, s>>>(S, seg_bank(p.pc_iters), c.sm_pa, c.y, c.g_ext, c.inv_M_prec, c.sm_pb);
    k_sg_reduce<1, true><<<G, kBlock, 0, s>>>(S, c.sm_pb, T_GS);
    // direction
    k_sg_beta_dots<<<W, kBlock, 0, s>>>(S, c.y, c.g_ext, c.g_prev, c.d_prev, c.sm_pa);
    k_sg_direction<<<W, kBlock, 0, s>>>(S, c.m, c.y, c.d_prev, c.g_ext, c.d, c.sm_pa, c.sm_pb);
    k_sg_direction_fix<<<W, kBlock, 0, s>>>(S, c.m, c.y, c.g_raw, c.d, c.H, c.sm_pb, c.sm_pa);
    k_sg_reduce<1, false><<<G, kBlock, 0, s>>>(S, c.sm_pa, T_PG);
    MM_TRY(get_tab());

    // line search, one trial per searching segment per round (make_armijo_ls_v2)
    std::vector<SegLS> ls(G);
    int searching = 0;
    for (int g = 0; g < G; ++g) {
      SegLS& L = ls[g];
      L.E0 = E[g];
      L.E = E[g];
      L.pg = T(g, T_PG);
      if (want[g] && L.pg < 0.0) {
        L.phase = 1;
        L.s = 1.0;
        ++searching;
      } else {
        L.phase = 3;
      }
    }
    MM_TRY(vcopy(c.U_t, c.U, N, s));
    const int maxe = p.ls_max_evals;
    while (searching > 0) {
      for (int g = 0; g < G; ++g) {
        const bool tr = ls[g].phase == 1 || ls[g].phase == 2;
        trial[g] = tr;
        in[(size_t)g * kSegIn + I_TRIAL] = tr;
        in[(size_t)g * kSegIn + I_STEP] = tr ? ls[g].s : 0.0;
        tolv[g] = phi_tol;
      }
      MM_TRY(put_in());
      MM_TRY(use(trial));
      k_sg_cayley<<<W, kBlock, 0, s>>>(S, c.m, c.H, c.m_t);
      MM_TRY(solve_U_seg_dev(c, c.m_t, c.U_t, tolv.data(), trial.data(), p.cg_maxiter, its.data(), rr.data()));
      MM_TRY(use(trial));  // the solve leaves every block enabled
      MM_TRY(energy(c.m_t, c.U_t, c.g_t));
      MM_TRY(get_tab());
      for (int g = 0; g < G; ++g) {
        if (!trial[g]) continue;
        SegLS& L = ls[g];
        L.demag += its[g];
        double En = T(g, T_E);
        if (!std::isfinite(En)) En = 1e20;
        const double dn = (En - L.E0) / (L.s * L.pg + 1e-30);
        bool finish = false;
        if (L.phase == 1) {
          const bool stop = std::fabs(1.0 - dn) >= p.ls_eta2 || dn < 0.0;
          const double s_next = stop ? L.s : p.ls_C * L.s;
          if (!stop) L.s_min = L.s;
          L.s = s_next;
          L.d_v = dn;
          L.E_v = En;
          ++L.it_exp;
          if (stop || L.it_exp >= maxe) {
            L.s_min_exp = L.s_min;
            const bool done = L.d_v >= p.ls_eta1 && L.d_v < 1e10;
            if (done || maxe <= 0) {
              finish = true;
            } else {
              L.phase = 2;
              L.s = L.s_min_exp + p.ls_c * (L.s - L.s_min_exp);
            }
          }
        } else {
          L.d_v = dn;
          L.E_v = En;
          ++L.it_con;
          const bool done = L.d_v >= p.ls_eta1 && L.d_v < 1e10;
          if (done || L.it_con >= maxe)
            finish = true;
          else
            L.s = L.s_min_exp + p.ls_c * (L.s - L.s_min_exp);
        }
        if (finish) {
          L.phase = 3;
          --searching;
          if (L.d_v >= 0.0) {
            L.tau = L.s;
            L.E = L.E_v;
            L.moved = true;
          }
        }
      }
    }

    // accept, convergence (src/minimizers.py:160-172), bookkeeping
    for (int g = 0; g < G; ++g) {
      in[(size_t)g * kSegIn + I_TRIAL] = 0.0;
      in[(size_t)g * kSegIn + I_ACCEPT] = ls[g].moved;
    }
    MM_TRY(put_in());
    MM_TRY(use(want));
    k_sg_accept<<<W, kBlock, 0, s>>>(S, c.m, c.m_t, c.U, c.U_t, c.g_raw, c.g_t, c.sm_pa);
    k_sg_reduce<1, true><<<G, kBlock, 0, s>>>(S, c.sm_pa, T_DM);
    MM_TRY(get_tab());
    std::swap(c.g_prev, c.g_ext);
    std::swap(c.d_prev, c.d);
    live = 0;
    for (int g = 0; g < G; ++g) {
      if (!want[g]) continue;
      const SegLS& L = ls[g];
      const double dm = L.moved ? T(g, T_DM) : 0.0;
      const double gs = T(g, T_GS);
      const double E_new = L.E;
      if (it[g] > 0) {
        const bool u1 = (E[g] - E_new) < p.tau_f * (1.0 + std::fabs(E_new));
        const bool u2 = dm < std::sqrt(p.tau_f) * 2.0;
        const bool u3 = gs <= std::pow(p.tau_f, 1.0 / 3.0) * (1.0 + std::fabs(E_new));
        const bool u4 = gs < p.eps_a;
        conv[g] = (u1 && u2 && u3) || u4;
      }
      gnorm[g] = gs;
      E[g] = E_new;
      evals[g] += L.it_exp + L.it_con;
      preco[g] += (int)T(g, T_PCIT);
      demag[g] += L.demag;
      ++it[g];
      want[g] = !conv[g] && it[g] < p.max_iter;
      live += want[g];
    }
  }
  MM_TRY(set_block_lists(c, nullptr));
  MM_HIP(hipStreamSynchronize(s));
  MM_HIP(hipGetLastError());
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int g = 0; g < G; ++g) {
    info[g].iters = it[g];
    info[g].evals = evals[g];
    info[g].demag_iters = demag[g];
    info[g].preco_iters = preco[g];
    info[g].converged = conv[g];
    info[g].E = E[g];
    info[g].gnorm = gnorm[g];
    info[g].time = el;
  }
  return 0;
}

// per-segment J_par along h: out[g] = sum_e Ve Js m.h / V_mag of segment g
int j_parallel_segmented(Ctx& c, const double* m, const double h[3], double* out) {
  MM_TRY(seg_min_alloc(c));
  MM_TRY(launch_jpar(c, m, h, c.dscal + 20, c.stream));
  MM_TRY(launch_seg_block_sums(c, c.partial, 2, false, c.sm_pa, 2, 0, c.stream));
  std::vector<double> v(2 * (size_t)c.nseg);
  MM_HIP(hipMemcpyAsync(v.data(), c.sm_pa, v.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  for (int g = 0; g < c.nseg; ++g) out[g] = v[2 * g] / (v[2 * g + 1] + 1e-30);
  return 0;
}

}  // namespace mm
