// Field-step minimiser on the unit sphere, device resident: projected L-BFGS
// (src/minimizers.py:1216-1381) and preconditioned Cohen-HS conjugate gradients
// (src/minimizers.py:950-1073) with the curvilinear Armijo line search
// (src/minimizers.py:265-385), the Steihaug-PCG preconditioner solve
// (src/minimizers.py:393-492) and the Gill-Murray convergence test (src/minimizers.py:160-172).
//
// Vectors (N x 3 interleaved) never leave the GPU. Reductions are per-workgroup partials
// (no fences): the next kernel sums the ones it needs, and the host reads back only what
// drives control flow: the slope pg and the convergence quantities once per iteration, the
// trial energy once per line-search evaluation. Loop state and step lengths live in the
// device scalar block.
#include <chrono>
#include <cmath>
#include <type_traits>
#include <utility>

#include "mm_device.h"
#include "mm_solver.h"

namespace mm {

int alloc_min_state(Ctx& c);

enum Slot {
  S_E = 16,
  // max |g_tan|, max |y|, the beta dots, d.g_ext, g_raw.d and max |m_new - m| are partial banks
  // (RedBank below)
  S_PCIT = 31, // inner iterations (copy for the host)
  S_RHO = 32, S_Q = 33, S_IT = 34, S_DONE = 35, S_ACT = 36, S_ALPHA = 37, S_TGT = 38, S_DQ = 39, S_DN = 40,
  S_PAP = 41, S_RHON = 42, S_YY = 43, S_ZZ = 44, S_YG = 45,
  S_ALPHAS = 64,  // + history slot
  S_RHOS = 96,    // + history slot
};

static inline int vg(long n) {
  long g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  return (int)(g < 1 ? 1 : g);
}

#define NODE_LOOP(i, n) for (long i = blockIdx.x * (long)kBlock + threadIdx.x; i < (n); i += (long)gridDim.x * kBlock)

// operands of one node for the chunked loops (grid_chunks) of the node kernels: up to four 3-vectors and a
// scalar
struct Node4x3 {
  double3 a, b, c, d;
  double s;
};
template <typename TV>
__device__ __forceinline__ double3 load3(const TV* __restrict__ v, long i) {
  return double3{(double)v[3 * i], (double)v[3 * i + 1], (double)v[3 * i + 2]};
}

__device__ __forceinline__ void tangent3(const double* m, const double* g, double s, double* out) {
  const double md = m[0] * g[0] + m[1] * g[1] + m[2] * g[2];
  out[0] = s * (g[0] - md * m[0]);
  out[1] = s * (g[1] - md * m[1]);
  out[2] = s * (g[2] - md * m[2]);
}

// m /= |m|
__global__ void k_normalize(double* __restrict__ m, long N) {
  NODE_LOOP(i, N) {
    const double a = m[3 * i], b = m[3 * i + 1], c = m[3 * i + 2];
    const double n = sqrt(a * a + b * b + c * c);
    m[3 * i] = a / n;
    m[3 * i + 1] = b / n;
    m[3 * i + 2] = c / n;
  }
}

// g_tan = T(m, g invMr), g_ext = T(m, g); per-workgroup max |g_tan| to part (src/minimizers.py:94-96,
// 963-966)
__global__ __launch_bounds__(kBlock) void k_tangents(const double* __restrict__ m, const double* __restrict__ g,
                                                     const double* __restrict__ invMr, double* __restrict__ g_tan,
                                                     double* __restrict__ g_ext, long N, double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks<2>(
      N, [&](long i) { return Node4x3{load3(m, i), load3(g, i), {}, {}, invMr[i]}; },
      [&](long i, const Node4x3& u) {
        const double mi[3] = {u.a.x, u.a.y, u.a.z}, gi[3] = {u.b.x, u.b.y, u.b.z};
        double t[3];
        tangent3(mi, gi, 1.0, t);
        g_ext[3 * i] = t[0];
        g_ext[3 * i + 1] = t[1];
        g_ext[3 * i + 2] = t[2];
        const double s = u.s;
        double gs[3] = {gi[0] * s, gi[1] * s, gi[2] * s};
        tangent3(mi, gs, 1.0, t);
        g_tan[3 * i] = t[0];
        g_tan[3 * i + 1] = t[1];
        g_tan[3 * i + 2] = t[2];
        v[0] = fmax(v[0], fmax(fabs(t[0]), fmax(fabs(t[1]), fabs(t[2]))));
      });
  block_partial<1, kBlock, true>(v, part);
}

// per-workgroup max |a| to part
__global__ __launch_bounds__(kBlock) void k_maxabs_part(const double* __restrict__ a, long n,
                                                        double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return a[i]; }, [&](long, double u) { v[0] = fmax(v[0], fabs(u)); });
  block_partial<1, kBlock, true>(v, part);
}

// ------------------------------------------------------------- Steihaug-PCG (solve_Py_g)
// The loop state (rho, Q, iteration, done, active, target) lives in two banks of device
// slots: iteration k reads bank k % 2 and its last kernel writes bank (k + 1) % 2, so no
// kernel overwrites a slot that other workgroups of the same launch still read. Dot
// products are per-workgroup partials summed by the consuming kernel (fixed order, no
// fences): p.q in part_a, r.z in part_b.
enum PcBank { B_RHO = 0, B_Q = 1, B_IT = 2, B_DONE = 3, B_ACT = 4, B_TGT = 5 };
constexpr int kPcBank0 = 224;  // banks at slots 224..229 and 232..237
__host__ __device__ constexpr int pc_bank(int k) { return kPcBank0 + 8 * (k & 1); }

// The loop's vectors r, z, p, q and C p are stored as TV: FP64, or FP32 in the mixed-precision
// forms (whose Hessian action is FP32 arithmetic already); y, g_ext, m, g and the scalars stay
// FP64, and every product is formed in FP64 from the stored values, so with TV = double the
// arithmetic is exactly the FP64 loop's.

// y = 0; r = g_ext; z = Mp r; p = z; partials of rho = r.z
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pc_init(const double* __restrict__ g_ext, const double* __restrict__ Mp,
                                                    double* __restrict__ y, TV* __restrict__ r,
                                                    TV* __restrict__ z, TV* __restrict__ p,
                                                    TV* __restrict__ cv, long N, double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      N, [&](long i) { return Node4x3{load3(g_ext, i), {}, {}, {}, Mp[i]}; },
      [&](long i, const Node4x3& u) {
        const double s = u.s;
        const double rr[3] = {u.a.x, u.a.y, u.a.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double ri = (double)(TV)rr[k];
          const double zi = (double)(TV)(ri * s);
          y[3 * i + k] = 0.0;
          r[3 * i + k] = (TV)ri;
          z[3 * i + k] = (TV)zi;
          p[3 * i + k] = (TV)zi;
          cv[3 * i + k] = (TV)0;
          if constexpr (sizeof(TV) == 8)
            v[0] += ri * ri * s;  // the FP64 loop's own order
          else
            v[0] += ri * zi;  // r.z of the stored vectors
        }
      });
  block_partial<1>(v, part);
}

__device__ __forceinline__ double pc_active(const double* bank, int max_iter) {
  const double rho = bank[B_RHO];
  const bool act = (bank[B_IT] < (double)max_iter) && (rho > bank[B_TGT]) && (rho > 1e-25) && (rho < 1e20) &&
                   (bank[B_DONE] == 0.0);
  return act ? 1.0 : 0.0;
}

// bank 0: rho = sum of the k_pc_init partials, forcing-sequence tolerance (src/minimizers.py:968-972,
// 445-450), Q = 0, iteration 0, active flag. One workgroup.
__global__ __launch_bounds__(kBlock) void k_pc_state_init(double* __restrict__ sc, const double* __restrict__ part,
                                                          int n_part, const double* __restrict__ part_gn, int n_gn,
                                                          int pc_auto, double eta, double alpha_f, int max_iter) {
  const double rho = sum_partials_all<kBlock>(part, n_part);
  const double gn = max_partials_all<kBlock>(part_gn, n_gn);  // max |g_tan| (k_tangents)
  if (threadIdx.x != 0) return;
  double* bk = sc + pc_bank(0);
  const double tol = pc_auto ? fmin(eta, pow(gn, alpha_f)) : 0.0;
  bk[B_RHO] = rho;
  bk[B_TGT] = tol * tol * rho;
  bk[B_Q] = 0.0;
  bk[B_IT] = 0.0;
  bk[B_DONE] = 0.0;
  bk[B_ACT] = pc_active(bk, max_iter);
}

// q = P p = Cv - (m.Cv) m - (m.g) p + reg p; partials of p.q
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pc_ap(const double* __restrict__ bank, const double* __restrict__ m,
                                                  const double* __restrict__ g, const TV* __restrict__ p,
                                                  const TV* __restrict__ cv, TV* __restrict__ q, double reg,
                                                  long N, double* __restrict__ part) {
  if (bank[B_ACT] == 0.0) return;
  double v[1] = {0.0};
  grid_chunks<2>(
      N, [&](long i) { return Node4x3{load3(m, i), load3(cv, i), load3(g, i), load3(p, i)}; },
      [&](long i, const Node4x3& u) {
        const double3 mi = u.a, c = u.b, gi = u.c, pi = u.d;
        const double mc = mi.x * c.x + mi.y * c.y + mi.z * c.z;
        const double mg = mi.x * gi.x + mi.y * gi.y + mi.z * gi.z;
        const double q0 = (double)(TV)(c.x - mc * mi.x - mg * pi.x + reg * pi.x);
        const double q1 = (double)(TV)(c.y - mc * mi.y - mg * pi.y + reg * pi.y);
        const double q2 = (double)(TV)(c.z - mc * mi.z - mg * pi.z + reg * pi.z);
        q[3 * i] = (TV)q0;
        q[3 * i + 1] = (TV)q1;
        q[3 * i + 2] = (TV)q2;
        v[0] += pi.x * q0;
        v[0] += pi.y * q1;
        v[0] += pi.z * q2;
      });
  block_partial<1>(v, part);
}

// p.q from the partials -> alpha, dq, negative curvature / stagnation (workgroup 0 stores them);
// y += alpha p; r -= alpha q (unless done now); z = Mp r; partials of rho_n = r.z
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pc_yr(double* __restrict__ sc, const double* __restrict__ bank,
                                                  const double* __restrict__ part_a, int n_part, double nu,
                                                  double* __restrict__ y, TV* __restrict__ r,
                                                  TV* __restrict__ z, const TV* __restrict__ p,
                                                  const TV* __restrict__ q, const double* __restrict__ Mp, long N,
                                                  double* __restrict__ part_b) {
  if (bank[B_ACT] == 0.0) return;
  const double pAp = sum_partials_all<kBlock>(part_a, n_part);
  const double rho = bank[B_RHO];
  const double alpha = rho / (pAp + 1e-30);
  const double dq = 0.5 * alpha * rho;
  const bool stag = (bank[B_IT] > 0.0) && (dq <= nu * bank[B_Q]);
  const bool dn = pAp <= 0.0 || stag;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_PAP] = pAp;
    sc[S_ALPHA] = alpha;
    sc[S_DQ] = dq;
    sc[S_DN] = dn ? 1.0 : 0.0;
  }
  double v[1] = {0.0};
  // z = Mp r and the r.z partials, after the update of y and r unless done (the branch on dn
  // stays outside the loops: inside, the compiler waits for each load)
  auto zr = [&](long i, const double (&rr)[3], double s) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double zk = (double)(TV)(rr[k] * s);
      z[3 * i + k] = (TV)zk;
      v[0] += rr[k] * zk;
    }
  };
  if (dn)
    grid_chunks<2>(
        N, [&](long i) { return Node4x3{load3(r, i), {}, {}, {}, Mp[i]}; },
        [&](long i, const Node4x3& u) {
          const double rr[3] = {u.a.x, u.a.y, u.a.z};
          zr(i, rr, u.s);
        });
  else
    grid_chunks<2>(
        N, [&](long i) { return Node4x3{load3(r, i), load3(y, i), load3(p, i), load3(q, i), Mp[i]}; },
        [&](long i, const Node4x3& u) {
          const double yy[3] = {u.b.x, u.b.y, u.b.z}, pp[3] = {u.c.x, u.c.y, u.c.z};
          const double r0[3] = {u.a.x, u.a.y, u.a.z}, qq[3] = {u.d.x, u.d.y, u.d.z};
          double rr[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            y[3 * i + k] = yy[k] + alpha * pp[k];
            rr[k] = (double)(TV)(r0[k] - alpha * qq[k]);
            r[3 * i + k] = (TV)rr[k];
          }
          zr(i, rr, u.s);
        });
  block_partial<1>(v, part_b);
}

// rho_n from the partials; p = z + beta p (unless done now); cv = 0 for the next iteration;
// workgroup 0 writes the next bank (a copy of this one when the loop is no longer active)
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_pc_p(const double* __restrict__ sc, const double* __restrict__ bank,
                                                 double* __restrict__ next, const double* __restrict__ part_b,
                                                 int n_part, const TV* __restrict__ z, TV* __restrict__ p,
                                                 TV* __restrict__ cv, int max_iter, long N) {
  if (bank[B_ACT] == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x < 6) next[threadIdx.x] = bank[threadIdx.x];
    return;
  }
  const double rhon = sum_partials_all<kBlock>(part_b, n_part);
  const bool upd = sc[S_DN] == 0.0;
  const double beta = rhon / (bank[B_RHO] + 1e-30);
  if (upd)
    grid_chunks(
        3 * N, [&](long j) { return double2{(double)z[j], (double)p[j]}; }, [&](long j, double2 u) {
          p[j] = (TV)(u.x + beta * u.y);
          cv[j] = (TV)0;
        });
  else
    grid_chunks(
        3 * N, [&](long) { return 0.0; }, [&](long j, double) { cv[j] = (TV)0; });
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    next[B_RHO] = rhon;
    next[B_Q] = bank[B_Q] + sc[S_DQ];
    next[B_IT] = bank[B_IT] + 1.0;
    next[B_DONE] = sc[S_DN];
    next[B_TGT] = bank[B_TGT];
    next[B_ACT] = pc_active(next, max_iter);
  }
}

// partials of y.y, zf.zf (zf = Mp g_ext), y.g_ext (3 per workgroup)
__global__ __launch_bounds__(kBlock) void k_pc_norms(const double* __restrict__ y, const double* __restrict__ g,
                                                     const double* __restrict__ Mp, long N,
                                                     double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  grid_chunks<2>(
      N, [&](long i) { return Node4x3{load3(y, i), load3(g, i), {}, {}, Mp[i]}; },
      [&](long, const Node4x3& u) {
        const double s = u.s;
        const double yy[3] = {u.a.x, u.a.y, u.a.z}, gg[3] = {u.b.x, u.b.y, u.b.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double yk = yy[k], gk = gg[k];
          v[0] += yk * yk;
          v[1] += gk * s * gk * s;
          v[2] += yk * gk;
        }
      });
  block_partial<3>(v, part);
}

// clip |y| <= 10 |zf|; fall back to zf when y is not a descent direction (src/minimizers.py:478-486);
// the three norms are summed from the k_pc_norms partials (stride 3)
__global__ __launch_bounds__(kBlock) void k_pc_finish(double* __restrict__ sc, const double* __restrict__ bank,
                                                      const double* __restrict__ part, int n_part,
                                                      double* __restrict__ y, const double* __restrict__ g,
                                                      const double* __restrict__ Mp, long N) {
  __shared__ double tot[3];
  {
    // the three strided sums, each in the fixed order of sum_partials_all
    const int tid = threadIdx.x;
    double s3[3] = {0.0, 0.0, 0.0};
    strided_sums<3, kBlock>(part, 0, n_part, s3);
    for (int k = 0; k < 3; ++k) {
      const double sk = wave_red<false>(s3[k]);
      __shared__ double sh[kBlock / 64];
      if ((tid & 63) == 0) sh[tid >> 6] = sk;
      __syncthreads();
      if (tid == 0) {
        double t = sh[0];
        for (int w = 1; w < kBlock / 64; ++w) t += sh[w];
        tot[k] = t;
      }
      __syncthreads();
    }
  }
  const double yy = tot[0], zz = tot[1], yg = tot[2];
  const double yn = sqrt(yy), zn = sqrt(zz);
  const double scale = (yn > 10.0 * zn) ? (10.0 * zn / (yn + 1e-30)) : 1.0;
  const bool fallback = !(scale * yg > 1e-12);
  if (fallback)  // uniform branches outside the loops, so the loads of a round go out together
    grid_chunks(
        N, [&](long i) { return Node4x3{load3(g, i), {}, {}, {}, Mp[i]}; },
        [&](long i, const Node4x3& u) {
          y[3 * i] = u.a.x * u.s;
          y[3 * i + 1] = u.a.y * u.s;
          y[3 * i + 2] = u.a.z * u.s;
        });
  else
    grid_chunks(
        3 * N, [&](long j) { return y[j]; }, [&](long j, double u) { y[j] = u * scale; });
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_YY] = yy;
    sc[S_ZZ] = zz;
    sc[S_YG] = yg;
    sc[S_PCIT] = bank[B_IT];
  }
}

// ----------------------------------------------------------------- direction kernels
// HS beta dots: y.(g_ext - g_prev), d_prev.(g_ext - g_prev)
__global__ __launch_bounds__(kBlock) void k_beta_dots(const double* __restrict__ y, const double* __restrict__ ge,
                                                      const double* __restrict__ gp, const double* __restrict__ dp,
                                                      long n3, double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  grid_chunks(
      n3, [&](long i) { return double4{ge[i], gp[i], y[i], dp[i]}; },
      [&](long, double4 u) {
        const double dg = u.x - u.y;
        v[0] += u.z * dg;
        v[1] += u.w * dg;
      });
  block_partial<2>(v, part);
}

// d = -y + beta T(m, d_prev) with beta from the k_beta_dots partials (y.dg, d_prev.dg); per-workgroup
// d.g_ext to part_dg  (src/minimizers.py:997-1012)
__global__ __launch_bounds__(kBlock) void k_direction(int restart, const double* __restrict__ m,
                                                      const double* __restrict__ y, const double* __restrict__ dp,
                                                      const double* __restrict__ ge, double* __restrict__ d, long N,
                                                      const double* __restrict__ part_nd, int n_nd,
                                                      double* __restrict__ part_dg) {
  double nd[2];
  sum_partials_k<2, kBlock>(part_nd, n_nd, nd);
  const double beta = restart ? 0.0 : fmax(0.0, nd[0] / (nd[1] + 1e-30));
  double v[1] = {0.0};
  grid_chunks<2>(
      N, [&](long i) { return Node4x3{load3(m, i), load3(dp, i), load3(y, i), load3(ge, i)}; },
      [&](long i, const Node4x3& u) {
        const double mi[3] = {u.a.x, u.a.y, u.a.z}, di[3] = {u.b.x, u.b.y, u.b.z};
        const double yi[3] = {u.c.x, u.c.y, u.c.z}, gi[3] = {u.d.x, u.d.y, u.d.z};
        double t[3];
        tangent3(mi, di, 1.0, t);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double dk = -yi[k] + beta * t[k];
          d[3 * i + k] = dk;
          v[0] += dk * gi[k];
        }
      });
  block_partial<1>(v, part_dg);
}

// if (force || d.g_ext > 0) d = -alt (d.g_ext from the part_dg partials); H = m x d; per-workgroup
// g_raw . d to part_pg
__global__ __launch_bounds__(kBlock) void k_direction_fix(int force, const double* __restrict__ m,
                                                          const double* __restrict__ alt, const double* __restrict__ g,
                                                          double* __restrict__ d, double* __restrict__ H, long N,
                                                          const double* __restrict__ part_dg, int n_dg,
                                                          double* __restrict__ part_pg) {
  const bool flip = force || sum_partials_all<kBlock>(part_dg, n_dg) > 0.0;
  double v[1] = {0.0};
  // the uniform flip is applied after the loads, which all go out first
  const double* src = flip ? alt : d;
  const double sgn = flip ? -1.0 : 1.0;
  grid_chunks<2>(
      N, [&](long i) { return Node4x3{load3(src, i), load3(g, i), load3(m, i), {}}; },
      [&](long i, const Node4x3& u) {
        const double dd[3] = {sgn * u.a.x, sgn * u.a.y, sgn * u.a.z};
        const double gi[3] = {u.b.x, u.b.y, u.b.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          d[3 * i + k] = dd[k];
          v[0] += gi[k] * dd[k];
        }
        const double mi[3] = {u.c.x, u.c.y, u.c.z};
        H[3 * i + 0] = mi[1] * dd[2] - mi[2] * dd[1];
        H[3 * i + 1] = mi[2] * dd[0] - mi[0] * dd[2];
        H[3 * i + 2] = mi[0] * dd[1] - mi[1] * dd[0];
      });
  block_partial<1>(v, part_pg);
}

// Cayley update m_t = cayley(m, H, s) (src/minimizers.py:71-79)
__global__ void k_cayley(const double* __restrict__ m, const double* __restrict__ H, double s,
                         double* __restrict__ out, long N) {
  NODE_LOOP(i, N) {
    const double k0 = 0.5 * s * H[3 * i], k1 = 0.5 * s * H[3 * i + 1], k2 = 0.5 * s * H[3 * i + 2];
    const double m0 = m[3 * i], m1 = m[3 * i + 1], m2 = m[3 * i + 2];
    const double kk = k0 * k0 + k1 * k1 + k2 * k2;
    const double km0 = k1 * m2 - k2 * m1, km1 = k2 * m0 - k0 * m2, km2 = k0 * m1 - k1 * m0;
    const double kd = k0 * m0 + k1 * m1 + k2 * m2;
    const double inv = 1.0 / (1.0 + kk);
    double a = ((1.0 - kk) * m0 + 2.0 * km0 + 2.0 * kd * k0) * inv;
    double b = ((1.0 - kk) * m1 + 2.0 * km1 + 2.0 * kd * k1) * inv;
    double c = ((1.0 - kk) * m2 + 2.0 * km2 + 2.0 * kd * k2) * inv;
    const double n = sqrt(a * a + b * b + c * c);
    out[3 * i] = a / n;
    out[3 * i + 1] = b / n;
    out[3 * i + 2] = c / n;
  }
}

__global__ __launch_bounds__(kBlock) void k_maxdiff(const double* __restrict__ a, const double* __restrict__ b,
                                                    long n, double* __restrict__ part) {
  double v[1] = {0.0};
  grid_chunks(
      n, [&](long i) { return double2{a[i], b[i]}; }, [&](long, double2 u) { v[0] = fmax(v[0], fabs(u.x - u.y)); });
  block_partial<1, kBlock, true>(v, part);
}

// ----------------------------------------------------------------------- L-BFGS kernels
// The history vectors S, Y are FP64, or FP32 in the mixed-precision forms (TS = float): the
// two-loop recursion reads them in FP64 arithmetic and keeps q, r, alpha, beta, rho and gamma
// in FP64.

// per-workgroup partials of a.b
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_lb_dot(const TS* __restrict__ a, const double* __restrict__ b, long n,
                                                   double* __restrict__ part) {
  double v[1] = {0.0};
  NODE_LOOP(i, n) v[0] += (double)a[i] * b[i];
  block_partial<1>(v, part);
}

// first loop: alpha_idx = rho_idx * (S_idx . q) from the k_lb_dot partials; q -= alpha_idx Y_idx
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_lb_first(double* __restrict__ sc, int idx, double* __restrict__ q,
                                                     const TS* __restrict__ Yi, long n,
                                                     const double* __restrict__ part, int n_part) {
  const double alpha = sc[S_RHOS + idx] * sum_partials_all<kBlock>(part, n_part);
  NODE_LOOP(i, n) q[i] -= alpha * (double)Yi[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_ALPHAS + idx] = alpha;
}

// per-workgroup partials of the gamma dots S_last.Y_last and Y_last.(Y_last invMr)
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_lb_gamma_dots(const TS* __restrict__ S, const TS* __restrict__ Y,
                                                          const double* __restrict__ invMr, long N,
                                                          double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  NODE_LOOP(i, N) {
    const double s = invMr[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double yk = (double)Y[3 * i + k];
      v[0] += (double)S[3 * i + k] * yk;
      v[1] += yk * yk * s;
    }
  }
  block_partial<2>(v, part);
}

// r = gamma q invMr, gamma = clip(sy/(ymy + 1e-30), 1e-3, 1e3) from the k_lb_gamma_dots partials, or 1
// without history
__global__ __launch_bounds__(kBlock) void k_lb_scale(int have_hist, const double* __restrict__ q,
                                                     const double* __restrict__ invMr, double* __restrict__ r, long N,
                                                     const double* __restrict__ part, int n_part) {
  double sy[2] = {0.0, 1.0};
  if (have_hist) sum_partials_k<2, kBlock>(part, n_part, sy);  // uniform per launch
  double gam = have_hist ? sy[0] / (sy[1] + 1e-30) : 1.0;
  gam = fmin(fmax(gam, 1e-3), 1e3);
  NODE_LOOP(i, N) {
    const double s = gam * invMr[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[3 * i + k] = s * q[3 * i + k];
  }
}

// second loop: beta = rho_idx * (Y_idx . r) from the k_lb_dot partials; r += S_idx (alpha_idx - beta)
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_lb_second(const double* __restrict__ sc, int idx, double* __restrict__ r,
                                                      const TS* __restrict__ Si, long n,
                                                      const double* __restrict__ part, int n_part) {
  const double f = sc[S_ALPHAS + idx] - sc[S_RHOS + idx] * sum_partials_all<kBlock>(part, n_part);
  NODE_LOOP(i, n) r[i] += (double)Si[i] * f;
}

template <typename TS>
__global__ void k_lb_zero(TS* __restrict__ dst, long n) {
  NODE_LOOP(i, n) dst[i] = (TS)0;
}

// history slot = (TS) v
template <typename TS>
__global__ void k_lb_store(TS* __restrict__ dst, const double* __restrict__ v, long n) {
  NODE_LOOP(i, n) dst[i] = (TS)v[i];
}

// d = -r and d.g_ext
__global__ __launch_bounds__(kBlock) void k_neg_dot(const double* __restrict__ r, const double* __restrict__ ge,
                                                    double* __restrict__ d, long n, double* __restrict__ part_dg) {
  double v[1] = {0.0};
  NODE_LOOP(i, n) {
    const double dk = -r[i];
    d[i] = dk;
    v[0] += dk * ge[i];
  }
  block_partial<1>(v, part_dg);
}

// secant pair at m_new: s = T(m, d tau), y = T(m, g_new) - T(m, g_ext_old); curv = y.s, ss = s.s
__global__ __launch_bounds__(kBlock) void k_lb_pair(const double* __restrict__ m, const double* __restrict__ d,
                                                    double tau, const double* __restrict__ gn,
                                                    const double* __restrict__ ge, double* __restrict__ s_out,
                                                    double* __restrict__ y_out, long N, double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  NODE_LOOP(i, N) {
    double s3[3], a[3], b[3];
    tangent3(m + 3 * i, d + 3 * i, tau, s3);
    // T(m, d tau) = tau T(m, d) up to rounding; compute exactly as written: T(m, d*tau)
    const double dt[3] = {d[3 * i] * tau, d[3 * i + 1] * tau, d[3 * i + 2] * tau};
    tangent3(m + 3 * i, dt, 1.0, s3);
    tangent3(m + 3 * i, gn + 3 * i, 1.0, a);
    tangent3(m + 3 * i, ge + 3 * i, 1.0, b);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double yk = a[k] - b[k];
      s_out[3 * i + k] = s3[k];
      y_out[3 * i + k] = yk;
      v[0] += yk * s3[k];
      v[1] += s3[k] * s3[k];
    }
  }
  block_partial<2>(v, part);
}

// transport every history vector to the tangent space at m (src/minimizers.py:1342-1347)
template <typename TS>
__global__ void k_lb_transport(const double* __restrict__ m, TS* __restrict__ S, TS* __restrict__ Y, int mem,
                               long N) {
  NODE_LOOP(i, N) {
    const double m0 = m[3 * i], m1 = m[3 * i + 1], m2 = m[3 * i + 2];
    for (int h = 0; h < mem; ++h) {
      TS* s = S + (long)h * 3 * N + 3 * i;
      TS* y = Y + (long)h * 3 * N + 3 * i;
      const double s0 = s[0], s1 = s[1], s2 = s[2];
      const double ds = m0 * s0 + m1 * s1 + m2 * s2;
      s[0] = (TS)(s0 - ds * m0);
      s[1] = (TS)(s1 - ds * m1);
      s[2] = (TS)(s2 - ds * m2);
      const double y0 = y[0], y1 = y[1], y2 = y[2];
      const double dy = m0 * y0 + m1 * y1 + m2 * y2;
      y[0] = (TS)(y0 - dy * m0);
      y[1] = (TS)(y1 - dy * m1);
      y[2] = (TS)(y2 - dy * m2);
    }
  }
}

__global__ void k_set_slot(double* __restrict__ sc, int slot, double v) { sc[slot] = v; }

// ------------------------------------------------------------------------- host driver
static int read_slots(Ctx& c, int first, int count) {
  MM_HIP(hipMemcpyAsync(c.hscal + first, c.dscal + first, count * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

// Per-workgroup partial banks of the field-step reductions (no fences or counters: a consuming
// kernel sums a bank in a fixed order, the host reduces the banks it needs after the iteration's
// synchronisation). At most 2 values per workgroup, at most kRedGridMax workgroups.
enum RedBank { R_GN = 0, R_GS = 1, R_PG = 2, R_DM = 3, R_CURV = 4, R_ND = 5, R_DG = 6, R_TMP = 7, R_SY = 8 };
constexpr int kBankStride = 2 * kRedGridMax;
static inline double* rbank(const Ctx& c, int r) { return c.mpart + (size_t)r * kBankStride; }

// Copies banks [0, nb) and the scalar slots S_E .. S_PCIT to the host and waits.
static int fetch_host(Ctx& c, int nb) {
  MM_HIP(hipMemcpyAsync(c.hmpart, c.mpart, (size_t)nb * kBankStride * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  MM_HIP(hipMemcpyAsync(c.hscal + S_E, c.dscal + S_E, (S_PCIT - S_E + 1) * sizeof(double), hipMemcpyDeviceToHost,
                        c.stream));
  MM_HIP(hipStreamSynchronize(c.stream));
  return 0;
}
static double host_bank_max(const Ctx& c, int r, int n) {
  const double* b = c.hmpart + (size_t)r * kBankStride;
  double m = 0.0;
  for (int i = 0; i < n; ++i) m = std::fmax(m, b[i]);
  return m;
}
static double host_bank_sum(const Ctx& c, int r, int n) {
  const double* b = c.hmpart + (size_t)r * kBankStride;
  double t = 0.0;
  for (int i = 0; i < n; ++i) t += b[i];
  return t;
}

// Energy and gradient of (m, U) into g (zeroed here); energy left in slot S_E.
static int eval_energy(Ctx& c, const double* m, const double* U, const double B[3], double* g) {
  vfill(g, 0.0, 3L * c.N, c.stream);
  MM_TRY(launch_energy(c, m, U, B, g, c.dscal + S_E, c.stream));
  c.n_energy++;
  return 0;
}

// Steihaug-PCG solve of P y = g_ext into c.y; iterations in slot S_PCIT.
static int solve_P(Ctx& c, const MinParams& p) {
  const long N = c.N;
  hipStream_t s = c.stream;
  double* sc = c.dscal;
  const int gr = red_grid(N);
  double* part_a = c.part_rz;  // p.q partials (the PCG is not running meanwhile)
  double* part_b = c.part_rr;  // r.z partials
  // the loop on FP32 vectors in the mixed-precision forms, on FP64 vectors otherwise
  auto loop = [&](auto* r, auto* z, auto* pv, auto* q, auto* cv) -> int {
    using TV = std::remove_pointer_t<decltype(r)>;
    k_pc_init<TV><<<gr, kBlock, 0, s>>>(c.g_ext, c.inv_M_prec, c.y, r, z, pv, cv, N, part_b);
    k_pc_state_init<<<1, kBlock, 0, s>>>(sc, part_b, gr, rbank(c, R_GN), red_grid(N), p.pc_auto, p.pc_force_eta,
                                         p.pc_force_alpha, p.pc_iters);
    for (int k = 0; k < p.pc_iters; ++k) {
      double* bank = sc + pc_bank(k);
      MM_TRY(launch_local(c, pv, cv, bank + B_ACT, s));
      c.n_local++;
      k_pc_ap<TV><<<gr, kBlock, 0, s>>>(bank, c.m, c.g_raw, pv, cv, q, p.pc_reg, N, part_a);
      k_pc_yr<TV><<<gr, kBlock, 0, s>>>(sc, bank, part_a, gr, p.pc_stagnation_nu, c.y, r, z, pv, q, c.inv_M_prec, N,
                                         part_b);
      k_pc_p<TV><<<gr, kBlock, 0, s>>>(sc, bank, sc + pc_bank(k + 1), part_b, gr, z, pv, cv, p.pc_iters, N);
    }
    return 0;
  };
  if (c.precision >= 1 && c.r3f != nullptr)
    MM_TRY(loop(c.r3f, c.z3f, c.p3f, c.q3f, c.cv3f));
  else
    MM_TRY(loop(c.r3, c.z3, c.p3, c.q3, c.cv3));
  k_pc_norms<<<gr, kBlock, 0, s>>>(c.y, c.g_ext, c.inv_M_prec, N, c.partial);
  k_pc_finish<<<gr, kBlock, 0, s>>>(sc, sc + pc_bank(p.pc_iters), c.partial, gr, c.y, c.g_ext, c.inv_M_prec, N);
  return 0;
}

struct LSResult {
  double tau, E;
  int evals, demag;
  bool moved;  // the trial buffers hold the accepted state
};

// Curvilinear Armijo line search (make_armijo_ls_v2). Trials are written to m_t, U_t, g_t;
// on return with moved == true they hold the accepted state.
static int line_search(Ctx& c, const MinParams& p, const double B[3], double pg, double E0, double s0, LSResult& out) {
  out.tau = 0.0;
  out.E = E0;
  out.evals = 0;
  out.demag = 0;
  out.moved = false;
  if (!(pg < 0.0)) return 0;
  MM_TRY(vcopy(c.U_t, c.U, c.N, c.stream));
  auto D = [&](double s, double& d, double& E) -> int {
    k_cayley<<<vg(c.N), kBlock, 0, c.stream>>>(c.m, c.H, s, c.m_t, c.N);
    int it = 0;
    double rr = 0;
    MM_TRY(solve_U_guess(c, c.m_t, c.U_t, p.phi_tol, p.cg_maxiter, &it, &rr));
    out.demag += it;
    MM_TRY(eval_energy(c, c.m_t, c.U_t, B, c.g_t));
    MM_TRY(read_slots(c, S_E, 1));
    E = c.hscal[S_E];
    if (!std::isfinite(E)) E = 1e20;
    d = (E - E0) / (s * pg + 1e-30);
    return 0;
  };
  const int maxe = p.ls_max_evals;
  double s = s0, s_min = 0.0, d_v = 0.0, E_v = E0;
  int it_exp = 0;
  bool done = false;
  while (it_exp < maxe && !done) {
    double dn, En;
    MM_TRY(D(s, dn, En));
    const bool stop = std::fabs(1.0 - dn) >= p.ls_eta2 || dn < 0.0;
    const double s_next = stop ? s : p.ls_C * s;
    if (!stop) s_min = s;
    s = s_next;
    d_v = dn;
    E_v = En;
    ++it_exp;
    done = stop;
  }
  const double s_min_exp = s_min;
  done = d_v >= p.ls_eta1 && d_v < 1e10;
  int it_con = 0;
  while (it_con < maxe && !done) {
    s = s_min_exp + p.ls_c * (s - s_min_exp);
    MM_TRY(D(s, d_v, E_v));
    done = d_v >= p.ls_eta1 && d_v < 1e10;
    ++it_con;
  }
  out.evals = it_exp + it_con;
  if (d_v >= 0.0) {
    out.tau = s;
    out.E = E_v;
    out.moved = true;
  }
  return 0;
}

int minimize_resident(Ctx& c, const MinParams& p, const double B[3], MinInfo& info) {
  MM_TRY(alloc_min_state(c));
  const auto t0 = std::chrono::steady_clock::now();
  const long N = c.N, n3 = 3 * N;
  hipStream_t s = c.stream;
  double* sc = c.dscal;
  const int mem = p.memory > 0 ? p.memory : 10;
  // L-BFGS history: FP64, or FP32 in the mixed-precision forms
  const bool lb32 = c.precision >= 1;
  if (p.method == 1 && ((lb32 ? !c.Sf : !c.S) || c.mem != mem)) {
    if (c.mem != mem) {
      for (void* old : {(void*)c.S, (void*)c.Yh, (void*)c.Sf, (void*)c.Yf})
        if (old) {
          auto itp = std::find(c.allocs.begin(), c.allocs.end(), old);
          if (itp != c.allocs.end()) c.allocs.erase(itp);
          (void)hipFree(old);
        }
      c.S = c.Yh = nullptr;
      c.Sf = c.Yf = nullptr;
    }
    c.mem = mem;
    const size_t bytes = (size_t)mem * n3 * (lb32 ? sizeof(float) : sizeof(double));
    void *S = nullptr, *Y = nullptr;
    MM_HIP(hipMalloc(&S, bytes));
    MM_HIP(hipMalloc(&Y, bytes));
    c.allocs.push_back(S);
    c.allocs.push_back(Y);
    if (lb32) {
      c.Sf = (float*)S;
      c.Yf = (float*)Y;
    } else {
      c.S = (double*)S;
      c.Yh = (double*)Y;
    }
  }

  // solve_and_minimize (src/minimizers.py:3408-3435)
  k_normalize<<<vg(N), kBlock, 0, s>>>(c.m, N);
  int demag = 0;
  {
    int it = 0;
    double rr = 0;
    MM_TRY(solve_U_guess(c, c.m, c.U, p.cg_tol, p.cg_maxiter, &it, &rr));
    demag = it;
  }
  MM_TRY(eval_energy(c, c.m, c.U, B, c.g_raw));
  const int nN = red_grid(N), n3g = red_grid(n3);  // workgroups (= partials) of the N- and 3N-loops
  k_tangents<<<nN, kBlock, 0, s>>>(c.m, c.g_raw, c.inv_M_rel, c.g_tan, c.g_ext, N, rbank(c, R_GN));
  MM_TRY(fetch_host(c, R_GN + 1));
  double E = c.hscal[S_E];
  double gnorm = host_bank_max(c, R_GN, nN);
  int evals = 1, preco = 0, it = 0;
  bool conv = false;
  const int Lp = p.L > 0 ? p.L : (int)N;

  if (p.method == 0) {
    // PCGState init: g = g_ext, y = g_ext, d = -g_ext
    MM_TRY(vcopy(c.g_prev, c.g_ext, n3, s));
    k_neg_dot<<<n3g, kBlock, 0, s>>>(c.g_ext, c.g_ext, c.d_prev, n3, rbank(c, R_DG));
  } else {
    if (lb32) {
      k_lb_zero<float><<<vg(mem * n3), kBlock, 0, s>>>(c.Sf, mem * n3);
      k_lb_zero<float><<<vg(mem * n3), kBlock, 0, s>>>(c.Yf, mem * n3);
    } else {
      vfill(c.S, 0.0, (long)mem * n3, s);
      vfill(c.Yh, 0.0, (long)mem * n3, s);
    }
    for (int h = 0; h < mem; ++h) k_set_slot<<<1, 1, 0, s>>>(sc, S_RHOS + h, 0.0);
  }
  int hist = 0;
  std::vector<double> h_rho(mem, 0.0);

  while (!conv && it < p.max_iter) {
    k_tangents<<<nN, kBlock, 0, s>>>(c.m, c.g_raw, c.inv_M_rel, c.g_tan, c.g_ext, N, rbank(c, R_GN));
    double s0 = 1.0;
    if (p.method == 0) {
      MM_TRY(solve_P(c, p));
      k_maxabs_part<<<n3g, kBlock, 0, s>>>(c.y, n3, rbank(c, R_GS));
      k_beta_dots<<<n3g, kBlock, 0, s>>>(c.y, c.g_ext, c.g_prev, c.d_prev, n3, rbank(c, R_ND));
      k_direction<<<nN, kBlock, 0, s>>>((it % Lp) == 0, c.m, c.y, c.d_prev, c.g_ext, c.d, N, rbank(c, R_ND), n3g,
                                         rbank(c, R_DG));
      k_direction_fix<<<nN, kBlock, 0, s>>>(0, c.m, c.y, c.g_raw, c.d, c.H, N, rbank(c, R_DG), nN, rbank(c, R_PG));
    } else {
      // two-loop recursion on q = g_ext (src/minimizers.py:1284-1312); q in r3, r in z3
      MM_TRY(vcopy(c.r3, c.g_ext, n3, s));
      const int nh = std::min(hist, mem);
      auto two_loop = [&](auto* S, auto* Y) {
        using TS = std::remove_const_t<std::remove_pointer_t<decltype(S)>>;
        for (int i = 0; i < nh; ++i) {
          const int idx = ((hist - 1 - i) % mem + mem) % mem;
          k_lb_dot<TS><<<n3g, kBlock, 0, s>>>(S + (long)idx * n3, c.r3, n3, rbank(c, R_TMP));
          k_lb_first<TS><<<n3g, kBlock, 0, s>>>(sc, idx, c.r3, Y + (long)idx * n3, n3, rbank(c, R_TMP), n3g);
        }
        if (hist > 0) {
          const int last = ((hist - 1) % mem + mem) % mem;
          k_lb_gamma_dots<TS><<<nN, kBlock, 0, s>>>(S + (long)last * n3, Y + (long)last * n3, c.inv_M_rel, N,
                                                     rbank(c, R_SY));
        }
        k_lb_scale<<<nN, kBlock, 0, s>>>(hist > 0, c.r3, c.inv_M_rel, c.z3, N, rbank(c, R_SY), nN);
        for (int i = 0; i < nh; ++i) {
          const int idx = ((hist - nh + i) % mem + mem) % mem;
          k_lb_dot<TS><<<n3g, kBlock, 0, s>>>(Y + (long)idx * n3, c.z3, n3, rbank(c, R_TMP));
          k_lb_second<TS><<<n3g, kBlock, 0, s>>>(sc, idx, c.z3, S + (long)idx * n3, n3, rbank(c, R_TMP), n3g);
        }
      };
      if (lb32)
        two_loop(c.Sf, c.Yf);
      else
        two_loop(c.S, c.Yh);
      k_neg_dot<<<n3g, kBlock, 0, s>>>(c.z3, c.g_ext, c.d, n3, rbank(c, R_DG));
      k_direction_fix<<<nN, kBlock, 0, s>>>(it == 0, c.m, c.g_tan, c.g_raw, c.d, c.H, N, rbank(c, R_DG), n3g,
                                             rbank(c, R_PG));
      s0 = (it == 0) ? p.tau0 : 1.0;
    }
    MM_TRY(fetch_host(c, R_PG + 1));  // banks GN, GS, PG and the slots (PCIT)
    const double gn_tan = host_bank_max(c, R_GN, nN);
    const double gs = (p.method == 0) ? host_bank_max(c, R_GS, n3g) : 0.0;
    const double pg = host_bank_sum(c, R_PG, nN);
    const int pit = (p.method == 0) ? (int)c.hscal[S_PCIT] : 0;
    const double gconv = (p.method == 0) ? gs : gn_tan;

    LSResult ls;
    MM_TRY(line_search(c, p, B, pg, E, s0, ls));
    double dm = 0.0;
    if (ls.moved) {
      k_maxdiff<<<n3g, kBlock, 0, s>>>(c.m_t, c.m, n3, rbank(c, R_DM));
      MM_TRY(fetch_host(c, R_DM + 1));
      dm = host_bank_max(c, R_DM, n3g);
    }
    const double E_new = ls.E;
    // check_convergence (src/minimizers.py:160-172)
    if (it > 0) {
      const bool u1 = (E - E_new) < p.tau_f * (1.0 + std::fabs(E_new));
      const bool u2 = dm < std::sqrt(p.tau_f) * 2.0;
      const bool u3 = gconv <= std::pow(p.tau_f, 1.0 / 3.0) * (1.0 + std::fabs(E_new));
      const bool u4 = gconv < p.eps_a;
      conv = (u1 && u2 && u3) || u4;
    }
    if (ls.moved) {
      std::swap(c.m, c.m_t);
      std::swap(c.U, c.U_t);
      std::swap(c.g_raw, c.g_t);
    }
    if (p.method == 0) {
      std::swap(c.g_prev, c.g_ext);
      std::swap(c.d_prev, c.d);
      gnorm = gs;
    } else {
      // secant pair, transport of the history, conditional insertion (src/minimizers.py:1330-1360)
      k_lb_pair<<<nN, kBlock, 0, s>>>(c.m, c.d, ls.tau, c.g_raw, c.g_ext, c.q3, c.cv3, N, rbank(c, R_CURV));
      MM_TRY(fetch_host(c, R_CURV + 1));
      double curv = 0.0, ss = 0.0;
      for (int w = 0; w < nN; ++w) {  // fixed order
        curv += c.hmpart[(size_t)R_CURV * kBankStride + 2 * w];
        ss += c.hmpart[(size_t)R_CURV * kBankStride + 2 * w + 1];
      }
      const bool ok = curv > 1e-12 * ss;
      const int idx = hist % mem;
      if (lb32)
        k_lb_transport<float><<<vg(N), kBlock, 0, s>>>(c.m, c.Sf, c.Yf, mem, N);
      else
        k_lb_transport<double><<<vg(N), kBlock, 0, s>>>(c.m, c.S, c.Yh, mem, N);
      if (ok) {
        if (lb32) {
          k_lb_store<float><<<vg(n3), kBlock, 0, s>>>(c.Sf + (long)idx * n3, c.q3, n3);
          k_lb_store<float><<<vg(n3), kBlock, 0, s>>>(c.Yf + (long)idx * n3, c.cv3, n3);
        } else {
          MM_TRY(vcopy(c.S + (long)idx * n3, c.q3, n3, s));
          MM_TRY(vcopy(c.Yh + (long)idx * n3, c.cv3, n3, s));
        }
        k_set_slot<<<1, 1, 0, s>>>(sc, S_RHOS + idx, 1.0 / (curv + 1e-30));
        ++hist;
      }
      gnorm = gn_tan;
    }
    E = E_new;
    evals += ls.evals;
    preco += pit;
    demag += ls.demag;
    ++it;
  }
  MM_HIP(hipStreamSynchronize(s));
  MM_HIP(hipGetLastError());
  info.iters = it;
  info.evals = evals;
  info.demag_iters = demag;
  info.preco_iters = preco;
  info.converged = conv ? 1 : 0;
  info.E = E;
  info.gnorm = gnorm;
  info.time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace mm
