// CPU restatement (C++ / OpenMP) of the pcohen_hs field step on assembled operators.
// TEST INFRASTRUCTURE ONLY: the multi-core CPU baseline of bench.py and a checker in tests/.
// The product library never links it.
//
// It restates oracle/minimize.py and oracle/solvers.py line for line in compiled form, which in
// turn restate the reference:
//   PCG                     src/poisson_solve.py:422-478
//   V-cycle (SPAI0, amgcl)  src/amg_utils.py:557-623
//   energy / gradient       src/energy_kernels.py:147-180 (assembled: K_eff m + G U - 2 M B)
//   Cayley, tangent         src/minimizers.py:71-96
//   Armijo line search      src/minimizers.py:265-385
//   Steihaug-PCG            src/minimizers.py:393-492
//   pcohen_hs (HS beta)     src/minimizers.py:950-1073
//   convergence             src/minimizers.py:160-172
//   driver                  src/minimizers.py:3402-3452
// and plays the role of the reference's MKL C++ path (src/cpp/cpp_mkl_minimizer.cpp:542-763),
// with the AMG-preconditioned PCG in place of PARDISO.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

namespace {

struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> p, j;
  std::vector<double> x;
  void set(int r, int c, const int32_t* ip, const int32_t* ij, const double* ix) {
    rows = r;
    cols = c;
    p.assign(ip, ip + r + 1);
    j.assign(ij, ij + p[r]);
    x.assign(ix, ix + p[r]);
  }
  // y = A v
  void mv(const double* v, double* y) const {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; ++r) {
      double s = 0.0;
      for (int k = p[r]; k < p[r + 1]; ++k) s += x[k] * v[j[k]];
      y[r] = s;
    }
  }
};

struct Level {
  int n = 0;
  Csr A, P, R;  // A unused on level 0 (the fine operator is used)
  std::vector<double> M;
  std::vector<double> inv;  // coarsest level: dense inverse, row-major
};

struct Sys {
  int N = 0;
  Csr A, D, G, K;
  std::vector<double> Mn, invMr, invMp;
  double V = 1.0;
  std::vector<Level> lv;
};

std::mutex g_mu;
std::vector<std::unique_ptr<Sys>> g_sys;

using Vec = std::vector<double>;

double dot(const Vec& a, const Vec& b) {
  double s = 0.0;
  const long n = (long)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

double maxabs(const Vec& a) {
  double s = 0.0;
  const long n = (long)a.size();
#pragma omp parallel for reduction(max : s) schedule(static)
  for (long i = 0; i < n; ++i) s = std::max(s, std::fabs(a[i]));
  return s;
}

// ------------------------------------------------------------------ Poisson solve
// One V-cycle (oracle/solvers.py:vcycle, kind 'amgcl'): smoothing x += M (b - A x), one pre-
// and one post-sweep, zero start, exact coarsest solve.
Vec vcycle(const Sys& S, int l, const Vec& b) {
  const Level& L = S.lv[l];
  const int n = L.n;
  if (l == (int)S.lv.size() - 1) {
    Vec x(n, 0.0);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int q = 0; q < n; ++q) s += L.inv[(size_t)r * n + q] * b[q];
      x[r] = s;
    }
    return x;
  }
  const Csr& A = l == 0 ? S.A : L.A;
  Vec x(n), t(n), res(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) x[i] = L.M[i] * b[i];  // smooth from x = 0
  A.mv(x.data(), t.data());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) res[i] = b[i] - t[i];
  Vec bc(L.R.rows);
  L.R.mv(res.data(), bc.data());
  const Vec e = vcycle(S, l + 1, bc);
  L.P.mv(e.data(), t.data());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) x[i] += t[i];
  A.mv(x.data(), t.data());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) x[i] += L.M[i] * (b[i] - t[i]);
  return x;
}

// PCG (oracle/solvers.py:pcg, unmasked): returns iterations; U in/out
int pcg(const Sys& S, const Vec& b, Vec& x, double tol, int maxiter) {
  const int N = S.N;
  Vec r(N), Ap(N);
  S.A.mv(x.data(), Ap.data());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) r[i] = b[i] - Ap[i];
  Vec z = vcycle(S, 0, r);
  Vec p = z;
  double rz = dot(r, z), r2 = dot(r, r);
  const double tol2 = tol * tol * dot(b, b);
  int it = 0;
  while (it < maxiter && r2 > tol2) {
    S.A.mv(p.data(), Ap.data());
    const double alpha = rz / (dot(p, Ap) + 1e-30);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    z = vcycle(S, 0, r);
    const double rz_new = dot(r, z);
    const double beta = rz_new / (rz + 1e-30);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
    r2 = dot(r, r);
    rz = rz_new;
    ++it;
  }
  return it;
}

int solve_U(const Sys& S, const Vec& m, Vec& U, double tol, int maxiter) {
  Vec b(S.N);
  S.D.mv(m.data(), b.data());
  return pcg(S, b, U, tol, maxiter);
}

// ------------------------------------------------------------------ energy
// E and g of (m, U) at B (oracle/fem.py:energy_and_grad), both / V_mag
double energy_grad(const Sys& S, const Vec& m, const Vec& U, const double B[3], Vec& g) {
  const int N = S.N;
  Vec gu(3 * (size_t)N);
  S.K.mv(m.data(), g.data());
  S.G.mv(U.data(), gu.data());
  double E = 0.0;
  const double iv = 1.0 / S.V;
#pragma omp parallel for reduction(+ : E) schedule(static)
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 3; ++c) {
      const double gz = -2.0 * S.Mn[i] * B[c];
      const double gi = g[3 * i + c] + gu[3 * i + c] + gz;
      E += 0.5 * m[3 * i + c] * (gi + gz);
      g[3 * i + c] = gi * iv;
    }
  return E * iv;
}

// ------------------------------------------------------------------ manifold tools
void tangent(const Vec& m, const Vec& g, Vec& out) {
  const long N = (long)m.size() / 3;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) {
    const double md = m[3 * i] * g[3 * i] + m[3 * i + 1] * g[3 * i + 1] + m[3 * i + 2] * g[3 * i + 2];
    for (int c = 0; c < 3; ++c) out[3 * i + c] = g[3 * i + c] - md * m[3 * i + c];
  }
}

void cayley(const Vec& m, const Vec& H, double tau, Vec& out) {
  const long N = (long)m.size() / 3;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) {
    const double k0 = 0.5 * tau * H[3 * i], k1 = 0.5 * tau * H[3 * i + 1], k2 = 0.5 * tau * H[3 * i + 2];
    const double m0 = m[3 * i], m1 = m[3 * i + 1], m2 = m[3 * i + 2];
    const double kk = k0 * k0 + k1 * k1 + k2 * k2;
    const double km0 = k1 * m2 - k2 * m1, km1 = k2 * m0 - k0 * m2, km2 = k0 * m1 - k1 * m0;
    const double kd = k0 * m0 + k1 * m1 + k2 * m2;
    double a = ((1.0 - kk) * m0 + 2.0 * km0 + 2.0 * kd * k0) / (1.0 + kk);
    double b = ((1.0 - kk) * m1 + 2.0 * km1 + 2.0 * kd * k1) / (1.0 + kk);
    double c = ((1.0 - kk) * m2 + 2.0 * km2 + 2.0 * kd * k2) / (1.0 + kk);
    const double n = std::sqrt(a * a + b * b + c * c);
    out[3 * i] = a / n;
    out[3 * i + 1] = b / n;
    out[3 * i + 2] = c / n;
  }
}

struct Params {
  double tau_f, eps_a, ls_eta1, ls_eta2, ls_C, ls_c, pc_force_eta, pc_force_alpha, pc_stagnation_nu, pc_reg,
      cg_tol;
  int max_iter, ls_max_evals, pc_iters, pc_auto, cg_maxiter;
};

// Steihaug-PCG for P y = g_tan (oracle/minimize.py:solve_Py_g); returns the iterations
int solve_Py_g(const Sys& S, const Params& q, const Vec& m, const Vec& g_ext, const Vec& gte, double tol, Vec& y) {
  const int N = S.N;
  const long n3 = 3L * N;
  Vec mg(N);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i)
    mg[i] = m[3 * i] * g_ext[3 * i] + m[3 * i + 1] * g_ext[3 * i + 1] + m[3 * i + 2] * g_ext[3 * i + 2];
  Vec Cv(n3), Ap(n3);
  auto P = [&](const Vec& v) {
    S.K.mv(v.data(), Cv.data());
    const double iv = 1.0 / S.V;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      double c3[3];
      for (int c = 0; c < 3; ++c) c3[c] = Cv[3 * i + c] * iv;
      const double mc = m[3 * i] * c3[0] + m[3 * i + 1] * c3[1] + m[3 * i + 2] * c3[2];
      for (int c = 0; c < 3; ++c) Ap[3 * i + c] = c3[c] - mc * m[3 * i + c] - mg[i] * v[3 * i + c] + q.pc_reg * v[3 * i + c];
    }
  };
  y.assign(n3, 0.0);
  Vec r = gte, z(n3), p(n3);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n3; ++k) z[k] = r[k] * S.invMp[k / 3];
  p = z;
  double rho = dot(r, z);
  const double target = tol * tol * rho;
  double Q = 0.0;
  int it = 0;
  bool done = false;
  while (it < q.pc_iters && rho > target && rho > 1e-25 && rho < 1e20 && !done) {
    P(p);
    const double pAp = dot(p, Ap);
    const bool neg = pAp <= 0.0;
    const double alpha = rho / (pAp + 1e-30);
    const double dq = 0.5 * alpha * rho;
    const bool stag = it > 0 && dq <= q.pc_stagnation_nu * Q;
    done = neg || stag;
    if (!done) {
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n3; ++k) {
        y[k] += alpha * p[k];
        r[k] -= alpha * Ap[k];
      }
    }
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n3; ++k) z[k] = r[k] * S.invMp[k / 3];
    const double rho_n = dot(r, z);
    if (!done) {
      const double beta = rho_n / (rho + 1e-30);
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n3; ++k) p[k] = z[k] + beta * p[k];
    }
    rho = rho_n;
    Q += dq;
    ++it;
  }
  Vec zf(n3);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n3; ++k) zf[k] = gte[k] * S.invMp[k / 3];
  const double yn = std::sqrt(dot(y, y)), zn = std::sqrt(dot(zf, zf));
  if (yn > 10.0 * zn) {
    const double sc = 10.0 * zn / (yn + 1e-30);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n3; ++k) y[k] *= sc;
  }
  if (!(dot(y, gte) > 1e-12)) y = zf;
  return it;
}

struct Info {
  int iters, evals, demag, preco, converged;
  double E, gnorm;
};

// Field step, pcohen_hs (oracle/minimize.py:Minimizer.minimize)
int minimize(const Sys& S, const Params& q, Vec& m, Vec& U, const double B[3], Info& info) {
  const int N = S.N;
  const long n3 = 3L * N;
  const double phi_tol = std::min(q.cg_tol, q.tau_f * 0.1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; ++i) {
    const double n = std::sqrt(m[3 * i] * m[3 * i] + m[3 * i + 1] * m[3 * i + 1] + m[3 * i + 2] * m[3 * i + 2]);
    for (int c = 0; c < 3; ++c) m[3 * i + c] /= n;
  }
  int demag = solve_U(S, m, U, q.cg_tol, q.cg_maxiter);
  Vec g_raw(n3), g_tan(n3), gte(n3), scaled(n3);
  double E = energy_grad(S, m, U, B, g_raw);
  auto tangents = [&]() {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n3; ++k) scaled[k] = g_raw[k] * S.invMr[k / 3];
    tangent(m, scaled, g_tan);
    tangent(m, g_raw, gte);
  };
  tangents();
  double gnorm = maxabs(g_tan);
  int evals = 1, preco = 0, it = 0;
  bool conv = false;
  const int Lp = N;
  Vec g_prev = gte, d_prev(n3), y(n3), d(n3), H(n3), dg(n3), tdp(n3);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n3; ++k) d_prev[k] = -gte[k];
  Vec m_t(n3), g_t(n3), U_t(N), m_v(n3), g_v(n3), U_v(N);
  while (!conv && it < q.max_iter) {
    tangents();
    const double gi = maxabs(g_tan);
    const double pc_tol = q.pc_auto ? std::min(q.pc_force_eta, std::pow(gi, q.pc_force_alpha)) : 0.0;
    const int pit = solve_Py_g(S, q, m, g_raw, gte, pc_tol, y);
    const double gs = maxabs(y);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n3; ++k) dg[k] = gte[k] - g_prev[k];
    const double beta = (it % Lp == 0) ? 0.0 : std::max(0.0, dot(y, dg) / (dot(d_prev, dg) + 1e-30));
    tangent(m, d_prev, tdp);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n3; ++k) d[k] = -y[k] + beta * tdp[k];
    if (dot(d, gte) > 0) {
#pragma omp parallel for schedule(static)
      for (long k = 0; k < n3; ++k) d[k] = -y[k];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      const double* mi = &m[3 * i];
      const double* di = &d[3 * i];
      H[3 * i] = mi[1] * di[2] - mi[2] * di[1];
      H[3 * i + 1] = mi[2] * di[0] - mi[0] * di[2];
      H[3 * i + 2] = mi[0] * di[1] - mi[1] * di[0];
    }
    const double pg = dot(g_raw, d);
    // Armijo line search (oracle/minimize.py:armijo_ls), s0 = 1
    double tau = 0.0, E_new = E;
    bool moved = false;
    int le = 0, ld = 0;
    if (pg < 0) {
      U_v = U;
      auto D = [&](double s, double& dn, double& Et) {
        cayley(m, H, s, m_t);
        U_t = U_v;
        ld += solve_U(S, m_t, U_t, phi_tol, q.cg_maxiter);
        Et = energy_grad(S, m_t, U_t, B, g_t);
        if (!std::isfinite(Et)) Et = 1e20;
        dn = (Et - E) / (s * pg + 1e-30);
        m_v = m_t;
        g_v = g_t;
        U_v = U_t;
      };
      double s = 1.0, s_min = 0.0, d_v = 0.0, E_v = E;
      int ite = 0;
      bool done = false;
      while (ite < q.ls_max_evals && !done) {
        double dn, Et;
        D(s, dn, Et);
        const bool stop = std::fabs(1.0 - dn) >= q.ls_eta2 || dn < 0;
        const double s_next = stop ? s : q.ls_C * s;
        if (!stop) s_min = s;
        s = s_next;
        ++ite;
        done = stop;
        d_v = dn;
        E_v = Et;
      }
      const double s_min_exp = s_min;
      done = d_v >= q.ls_eta1 && d_v < 1e10;
      int itc = 0;
      while (itc < q.ls_max_evals && !done) {
        s = s_min_exp + q.ls_c * (s - s_min_exp);
        D(s, d_v, E_v);
        done = d_v >= q.ls_eta1 && d_v < 1e10;
        ++itc;
      }
      le = ite + itc;
      if (d_v >= 0) {
        tau = s;
        E_new = E_v;
        moved = true;
      }
    }
    (void)tau;
    // convergence (oracle/minimize.py:check_convergence)
    if (it > 0) {
      double dm = 0.0;
      if (moved) {
#pragma omp parallel for reduction(max : dm) schedule(static)
        for (long k = 0; k < n3; ++k) dm = std::max(dm, std::fabs(m_v[k] - m[k]));
      }
      const bool u1 = (E - E_new) < q.tau_f * (1.0 + std::fabs(E_new));
      const bool u2 = dm < std::sqrt(q.tau_f) * 2.0;
      const bool u3 = gs <= std::pow(q.tau_f, 1.0 / 3.0) * (1.0 + std::fabs(E_new));
      const bool u4 = gs < q.eps_a;
      conv = (u1 && u2 && u3) || u4;
    }
    g_prev = gte;
    d_prev = d;
    if (moved) {
      m.swap(m_v);
      U.swap(U_v);
      g_raw.swap(g_v);
    }
    E = E_new;
    gnorm = gs;
    ++it;
    evals += le;
    preco += pit;
    demag += ld;
  }
  info.iters = it;
  info.evals = evals;
  info.demag = demag;
  info.preco = preco;
  info.converged = conv ? 1 : 0;
  info.E = E;
  info.gnorm = gnorm;
  return 0;
}

Sys* get(int64_t h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (h < 1 || h > (int64_t)g_sys.size()) return nullptr;
  return g_sys[h - 1].get();
}

}  // namespace

extern "C" {

int ocpu_threads() { return omp_get_max_threads(); }

// Assembled operators: A (N x N), D (N x 3N, interleaved), G (3N x N), K_eff (3N x 3N), all CSR.
int64_t ocpu_create(int N, const int32_t* Ap, const int32_t* Aj, const double* Ax, const int32_t* Dp,
                    const int32_t* Dj, const double* Dx, const int32_t* Gp, const int32_t* Gj, const double* Gx,
                    const int32_t* Kp, const int32_t* Kj, const double* Kx, const double* M_nodal,
                    const double* inv_M_rel, const double* inv_M_prec, double V_mag) {
  auto S = std::make_unique<Sys>();
  S->N = N;
  S->A.set(N, N, Ap, Aj, Ax);
  S->D.set(N, 3 * N, Dp, Dj, Dx);
  S->G.set(3 * N, N, Gp, Gj, Gx);
  S->K.set(3 * N, 3 * N, Kp, Kj, Kx);
  S->Mn.assign(M_nodal, M_nodal + N);
  S->invMr.assign(inv_M_rel, inv_M_rel + N);
  S->invMp.assign(inv_M_prec, inv_M_prec + N);
  S->V = V_mag;
  std::lock_guard<std::mutex> lk(g_mu);
  g_sys.push_back(std::move(S));
  return (int64_t)g_sys.size();
}

// Level l of the V-cycle: n rows; A (levels > 0), smoother diagonal M; P (n x n_next) and R
// (n_next x n) above the last level; dense inverse (n x n, row-major) on the last level.
int ocpu_add_level(int64_t h, int n, const int32_t* Ap, const int32_t* Aj, const double* Ax, const double* M,
                   int n_next, const int32_t* Pp, const int32_t* Pj, const double* Px, const int32_t* Rp,
                   const int32_t* Rj, const double* Rx, const double* dense_inv) {
  Sys* S = get(h);
  if (!S || n <= 0 || !M) return -1;
  Level L;
  L.n = n;
  L.M.assign(M, M + n);
  if (Ap) L.A.set(n, n, Ap, Aj, Ax);
  if (n_next > 0) {
    if (!Pp || !Rp) return -1;
    L.P.set(n, n_next, Pp, Pj, Px);
    L.R.set(n_next, n, Rp, Rj, Rx);
  }
  if (dense_inv) L.inv.assign(dense_inv, dense_inv + (size_t)n * n);
  S->lv.push_back(std::move(L));
  return 0;
}

// params: tau_f, eps_a, ls_eta1, ls_eta2, ls_C, ls_c, pc_force_eta, pc_force_alpha,
// pc_stagnation_nu, pc_reg, cg_tol; iparams: max_iter, ls_max_evals, pc_iters, pc_auto, cg_maxiter.
// info_i: iters, evals, demag_iters, preco_iters, converged; info_d: E, gnorm.
int ocpu_minimize(int64_t h, const double* params, const int32_t* iparams, double* m_inout, double* U_inout,
                  const double* B, int32_t* info_i, double* info_d) {
  Sys* S = get(h);
  if (!S || S->lv.empty()) return -1;
  Params q{params[0], params[1], params[2], params[3], params[4], params[5], params[6], params[7],
           params[8], params[9], params[10], iparams[0], iparams[1], iparams[2], iparams[3], iparams[4]};
  const long n3 = 3L * S->N;
  Vec m(m_inout, m_inout + n3), U(U_inout, U_inout + S->N);
  Info info{};
  minimize(*S, q, m, U, B, info);
  std::memcpy(m_inout, m.data(), n3 * sizeof(double));
  std::memcpy(U_inout, U.data(), S->N * sizeof(double));
  info_i[0] = info.iters;
  info_i[1] = info.evals;
  info_i[2] = info.demag;
  info_i[3] = info.preco;
  info_i[4] = info.converged;
  info_d[0] = info.E;
  info_d[1] = info.gnorm;
  return 0;
}

int ocpu_destroy(int64_t h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (h < 1 || h > (int64_t)g_sys.size()) return -1;
  g_sys[h - 1].reset();
  return 0;
}

}  // extern "C"
