// Solver entry points shared between translation units.
#pragma once

#include "mm_kernels.h"

namespace mm {

// Device-resident PCG on A U = b (b, U device; U holds the initial guess).
// q_ready: the caller has left A U in c.pq (the initial apply is skipped).
int pcg_dev(Ctx& c, const double* b, double* U, double tol, int maxiter, int* it_out, double* r2_out,
            double* bb_out, bool q_ready = false);

// Minimiser parameters / statistics mirrored from the C-ABI structs.
struct MinParams {
  int method;            // 0 pcohen_hs, 1 lbfgs
  int max_iter;
  double tau_f, eps_a, tau0;
  double ls_eta1, ls_eta2, ls_C, ls_c;
  int ls_max_evals;
  int pc_iters, pc_auto;
  double pc_force_eta, pc_force_alpha, pc_stagnation_nu, pc_reg;
  int L;                 // HS restart period (0: N)
  double cg_tol, phi_tol;
  int cg_maxiter;
  int memory;
};

struct MinInfo {
  int iters, evals, demag_iters, preco_iters, converged;
  double E, gnorm, time;
};

int minimize_resident(Ctx& c, const MinParams& p, const double B[3], MinInfo& info);
// pcohen_hs field step of every segment of a segmented context at once; info has nseg entries
int minimize_segmented(Ctx& c, const MinParams& p, const double B[3], MinInfo* info);
// per-segment volume-averaged J along the unit vector h (nseg values)
int j_parallel_segmented(Ctx& c, const double* m, const double h[3], double* out);

}  // namespace mm
