/*
 * mumag_hip.h -- C-ABI of the MI355X-native micromagnetics hot path.
 *
 * One shared library, libmumag_hip.so, loaded with ctypes. Every entry point returns an
 * int status: 0 on success, a negative value on error; mm_last_error() returns the
 * message. All pointers are caller-owned host arrays (float64 / int32, C-contiguous)
 * unless noted. A context owns its device buffers and one HIP stream; calls on one
 * context are serialised by the caller, distinct contexts may be used concurrently.
 *
 * Units and meaning follow the reference: lengths in mesh units (nm), material constants
 * already reduced (A_red, K1_red, Js_red, src/loop.py:862-877), fields B in units of
 * Js_ref, energies and gradients divided by the magnetic volume V_mag.
 */
#ifndef MUMAG_HIP_H
#define MUMAG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Minimiser parameters: the ``params`` dict of ``minimize(m0, B_ext, **params)``
 * (src/minimizers.py:3437-3452) and the LoopParams fields it reads
 * (src/hysteresis_loop.py:34-125). method: 0 = pcohen_hs, 1 = lbfgs. */
typedef struct mm_params {
  int method;
  int max_iter;
  double tau_f, eps_a, tau0;
  double ls_eta1, ls_eta2, ls_C, ls_c;
  int ls_max_evals;
  int pc_iters, pc_auto;
  double pc_force_eta, pc_force_alpha, pc_stagnation_nu, pc_reg;
  int L;
  double cg_tol, phi_tol;
  int cg_maxiter;
  int memory;
} mm_params;

/* Per-field-step statistics: the ``info`` dict returned by ``minimize``
 * (src/minimizers.py:4084-4096) and by ``cpp_minimize`` (src/cpp_minimizer.py:181-189). */
typedef struct mm_info {
  int iters, evals, demag_iters, preco_iters, converged;
  double E, gnorm, time;
} mm_info;

/* Last error message of the calling process. Returns the message length. */
int mm_last_error(char* buf, int len);

/* Library / device description ("gfx950 ..."). */
int mm_device_info(char* buf, int len);

/* Context creation: replaces the host set-up of src/loop.py:856-939 (compute_volume_JinvT,
 * TetGeom, M_nodal) and the operator construction of src/loop.py:1089-1176.
 * knt (N,3); conn (E,4) 0-based; mat_id (E,) 1-based; A_red, K1_red, Js_red (G,);
 * k_easy (G,3); bnd_mask (N,) 1.0 interior / 0.0 Dirichlet, or NULL (no Dirichlet nodes);
 * reg: Poisson diagonal regularisation (src/amg_utils.py:73).
 * flags bit 0..1: geometry source of the Poisson apply (0 on-the-fly from coordinates,
 * 1 stored FP64 gradients, 2 stored FP32 gradients); bits 8..16: distinct nodes per element
 * block, 16..256 (0: 256). */
int mm_ctx_create(int N, int E, const double* knt, const int32_t* conn, const int32_t* mat_id, int G,
                  const double* A_red, const double* K1_red, const double* Js_red, const double* k_easy,
                  const double* bnd_mask, double reg, int flags, int64_t* out_ctx);
int mm_ctx_destroy(int64_t ctx);

/* Context over several independent systems (segments) at once, e.g. the grains of a grain set
 * concatenated into one mesh: segment g owns nodes seg_node_off[g] .. seg_node_off[g+1]-1
 * (seg_node_off[0] = 0, seg_node_off[nseg] = N) and every element lies in one segment. Element
 * blocks never span two segments, so mm_solve_U_segmented and mm_minimize_segmented run all
 * segments in one launch per kernel with per-segment scalars. Each segment keeps its own
 * magnetic volume V_mag (energies, gradients and the minimiser's scalings are per segment), so
 * a segment behaves as a context of its own mesh; mm_energy_grad on the whole context returns
 * the sum of the segments' energies. Other arguments as mm_ctx_create. No reference
 * counterpart: the reference solves one system per solve_U call (src/poisson_solve.py:748-771). */
int mm_ctx_create_segmented(int N, int E, const double* knt, const int32_t* conn, const int32_t* mat_id, int G,
                            const double* A_red, const double* K1_red, const double* Js_red, const double* k_easy,
                            const double* bnd_mask, double reg, int flags, int nseg, const int32_t* seg_node_off,
                            int64_t* out_ctx);

/* info[0..14] = V_mag, N, E, element blocks (all), element blocks (magnetic),
 * largest node list of a block, magnetic elements, node-list entries of all blocks,
 * cluster blocks, node-list entries of the cluster blocks, edge-star clusters, block node cap,
 * cluster block node cap, cluster staging mode (mm_set_cluster_staging), block colours of the
 * colour-sequenced scatter (0 when off; mm_set_apply_scatter) */
int mm_ctx_info(int64_t ctx, double* info);

/* Oriented connectivity (E,4), Ve (E,) and shape-function gradients of corners 1..3 (E,3,3)
 * in the context's internal element order (src/loop.py:56-102). */
int mm_ctx_get_geometry(int64_t ctx, int32_t* conn, double* Ve, double* grad);

/* M_nodal, inv_M_rel and inv_M_prec (N,) (src/loop.py:929-939, src/hysteresis_loop.py:340-353). */
int mm_ctx_get_nodal(int64_t ctx, double* M_nodal, double* inv_M_rel, double* inv_M_prec);

/* diag(K_ex)/V_mag (N,): compute_exchange_diagonal (src/energy_kernels.py:492-566). */
int mm_ctx_get_exchange_diag(int64_t ctx, double* d);

/* Precision of the Poisson solve: 0 = FP64; 1 = mixed: the whole V-cycle in FP32 (FP32
 * matrix-free applies and coarse levels) while the PCG recurrence, its operator apply, the dot
 * products and alpha/beta stay FP64; 2 = FP32 PCG with FP64 residual replacement: the PCG runs on
 * FP32 vectors with the FP32 apply and V-cycle (dots, alpha and beta in FP64); its recursive
 * residual is replaced by the true residual b - A U of the FP64 apply after each reduction by
 * mm_set_refinement's factor, the Krylov direction carrying on, and it stops when the true
 * r.r <= tol^2 b.b (needs the AMG V-cycle and the on-the-fly cluster apply); its V-cycle reads
 * the transfer and coarse-level matrix values from FP16 copies, each matrix scaled by a power of
 * two that puts its largest magnitude in [2^13, 2^14) (products and sums in FP32, the row sums
 * scaled back; made when every value is finite); 3 = form 2 with the
 * energy densities and gradient terms of mm_energy_grad and of the minimiser in FP32 element
 * arithmetic (FP64 accumulation, FP64 Zeeman term, compensated FP64 energy sum). Every form meets
 * the FP64 solver tolerance; forms 1-3 run the minimiser's Steihaug loop on FP32 vectors (r, z,
 * p, q and the local Hessian action C p, which is FP32 element arithmetic; FP64 dot products,
 * FP64 result y) and keep the L-BFGS history in FP32. */
int mm_set_precision(int64_t ctx, int precision);

/* Select the geometry source of the Poisson apply (see flags of mm_ctx_create). */
int mm_set_geometry_mode(int64_t ctx, int mode);

/* Work shape of the on-the-fly Poisson apply: unit 0 = one tet per thread, 1 = one edge-star
 * cluster per thread (see mm_edge_clusters), with param = workgroup size (64, 128 or 256);
 * unit 2 = clusters on a persistent grid of param (1..4) workgroups per CU that prefetches
 * its next block while computing the current one (FP64 applies; the FP32 V-cycle keeps unit 1).
 * Results agree to rounding; the choice is a tuning knob with no reference counterpart. */
int mm_set_apply_shape(int64_t ctx, int unit, int param);

/* Distinct nodes per block of the edge-star cluster partition of the Poisson apply: 256 (the
 * default), 512, 1024 or 2048. Larger blocks stage fewer node entries and share fewer nodes with
 * other blocks (fewer global atomics); they run 256 threads per workgroup (512 for 2048-node
 * blocks) and not on the persistent grid. Rebuilds the cluster partition (not allowed on
 * segmented contexts). Results agree to rounding; a tuning knob with no reference counterpart. */
int mm_set_cluster_block_nodes(int64_t ctx, int nb);
/* How the FP32 applies (V-cycle, FP32 PCG) on cluster blocks of 512 to 2048 nodes stage their
 * nodes: 0 = load the node ids, then gather the FP64 coordinates; 1 (default) = stream per-entry
 * records (block-relative FP32 coordinates with the node id, in block order, 16 bytes per
 * node-list entry kept on the device). Same staged values, same results; the FP64 apply always
 * gathers. A tuning knob with no reference counterpart. */
int mm_set_cluster_staging(int64_t ctx, int mode);
/* How the cluster apply adds the block sums of nodes that several blocks touch: 0 (default) =
 * hardware FP64 atomics (global_atomic_add_f64) in one launch; 1 = colour-sequenced: the blocks
 * are coloured so that no two blocks of a colour share a node and run as one launch per colour,
 * adding shared-node sums with plain loads and stores (no global atomics; the sums inside a block
 * still come from LDS atomics). Not on segmented contexts or the persistent grid. A tuning knob with no reference
 * counterpart: the two scatters the path may use, kept selectable so either can be measured. */
int mm_set_apply_scatter(int64_t ctx, int mode);
/* FP32 PCG of precision forms 2 and 3: the recursive FP32 residual is replaced by the true FP64
 * residual b - A U (and the FP32 correction added to U) each time its norm has dropped by the
 * factor eta, in (0, 1), since the last replacement (default 1e-3); the Krylov direction carries
 * on. The stopping test is the true residual's, as in the FP64 PCG. A tuning knob with no reference
 * counterpart. */
int mm_set_refinement(int64_t ctx, double eta);

/* y = A x: assembled-mode apply_A (src/poisson_solve.py:166-171) computed matrix-free. */
int mm_poisson_apply(int64_t ctx, const double* x, double* y);
/* y_r = A x_r for k in {1, 2, 4, 8} vectors at once: x, y (k, N) row-major. One staging of the
 * mesh serves all k vectors (the unit of a multi-right-hand-side solve); needs the on-the-fly
 * cluster apply on 256-node cluster blocks. No reference counterpart (the reference applies A
 * to one vector per call, src/poisson_solve.py:166-171). */
int mm_poisson_apply_multi(int64_t ctx, int k, const double* x, double* y);
/* b = D m: rhs_from_m (src/poisson_solve.py:173-186, 254-280); m (N,3). */
int mm_poisson_rhs(int64_t ctx, const double* m, double* b);
/* diag(A): assemble_diag (src/poisson_solve.py:282-299) of the masked, regularised A. */
int mm_poisson_diag(int64_t ctx, double* d);

/* Preconditioner set-up. Level 0 uses the matrix-free apply; levels l >= 1 carry CSR A.
 * mm_amg_set_level(l): n rows; M (n,) smoother multiplier (SPAI0 diagonal for amgcl,
 * 0.6667/diag for amg); P (n x n_next) and R (n_next x n) in CSR (NULL on the last level);
 * A in CSR for l >= 1 (NULL for l == 0); dense_inv (n,n) row-major on the last level.
 * Replaces the hierarchy transfer of src/poisson_solve.py:600-639. */
int mm_amg_set_level(int64_t ctx, int l, int n, int n_next, const int32_t* Ap, const int32_t* Ai, const double* Ax,
                     const double* M, const int32_t* Pp, const int32_t* Pi, const double* Px, const int32_t* Rp,
                     const int32_t* Ri, const double* Rx, const double* dense_inv);
/* kind 1 = amgcl (SPAI0), 2 = amg (damped Jacobi); nlev levels set above. */
int mm_amg_commit(int64_t ctx, int kind, int nlev);
/* Lanes per row (1, 2, 4, 8, 16, 32 or 64) of the CSR products of level l: which = 0 its A, 1 its P,
 * 2 its R. mm_amg_set_level chooses them from the average row length; the SpMV tuner sets them
 * from measured V-cycle times. Results agree to rounding; no reference counterpart. */
int mm_amg_set_lanes(int64_t ctx, int l, int which, int lanes);
int mm_amg_get_lanes(int64_t ctx, int l, int32_t* out3);
/* Jacobi preconditioner on diag(A) (precond_type 'jacobi', src/poisson_solve.py:409-421). */
int mm_precond_jacobi(int64_t ctx);

/* One preconditioner application z = M^-1 r (src/amg_utils.py:557-623). */
int mm_precond_apply(int64_t ctx, const double* r, double* z);

/* PCG on A x = b for a caller-given b: make_pcg_solve's solve (src/poisson_solve.py:422-478);
 * x_inout holds x0 on entry; r2 receives the final r.r. */
int mm_pcg_solve(int64_t ctx, const double* b, double* x_inout, double tol, int maxiter, int* it, double* r2);

/* solve_U (src/poisson_solve.py:748-771): U_inout holds x0 on entry. */
int mm_solve_U(int64_t ctx, const double* m, double* U_inout, double tol, int maxiter, int* it, double* rel_res);

/* solve_U (src/poisson_solve.py:748-771) for every segment of a segmented context at once:
 * one launch per kernel for all segments, while each segment keeps its own alpha, beta and
 * stopping test (r_g.r_g <= tol[g]^2 b_g.b_g); it[g] and rel_res[g] per segment, as
 * mm_solve_U would return for the segment alone. Needs the AMG V-cycle (FP64 or mixed). */
int mm_solve_U_segmented(int64_t ctx, const double* m, double* U_inout, const double* tol, int maxiter, int* it,
                         double* rel_res);

/* energy_and_grad (src/energy_kernels.py:340-459): E scalar, g (N,3). */
int mm_energy_grad(int64_t ctx, const double* m, const double* U, const double* B, double* E, double* g);
/* energy_and_grad on device-resident arrays, allocating nothing: m_dev (N,3), U_dev (N,), E_dev
 * (1 double) and g_dev (N,3) are device pointers, B three host doubles; launched on `stream`
 * (a hipStream_t; NULL: the context's stream) without synchronising. The energy's block partials
 * go to a fixed 512 KiB device array of the library instead of a context buffer, so meshes with
 * more than 32768 magnetic element blocks are rejected. Every call on one GPU shares that array:
 * calls that overlap on one GPU (two streams, two threads) race. No reference counterpart; for
 * callers that cannot allocate device memory (HIP graph capture, serving loops). */
int mm_energy_grad_noalloc(int64_t ctx, const double* m_dev, const double* U_dev, const double* B, double* E_dev,
                           double* g_dev, void* stream);
/* local_grad_only (src/energy_kernels.py:272-338): g = (K_ex + K_an) v / V_mag. */
int mm_local_grad(int64_t ctx, const double* v, double* g);

/* Volume-averaged polarisation along h (src/hysteresis_loop.py:153-190), reduced units. */
int mm_j_parallel(int64_t ctx, const double* m, const double* h, double* J);

/* Device-resident state for a field sweep: upload / download m (N,3) and U (N,). */
int mm_state_set(int64_t ctx, const double* m, const double* U);
int mm_state_get(int64_t ctx, double* m, double* U);
/* Validate a parameter struct without touching a device: method 0 or 1, memory in [1, 32]
 * (0 selects 10), non-negative iteration limits, finite non-negative tolerances. The
 * minimiser entry points run the same checks. The reference validates nothing here; the cap on
 * memory is this library's (the history's scalars live in a fixed device table). */
int mm_params_check(const mm_params* p);
/* Minimise the resident state at field B (one field step, src/minimizers.py:3402-3452). */
int mm_minimize_resident(int64_t ctx, const mm_params* p, const double* B, mm_info* info);
/* mm_state_set + mm_minimize_resident + mm_state_get; the role of
 * run_cpp_pcohen_hs_minimization (src/cpp/cpp_mkl_minimizer.cpp:542-763). */
int mm_minimize(int64_t ctx, const mm_params* p, double* m_inout, double* U_inout, const double* B, mm_info* info);
/* J_parallel of the resident state. */
int mm_state_j_parallel(int64_t ctx, const double* h, double* J);

/* Initial guess of the Poisson solves inside mm_minimize / mm_minimize_resident (the line-
 * search trials and the first solve of each field step): k = 0 (default) starts from the
 * previous solution, as the reference warm-starts each trial from the previous trial's U
 * (src/minimizers.py:305,348); k in 1..8 also forms the combination of the last k solutions
 * whose right-hand sides best match the new one in the least-squares sense (Fischer 1998,
 * successive right-hand sides) and starts from whichever of the two has the smaller true
 * residual; the PCG then runs to the same tolerance. Resets the stored solutions. No
 * reference counterpart. */
int mm_set_solve_guess(int64_t ctx, int k);

/* Field step of every segment of a segmented context at once (pcohen_hs only): the resident
 * state of segment g is minimised at field B as mm_minimize_resident would minimise it in a
 * context of its own, with its own line search, Steihaug loop and convergence test, while all
 * segments share each kernel launch; info has nseg entries (time: the whole call). Segments
 * that converge keep their state while the others iterate. Needs the AMG V-cycle (FP64 or mixed).
 * mm_minimize_resident rejects segmented contexts. No reference counterpart (the reference
 * minimises one system per call, src/minimizers.py:4053-4096). */
int mm_minimize_segmented(int64_t ctx, const mm_params* p, const double* B, mm_info* info);
/* J_parallel of each segment's resident state (nseg values). */
int mm_state_j_parallel_segmented(int64_t ctx, const double* h, double* J);
/* Magnetic volume V_mag of each segment (nseg values). */
int mm_seg_volumes(int64_t ctx, double* V_mag);

/* Batched field step over n independent contexts (the grains of a grain set, each its own mesh
 * and airbox): context i minimises its resident state at field B[3i..3i+2] and fills out[i].
 * Replaces a Python loop of minimize(m0, B_ext) calls over systems (src/minimizers.py:4053-4096);
 * n_threads host threads (<= 0: min(n, 8)) each drive whole contexts on the contexts' own streams,
 * so the small systems' kernels overlap on the GPU. Results are those of n sequential
 * mm_minimize_resident calls. The contexts must be distinct. */
int mm_minimize_batched(int n, const int64_t* ctxs, const mm_params* p, const double* B, int n_threads,
                        mm_info* out);
/* Batched solve_U over n independent contexts (src/poisson_solve.py:724-771 per system):
 * m[i] (N_i,3) and U_inout[i] (N_i,) host arrays, per-system tolerance tol[i]; it[i] and
 * rel_res[i] as in mm_solve_U. */
int mm_solve_U_batched(int n, const int64_t* ctxs, const double* const* m, double* const* U_inout, const double* tol,
                       int maxiter, int n_threads, int* it, double* rel_res);

/* Host-side aggregation for the smoothed-aggregation AMG set-up (the role of
 * pyamg.smoothed_aggregation_solver's aggregation step, src/poisson_solve.py:596-598):
 * S is the strength graph in CSR (n rows); agg[i] receives the aggregate of node i or -1 for
 * an isolated node; n_agg the number of aggregates. No device work. */
int mm_sa_aggregate(int n, const int32_t* Sp, const int32_t* Sj, int32_t* agg, int32_t* n_agg);

/* Host-side grouping of the tets into edge-star clusters, the unit of work of the Poisson
 * apply: up to 6 tets sharing one edge on 8 node slots [E0, E1, R0..R5], tet j = (E0, E1, Rj,
 * R(j+1) mod 6) when bit j of tets[i] is set; unused slots are -1. nodes8 (at least E x 8)
 * and tets (at least E) receive the clusters, n_clusters their count. No device work; no
 * reference counterpart (the reference applies the operator per element or as CSR,
 * src/poisson_solve.py:166-171,210-252). */
int mm_edge_clusters(int N, int E, const double* knt, const int32_t* conn, int32_t* nodes8, uint32_t* tets,
                     int32_t* n_clusters);

/* Live timing of the Poisson apply kernel inside the PCG: HIP events on the context's stream
 * bracket every 8th launch. Returns the accumulated milliseconds and count of the timed
 * launches since the last call, then resets them and enables (enable != 0) or disables the
 * timing. */
int mm_apply_timing(int64_t ctx, int enable, double* total_ms, int64_t* count);

/* Counters of kernel launches since creation: applies, rhs, energy, local, V-cycles, PCG iterations. */
int mm_counters(int64_t ctx, int64_t* out6);

/* Benchmark hook: `reps` back-to-back Poisson applies on device-resident vectors, timed with
 * HIP events on the context's stream; returns the mean milliseconds per apply. */
int mm_bench_apply(int64_t ctx, int geom, int reps, double* ms_per_apply);

/* Benchmark hook: `reps` back-to-back k-vector applies (with their x.Ax partials) on device-resident
 * vectors; mean milliseconds per k-vector apply. */
int mm_bench_apply_multi(int64_t ctx, int k, int reps, double* ms);

/* Benchmark hook: `reps` copies of N doubles between two context vectors with a plain copy kernel
 * (k_copy_calib, 16 N bytes per launch), the known-size reference for HBM byte counters; mean ms. */
int mm_bench_copy(int64_t ctx, int reps, double* ms_per_copy);

/* Benchmark hook: `reps` preconditioner applications on device-resident vectors, launched
 * directly (use_graph = 0) or replayed from a captured hipGraph (use_graph = 1); mean ms. */
int mm_bench_precond(int64_t ctx, int use_graph, int reps, double* ms_per_apply);

#ifdef __cplusplus
}
#endif

#endif /* MUMAG_HIP_H */
