// Internal declarations of the MI355X micromagnetics library (not part of the C-ABI).
//
// Data layout in HBM (one context = one mesh on one GPU):
//   * nodes keep the caller's (Morton) order; nodal vectors are N doubles (potential) or
//     3N interleaved doubles (m, g, ...); node coordinates are 3N interleaved doubles;
//   * elements are re-sorted by (smallest node id, largest node id) and cut greedily into
//     blocks of consecutive elements touching at most 256 distinct nodes; one workgroup
//     owns one block;
//   * per block: the sorted list of its distinct nodes (`nodes`, owner/private/boundary
//     flags in the top three bits) and, per element, the four corner indices into that
//     list packed as 4 x uint8 (`lconn`);
//   * two partitions: all elements (Poisson apply) and magnetic elements only (Js, A or
//     K1 nonzero: right-hand side, energy, local Hessian), each with its own material
//     index stream;
//   * optional stored per-element geometry (gradients of corners 1..3, 9 x E, and Ve) in
//     FP64 and FP32, structure-of-arrays, for the streamed-geometry apply variant.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace mm {

constexpr int kBlock = 256;              // threads per workgroup
constexpr int kMaxBlockNodes = 256;      // distinct nodes per element block (uint8 indices)
constexpr int kMaxBlockElems = 2048;     // elements per block cap
constexpr int kMaxClusterBlockNodes = 2048;  // distinct nodes per cluster block (11-bit slot indices)
constexpr uint32_t kOwner = 1u << 31;    // this block adds reg*x / writes the boundary row / Zeeman
constexpr uint32_t kPrivate = 1u << 30;  // no other block touches the node: plain store
constexpr uint32_t kBoundary = 1u << 29; // Dirichlet node (outer surface)
constexpr uint32_t kIdMask = (1u << 29) - 1;
constexpr int kMaxLbfgsMemory = 32;       // L-BFGS history pairs (device scalar slots)
constexpr int kMaxMat = 255;               // materials (uint8 index per element)
constexpr int kRedBlocks = 2048;         // max workgroups of a reducing kernel
constexpr int kNumScal = 256;            // device scalar slots
constexpr int kSegStageSlots = 4;        // pinned staging slots of the segment tables

enum Geom { kGeomCoords = 0, kGeomStored64 = 1, kGeomStored32 = 2 };

void set_error(const std::string& msg);

#define MM_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess) {                                                               \
      mm::set_error(std::string(#call) + " failed: " + hipGetErrorString(_e) + " at " + \
                    __FILE__ + ":" + std::to_string(__LINE__));                           \
      return -2;                                                                          \
    }                                                                                     \
  } while (0)

#define MM_CHECK(cond, msg) \
  do {                      \
    if (!(cond)) {          \
      mm::set_error(msg);   \
      return -1;            \
    }                       \
  } while (0)

#define MM_TRY(call)        \
  do {                      \
    int _r = (call);        \
    if (_r != 0) return _r; \
  } while (0)

struct MatTable {
  double A[kMaxMat];
  double K1[kMaxMat];
  double Js[kMaxMat];
  double k[kMaxMat][3];
};

struct Part {
  int nblk = 0;
  int E = 0;                          // elements in this partition
  int Es = 0;                         // padded stride of stored-geometry streams
  const int* eptr = nullptr;          // nblk + 1 element offsets
  const int* nptr = nullptr;          // nblk + 1 node-list offsets
  const uint32_t* nodes = nullptr;    // node ids | flags
  const uint32_t* lconn = nullptr;    // E packed local corners
  const uint8_t* mat = nullptr;       // E material index (0-based)
  // edge-star cluster partition only: eptr are cluster offsets, lconn / lconn2 the low 8 bits
  // of the local indices of slots 0-3 / 4-7, cmask bits 0-5 the tets and bits 8 + 3q .. 10 + 3q
  // the high bits of slot q's index (blocks of up to 2048 nodes); the used slots follow from
  // the tets (cluster_slots)
  const uint32_t* lconn2 = nullptr;
  const uint32_t* cmask = nullptr;
  // cluster partition only, one record per node-list entry in block order for the FP32
  // kernels: x, y, z relative to the block's first node, id | flags as the bits of w
  const float4* rec32 = nullptr;
  // optional subset of the blocks (the blocks of some segments of a segmented context):
  // workgroup position i < nlist works on block blist[i]; null: every block
  const int* blist = nullptr;
  int nlist = 0;
  // colour-sequenced launch (cluster apply): no two blocks of the launch share a node, so a
  // shared node's block sum is added with a plain load and store instead of an FP64 atomic;
  // the launches of the colours follow each other on one stream
  int seq = 0;
};

// Workgroup tables of a segmented vector kernel over a subset of the segments: workgroup w
// covers chunk wg_chunk[w] (kSegChunk nodes) of segment wg_seg[w]; the workgroups of segment g
// are wg_off[g] .. wg_off[g + 1] - 1 (empty when g is not in the subset).
struct WgTables {
  int n = 0;                          // workgroups
  int *seg = nullptr, *chunk = nullptr, *off = nullptr;  // device
  std::vector<int> h_off;             // host copy of off
  std::vector<int> key;               // segment mask the tables were built for
};

struct Geo {
  const double* knt = nullptr;        // 3N
  const double* grad = nullptr;       // 9 x Es (part_all order)
  const double* Ve = nullptr;         // Es
  const float* grad32 = nullptr;
  const float* Ve32 = nullptr;
  const MatTable* mt = nullptr;
  const double* M_nodal = nullptr;    // N
  // N: 1 / V_mag of the node's segment (contexts of several independent systems); null: the
  // scalar 1 / V_mag passed to the kernel
  const double* inv_V_node = nullptr;
};

struct BlkInfo {
  int b;                              // logical block id (after the XCD remap)
  int e0, e1, n0, nb;
  bool valid;
};

using Runs = std::vector<std::pair<int, int>>;

// One AMG level; level 0 uses the matrix-free apply (A arrays unused there).
struct Level {
  int n = 0, n_next = 0;
  int *Ap = nullptr, *Ai = nullptr;
  double* Ax = nullptr;
  int *Pp = nullptr, *Pi = nullptr;
  double* Px = nullptr;
  int *Rp = nullptr, *Ri = nullptr;
  double* Rx = nullptr;
  double* M = nullptr;                // smoother multiplier: x += M * (b - A x)
  // coarsest level: the dense inverse as the lower-triangle 64 x 64 tiles of a symmetric matrix
  // (tile (I, J), J <= I, at index I (I + 1) / 2 + J, row-major inside), FP64 and FP32, and the
  // per-tile-pair partial products of the symmetric multiply
  int tiles = 0;                      // tile rows: ceil(n / 64)
  double* dense_inv = nullptr;
  float* dense_invf_t = nullptr;
  double* symv_part = nullptr;        // tiles x tiles x 64
  int lanesA = 8, lanesP = 8, lanesR = 8;  // SpMV lanes per row
  double *b = nullptr, *x = nullptr, *t = nullptr, *xs = nullptr;  // xs: post-smoothed result
  // FP32 copies for the mixed-precision V-cycle
  float *Axf = nullptr, *Pxf = nullptr, *Rxf = nullptr, *Mf = nullptr;
  // FP16 copies of the matrix values for the FP32 PCG's V-cycle (products and sums in FP32),
  // each matrix scaled by a power of two that puts its largest magnitude in [2^13, 2^14); the
  // *_s fields hold the inverse scales, applied to each row sum. Null where a value is not finite.
  __half *Axh = nullptr, *Pxh = nullptr, *Rxh = nullptr;
  float Axh_s = 1.0f, Pxh_s = 1.0f, Rxh_s = 1.0f;
  float *bf = nullptr, *xf = nullptr, *tf = nullptr, *xsf = nullptr;
  // segmented contexts: the rows of segment g are seg_off[g] .. seg_off[g + 1] - 1 (empty when
  // the level's unknowns are not ordered by segment); runs: the row ranges of the segments a
  // segmented solve still iterates (empty: every row)
  std::vector<int> seg_off;
  Runs runs;
};

struct Ctx {
  int device = 0;
  int N = 0, E = 0, G = 0;
  double V_mag = 1.0, reg = 1e-12;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  Part part_all, part_mag;
  Part part_cl;                       // edge-star clusters of all elements (Poisson apply)
  long long nodes_cl = 0;             // node-list entries of the cluster partition
  bool apply_clusters = true;         // Poisson apply over clusters (else per element)
  int apply_persist = 0;              // > 0: FP64 cluster apply as a persistent grid, this many workgroups per CU
  int num_cu = 256;                   // compute units of the device
  Geo geo;
  int n_mag = 0;                      // magnetic elements
  long long nodes_all = 0;            // node-list entries of the all-element partition
  int max_block_nodes = 0;
  int geom_mode = kGeomCoords;        // geometry source of the Poisson apply
  int elem_threads = 256;             // workgroup size of the Poisson apply (64, 128, 256)
  int node_cap = kMaxBlockNodes;      // distinct nodes per element block
  int cl_cap = kMaxBlockNodes;        // distinct nodes per cluster block (256, 512, 1024 or 2048)
  bool cl_records = true;             // FP32 applies on 512..2048-node cluster blocks stage entry records
  // colour-sequenced scatter of the cluster apply (instead of FP64 atomics on shared nodes): the
  // cluster blocks grouped by colour (no two blocks of a colour share a node), one launch each
  bool apply_seq = false;
  int* d_colour_blk = nullptr;        // part_cl.nblk block ids, colour by colour, block order inside
  std::vector<int> h_colour_off;      // colours + 1 offsets into d_colour_blk
  std::vector<int> h_cl_nodes;        // host copy of the clusters: 8 slots each (-1 unused)
  std::vector<uint32_t> h_cl_tets;    // tet mask of each cluster
  std::vector<char> h_is_bnd;         // Dirichlet flag per node
  std::vector<void*> allocs;          // every device allocation, freed at destroy
  MatTable hmt;

  double* knt = nullptr;
  double* M_nodal = nullptr;          // N
  double* inv_M_rel = nullptr;        // N
  double* inv_M_prec = nullptr;       // N
  std::vector<int> h_conn;            // E x 4 (part_all order)
  std::vector<double> h_Ve, h_grad;   // E, E x 9
  std::vector<double> h_M_nodal, h_inv_M_rel, h_inv_M_prec, h_exdiag;

  // reductions / scalars
  double* partial = nullptr;          // kRedBlocks x 4
  double* dscal = nullptr;            // kNumScal device scalars
  double* hscal = nullptr;            // pinned host mirror
  // fence-free PCG reductions: per-workgroup partials consumed by the next kernel or the host
  double* part_pq = nullptr;          // p.Ap partials of the Poisson apply (one per apply workgroup)
  double* part_rr = nullptr;          // r.r (and b.b) partials, 2 x kRedGridMax
  double* part_rz = nullptr;          // r.z partials, kRedGridMax
  double* hpart = nullptr;            // host-coherent pinned r.r (and b.b) partials of the PCG
  double* hpart_dev = nullptr;        // its device address

  // Poisson / PCG vectors (N)
  double *pb = nullptr, *pr = nullptr, *pz = nullptr, *pp = nullptr, *pq = nullptr, *pt = nullptr, *px0 = nullptr;
  double* work_U = nullptr;
  // AMG
  int amg_kind = 0;                   // 0: Jacobi on diag(A) only, 1: SPAI0 V-cycle, 2: damped-Jacobi V-cycle
  std::vector<Level> lv;
  std::vector<std::vector<int>> amg_seg_off;  // per level: rows of each segment (set-up)
  double* diagA = nullptr;            // N
  int precision = 0;                  // 0: FP64; 1: FP32 V-cycle; 2: FP32 PCG with FP64 refinement
  bool lv_half = false;               // every level holds its FP16 matrix values (the FP32 PCG uses them)
  float *px0f = nullptr, *ptf = nullptr;  // level-0 FP32 work vectors
  // FP32 PCG with FP64 refinement (precision 2): correction, residual, direction, A p, V-cycle output
  float *rxf = nullptr, *rrf = nullptr, *rpf = nullptr, *rqf = nullptr, *rzf = nullptr;
  int ref_cycles = 0;                 // residual replacements of the last solve
  double ref_eta = 1e-3;              // FP32 PCG: replace the residual after this reduction of its norm

  // minimiser state (3N unless noted)
  double *m = nullptr, *U = nullptr /*N*/, *g_raw = nullptr, *g_tan = nullptr, *g_ext = nullptr, *g_prev = nullptr;
  double *y = nullptr, *d = nullptr, *d_prev = nullptr, *H = nullptr;
  double *m_t = nullptr, *g_t = nullptr, *U_t = nullptr /*N*/;
  double *r3 = nullptr, *z3 = nullptr, *p3 = nullptr, *q3 = nullptr, *cv3 = nullptr;
  // the Steihaug loop's r, z, p, q and C p in FP32 (the mixed-precision forms; allocated with them)
  float *r3f = nullptr, *z3f = nullptr, *p3f = nullptr, *q3f = nullptr, *cv3f = nullptr;
  int mem = 10;
  double *S = nullptr, *Yh = nullptr;  // mem x 3N (FP64 history)
  float *Sf = nullptr, *Yf = nullptr;  // mem x 3N (FP32 history of the mixed-precision forms)
  double *mx = nullptr, *my = nullptr;  // 8N each: node-major vectors of the multi-vector apply
  double* mdot = nullptr;               // 8 x (cluster blocks at 256 nodes): its x.Ax partials
  int mdot_n = 0;
  // per-workgroup partials of the field-step reductions (9 banks of 2 x kRedGridMax) and a pinned
  // host copy of the first banks
  double *mpart = nullptr, *hmpart = nullptr;
  bool have_state = false;

  // independent systems (segments) sharing the context: contiguous node ranges; no element
  // couples two segments and no element block spans two
  int nseg = 1;
  std::vector<int> h_seg_node_off;    // nseg + 1
  std::vector<int> h_seg_blk_all;     // nseg + 1 block offsets of part_all
  std::vector<int> h_seg_blk_cl;      // nseg + 1 block offsets of part_cl
  std::vector<int> h_seg_blk_mag;     // nseg + 1 block offsets of part_mag
  std::vector<double> h_seg_vmag;     // magnetic volume per segment
  // segmented PCG work decomposition (set up on first use): workgroup w handles nodes of
  // segment wg_seg[w] only, kSegChunk at a time
  int seg_nwg = 0;
  int *d_seg_node_off = nullptr, *d_seg_blk_all = nullptr, *d_seg_blk_cl = nullptr;
  int* d_seg_active = nullptr;
  // The tables of a segment subset (d_seg_active, the workgroup tables wg_pcg / wg_min, the three
  // block lists) are sub-arrays of one device buffer seg_tab with a host master copy. An update
  // writes the master and widens the dirty range; a flush copies that range through a ring of
  // pinned staging slots, a slot being reused only after its previous copy has completed. While
  // seg_defer is set, updates only mark the range, so several updates travel in one copy.
  int* seg_tab = nullptr;
  int seg_tab_n = 0;
  std::vector<int> seg_master;
  int* seg_stage = nullptr;                       // kSegStageSlots x seg_tab_n, pinned
  hipEvent_t seg_stage_ev[kSegStageSlots] = {};
  bool seg_stage_rec[kSegStageSlots] = {};
  int seg_stage_next = 0;
  long seg_dirty_lo = -1, seg_dirty_hi = -1;
  bool seg_defer = false;
  double* d_seg_rz = nullptr;         // 2 x nseg r.z ping-pong
  double* seg_part_rz = nullptr;      // seg_nwg r.z partials
  double *hseg = nullptr, *hseg_dev = nullptr;  // host-coherent r.r / b.b partials, 2 x seg_nwg
  std::vector<int> h_seg_wg_off, h_seg_active;
  bool seg_running = false;           // the preconditioner ends with per-segment r.z partials
  int* d_seg_blk_mag = nullptr;
  // block subsets of the three partitions (set_block_lists) and the segment mask they hold
  int *d_blist_cl = nullptr, *d_blist_all = nullptr, *d_blist_mag = nullptr;
  std::vector<int> h_blist_cl, h_blist_all, h_blist_mag, blist_key;
  WgTables wg_pcg, wg_min;            // vector-kernel subsets of the segmented PCG / minimiser
  double* inv_V_node = nullptr;       // N (segmented contexts only)
  // segmented minimiser: per-segment scalar table, host-set inputs, partials
  double* sm_tab = nullptr;           // nseg x kSegSlots
  double* sm_in = nullptr;            // nseg x kSegIn
  double *sm_pa = nullptr, *sm_pb = nullptr;  // seg_nwg x 4 partials each

  // initial guess of the minimiser's Poisson solves: the combination of the last guess_k
  // solutions whose right-hand sides best match the new one (0: the previous solution as
  // given). Ring of solutions U_i and right-hand sides b_i; G[i][j] = b_i . b_j on the host.
  int guess_k = 0, guess_n = 0, guess_head = 0, guess_new = -1;
  double *guess_U = nullptr, *guess_b = nullptr;  // kGuessMax x N each
  double* guess_x = nullptr;                      // N: the candidate guess
  std::vector<double> guess_G;                    // kGuessMax x kGuessMax
  double *hguess = nullptr, *hguess_dev = nullptr;  // host-coherent partials

  long long n_apply = 0, n_rhs = 0, n_energy = 0, n_local = 0, n_vcycle = 0, n_pcg_it = 0;

  // live timing of the Poisson apply kernel: event pairs recorded on `stream` around each
  // launch inside the PCG, resolved at the PCG's own synchronisation points
  bool time_apply = false;
  std::vector<hipEvent_t> ev_pool;
  int ev_used = 0;
  double apply_ms = 0.0;
  long long apply_timed = 0;
  long long apply_seen = 0;           // applies issued while timing is on (every 8th is timed)
};

int timed_apply(Ctx& c, const double* x, double* y, hipStream_t s, double* dot_part = nullptr);
int resolve_apply_timing(Ctx& c);

// element kernels (mm_kernels.hip)
// dot_part != nullptr: also write per-block partials of x . (A x) (apply_dot_count(c) of them)
int launch_apply(const Ctx& c, int geom, const double* x, double* y, hipStream_t s, double* dot_part = nullptr);
// number of x . (A x) partials the on-the-fly apply writes
int apply_dot_count(const Ctx& c);
// greedy colouring of the cluster blocks (blocks sharing a node get different colours), uploaded
// to d_colour_blk / h_colour_off; called when the colour-sequenced scatter is switched on and
// whenever the cluster partition is rebuilt while it is on
int build_block_colours(Ctx& c);
int launch_apply_f32(const Ctx& c, const float* x, float* y, hipStream_t s);
// y = A (w .* r) with xo = w .* r written by the node owners (fused pre-smoothing); y zeroed first
int launch_apply_scaled(const Ctx& c, const double* r, const double* w, double* xo, double* y, hipStream_t s);
int launch_apply_f32_scaled(const Ctx& c, const double* r, const float* w, float* xo, float* y, hipStream_t s);
int launch_apply_f32_scaled_f(const Ctx& c, const float* r, const float* w, float* xo, float* y, hipStream_t s);
int launch_apply_f32_dot(const Ctx& c, const float* x, float* y, double* dot_part, hipStream_t s);
// y_r = A x_r for K in {1, 2, 4, 8} node-major vectors (x[K id + r]); y zeroed first; dot_part
// (nullable) receives K partials per cluster block
int launch_apply_multi(const Ctx& c, int K, const double* x, double* y, double* dot_part, hipStream_t s);
int launch_rhs(const Ctx& c, const double* m, double* b, hipStream_t s);
// launch_energy with the block partials (2 doubles per magnetic block) in a caller-given buffer
int launch_energy_buf(const Ctx& c, const double* m, const double* U, const double B[3], double* g, double* E_out,
                      double* partial, hipStream_t s);
int launch_energy(const Ctx& c, const double* m, const double* U, const double B[3], double* g, double* E_out,
                  hipStream_t s);
int launch_local(const Ctx& c, const double* v, double* out, const double* gate, hipStream_t s);
// the same on FP32 vectors (the mixed-precision forms' Steihaug loop)
int launch_local(const Ctx& c, const float* v, float* out, const double* gate, hipStream_t s);
int launch_geometry(const Ctx& c, const int* conn_dev, hipStream_t s);
int launch_diag(const Ctx& c, double* d, hipStream_t s);
int launch_jpar(const Ctx& c, const double* m, const double h[3], double* out, hipStream_t s);

// vector kernels (mm_vec.hip)
int vfill(double* x, double v, long n, hipStream_t s);
int vfill_f(float* x, float v, long n, hipStream_t s);
int vcopy(double* dst, const double* src, long n, hipStream_t s);
int launch_copy_calib(double* dst, const double* src, long n, hipStream_t s);
// out[0] = a.b (deterministic)

// solver (mm_solver.hip)
// z = M^-1 r; rz_part receives red_grid(N) partials of r.z
int precond_apply(Ctx& c, const double* r, double* z, double* rz_part, hipStream_t s);
int precond(Ctx& c, const double* r, double* z, double* rz_part, hipStream_t s);
int solve_U_dev(Ctx& c, const double* m, double* U, double tol, int maxiter, int* it, double* rel_res);
// solve_U_dev whose initial guess is the projection onto the last guess_k solutions (when
// guess_k > 0; U holds the fallback guess), and which adds its solution to them
constexpr int kGuessMax = 8;
int solve_U_guess(Ctx& c, const double* m, double* U, double tol, int maxiter, int* it, double* rel_res);
int set_solve_guess(Ctx& c, int k);
// solve_U for every segment of the context at once (one launch per kernel for all segments,
// per-segment PCG scalars and stopping test): tol, it, rel_res have one entry per segment
int solve_U_seg(Ctx& c, const double* m, double* U, const double* tol, int maxiter, int* it, double* rel_res);
// device-resident form: b = rhs(m) already in c.pb is not assumed; only segments with want[g] != 0
// iterate (the others keep U); it / rel_res per segment
int solve_U_seg_dev(Ctx& c, const double* m, double* U, const double* tol, const int* want, int maxiter, int* it,
                    double* rel_res);
// segmented decomposition tables (once per context)
int ensure_seg(Ctx& c);
// restrict the element kernels to the blocks of the segments with mask[g] != 0 (null: all)
int set_block_lists(Ctx& c, const int* mask);
// workgroup tables of the segments with mask[g] != 0 (null: all)
int build_wg_tables(Ctx& c, const int* mask, WgTables& T);
// n ints of src to the sub-array dev of seg_tab (through the master copy; flushed now unless
// seg_defer is set); seg_flush copies the dirty range
int seg_upload(Ctx& c, int* dev, const int* src, int n);
int seg_flush(Ctx& c);
// defers the flushes of the segment-table updates made while it lives; restores the previous
// setting on every path (batches nest), after which the owner calls seg_flush
struct SegBatch {
  Ctx& c;
  bool prev;
  explicit SegBatch(Ctx& c_) : c(c_), prev(c_.seg_defer) { c.seg_defer = true; }
  ~SegBatch() { c.seg_defer = prev; }
};
// per-segment sums of the magnetic-partition block partials: out[g * stride + k] for k < K
// (part holds K values per logical block; compensated pairs when CS)
int launch_seg_block_sums(const Ctx& c, const double* part, int K, bool cs, double* out, int stride, int slot,
                          hipStream_t s);

}  // namespace mm
