"""``MeshContext``: one mesh + material set resident on the GPU, behind the C-ABI.

It owns the device context (element-block partitions, node coordinates, nodal
scalings, AMG hierarchy, minimiser state) and exposes thin NumPy-in / NumPy-out
wrappers. The reference-shaped factories (``make_poisson_ops``, ``make_solve_U``,
``make_energy_kernels``, ``make_minimizer``) are built on it.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _hip
from ._hip import call, f64, i32, ptr

GEOM_MODES = {"on_the_fly": 0, "stored_f64": 1, "stored_f32": 2, "on_the_fly_f32": 3}
METHODS = {"pcohen_hs": 0, "lbfgs": 1}
APPLY_UNITS = {"element": 0, "cluster": 1, "persistent": 2}
CLUSTER_STAGING = {"gather": 0, "records": 1}
PRECISIONS = {"fp64": 0, "mixed": 1, "fp32pcg": 2, "fp32all": 3}


class MeshContext:
    """Device-resident mesh, materials and solver state.

    Args:
        knt: node coordinates (N, 3) in mesh units.
        conn: connectivity (E, 4), 0-based.
        mat_id: material ids (E,), 1-based.
        A_red, K1_red, Js_red: reduced material constants (G,).
        k_easy: easy axes (G, 3).
        boundary_mask: (N,) 1.0 interior / 0.0 Dirichlet, or None.
        reg: Poisson diagonal regularisation.
        geom: geometry source of the Poisson apply ('on_the_fly', 'stored_f64', 'stored_f32').
        block_nodes: distinct nodes per element block (16..256), the LDS staging size.
        segments: optional node offsets (nseg + 1,) of independent systems sharing the mesh
            (``mm_ctx_create_segmented``); enables ``solve_U_segmented``.
    """

    def __init__(self, knt, conn, mat_id, A_red, K1_red, Js_red, k_easy, boundary_mask=None, reg=1e-12,
                 geom="on_the_fly", block_nodes: int = 256, segments=None):
        self.knt = f64(knt).reshape(-1, 3)
        conn = i32(np.asarray(conn)[:, :4])
        mat_id = i32(mat_id)
        self.N = self.knt.shape[0]
        self.E = conn.shape[0]
        A_red, K1_red, Js_red = f64(A_red), f64(K1_red), f64(Js_red)
        k_easy = f64(np.asarray(k_easy, dtype=np.float64).reshape(-1, 3))
        G = A_red.size
        if not (K1_red.size == G and Js_red.size == G and k_easy.shape[0] == G):
            raise ValueError("material tables must all have G entries")
        if conn.size and (conn.min() < 0 or conn.max() >= self.N):
            raise ValueError("connectivity out of range")
        bnd = None if boundary_mask is None else f64(boundary_mask)
        if bnd is not None and bnd.size != self.N:
            raise ValueError("boundary_mask must have N entries")
        self.boundary_mask = bnd
        self.reg = float(reg)
        self.Js_red = Js_red
        self._conn = conn
        self._mat_id = mat_id
        h = C.c_int64(0)
        if not 16 <= int(block_nodes) <= 256:
            raise ValueError("block_nodes must be in [16, 256]")
        flags = GEOM_MODES[geom] | (int(block_nodes) % 256) << 8
        self.segments = None if segments is None else i32(segments)
        if self.segments is None:
            call("mm_ctx_create", self.N, self.E, ptr(self.knt), ptr(conn), ptr(mat_id), G, ptr(A_red), ptr(K1_red),
                 ptr(Js_red), ptr(k_easy), ptr(bnd), self.reg, flags, C.byref(h))
        else:
            call("mm_ctx_create_segmented", self.N, self.E, ptr(self.knt), ptr(conn), ptr(mat_id), G, ptr(A_red),
                 ptr(K1_red), ptr(Js_red), ptr(k_easy), ptr(bnd), self.reg, flags, self.segments.size - 1,
                 ptr(self.segments), C.byref(h))
        self.handle = h.value
        info = np.zeros(15)
        call("mm_ctx_info", self.handle, ptr(info))
        self.V_mag = float(info[0])
        self.n_blocks_all = int(info[3])
        self.n_blocks_mag = int(info[4])
        self.max_block_nodes = int(info[5])
        self.n_mag_elems = int(info[6])
        self.node_list_entries = int(info[7])
        self.n_blocks_cl = int(info[8])
        self.cluster_node_list_entries = int(info[9])
        self.n_clusters = int(info[10])
        self.block_nodes = int(info[11])
        self.cluster_block_nodes = int(info[12])
        self.cluster_staging = "records" if info[13] else "gather"
        self.apply_scatter = "atomic"
        self.n_block_colours = 0
        self.apply_shape = {"unit": "cluster", "param": 256}
        self.solve_guess = 0
        self.precond_type = None
        self.amg_levels = None
        self.precision = "fp64"

    # ------------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "handle", 0):
            _hip.load().mm_ctx_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ set-up data
    def geometry(self):
        """(conn (E,4), Ve (E,), grad (E,3,3)) in the context's element order."""
        conn = np.zeros((self.E, 4), dtype=np.int32)
        Ve = np.zeros(self.E)
        grad = np.zeros((self.E, 3, 3))
        call("mm_ctx_get_geometry", self.handle, ptr(conn), ptr(Ve), ptr(grad))
        return conn, Ve, grad

    def element_table(self):
        """(conn (E,4), Ve (E,), magnetic (E,) bool) in the caller's element order, on the host:
        Ve = |det J| / 6 from the coordinates, magnetic = the element's material has Js > 0."""
        x = self.knt[self._conn]
        d = np.linalg.det(x[:, 1:, :] - x[:, :1, :])
        return self._conn, np.abs(d) / 6.0, self.Js_red[self._mat_id - 1] > 0

    def nodal(self):
        """(M_nodal, inv_M_rel, inv_M_prec), each (N,)."""
        M = np.zeros(self.N)
        a = np.zeros(self.N)
        b = np.zeros(self.N)
        call("mm_ctx_get_nodal", self.handle, ptr(M), ptr(a), ptr(b))
        return M, a, b

    def exchange_diagonal(self):
        """diag(K_ex)/V_mag (N,)."""
        d = np.zeros(self.N)
        call("mm_ctx_get_exchange_diag", self.handle, ptr(d))
        return d

    def set_precision(self, precision: str):
        """'fp64'; 'mixed' (FP32 V-cycle under an FP64 PCG); 'fp32pcg' (the PCG itself on FP32
        vectors with the FP32 apply and V-cycle, FP64 dots and alpha/beta, and an FP64
        true-residual refinement loop to the same tolerance); 'fp32all' ('fp32pcg' with FP32
        energy densities). The mixed forms also run the Steihaug Hessian action in FP32 and keep
        the L-BFGS history in FP32."""
        call("mm_set_precision", self.handle, PRECISIONS[precision])
        self.precision = precision

    def set_apply_shape(self, unit: str = "cluster", param: int = 256):
        """Work shape of the on-the-fly Poisson apply: ``unit`` 'element' (one tet per thread) or
        'cluster' (one edge-star cluster per thread) with ``param`` = workgroup size (64, 128,
        256), or 'persistent' (clusters on a persistent grid that prefetches its next block) with
        ``param`` = workgroups per CU (1..4)."""
        call("mm_set_apply_shape", self.handle, APPLY_UNITS[unit], int(param))
        self.apply_shape = {"unit": unit, "param": int(param)}

    def set_cluster_block_nodes(self, nb: int):
        """Distinct nodes per cluster block of the Poisson apply (256, 512, 1024 or 2048);
        rebuilds the cluster partition and refreshes the partition counts."""
        call("mm_set_cluster_block_nodes", self.handle, int(nb))
        self._refresh_cluster_info()
        if self.cluster_block_nodes != 256 and self.apply_shape["unit"] == "persistent":
            self.apply_shape = {"unit": "cluster", "param": 256}
        if self.cluster_block_nodes != 256 and self.apply_shape["unit"] == "cluster":
            self.apply_shape = {"unit": "cluster", "param": 256}

    def set_apply_scatter(self, mode: str):
        """How the cluster apply adds the sums of nodes that several blocks touch: "atomic" (the
        default: hardware FP64 atomics, one launch) or "colour" (blocks coloured so that no two
        blocks of a colour share a node, one launch per colour, plain loads and stores). Same
        operator to rounding."""
        call("mm_set_apply_scatter", self.handle, {"atomic": 0, "colour": 1}[mode])
        self.apply_scatter = mode
        self._refresh_cluster_info()

    def set_refinement(self, eta: float):
        """FP32 PCG ('fp32pcg', 'fp32all'): replace the recursive FP32 residual by the true FP64
        residual each time its norm has dropped by ``eta`` since the last replacement."""
        call("mm_set_refinement", self.handle, float(eta))

    def _refresh_cluster_info(self):
        info = np.zeros(15)
        call("mm_ctx_info", self.handle, ptr(info))
        self.n_blocks_cl = int(info[8])
        self.cluster_node_list_entries = int(info[9])
        self.cluster_block_nodes = int(info[12])
        self.n_block_colours = int(info[14])

    def set_cluster_staging(self, mode: str):
        """How the FP32 applies on cluster blocks of 512 to 2048 nodes stage their nodes: "gather"
        (node ids, then the FP64 coordinates gathered by id) or "records" (the default: per-entry
        block-relative FP32 coordinate records streamed in block order). The staged values, and
        so the results, are the same; the FP64 apply always gathers."""
        call("mm_set_cluster_staging", self.handle, CLUSTER_STAGING[mode])
        self.cluster_staging = mode

    def set_geometry_mode(self, geom: str):
        call("mm_set_geometry_mode", self.handle, GEOM_MODES[geom])

    # ------------------------------------------------------------------ operators
    def apply_A(self, x):
        x = f64(x)
        y = np.empty(self.N)
        call("mm_poisson_apply", self.handle, ptr(x), ptr(y))
        return y

    def apply_A_multi(self, X):
        """Y = A X for k in {1, 2, 4, 8} vectors X (k, N) in one multi-vector apply."""
        X = f64(X)
        Y = np.empty_like(X)
        call("mm_poisson_apply_multi", self.handle, X.shape[0], ptr(X), ptr(Y))
        return Y

    def bench_apply_multi(self, k, reps=40):
        """Mean ms of a k-vector apply (device-resident vectors)."""
        ms = C.c_double(0.0)
        call("mm_bench_apply_multi", self.handle, int(k), int(reps), C.byref(ms))
        return float(ms.value)

    def rhs(self, m):
        m = f64(m).reshape(-1)
        b = np.empty(self.N)
        call("mm_poisson_rhs", self.handle, ptr(m), ptr(b))
        return b

    def diag(self):
        d = np.empty(self.N)
        call("mm_poisson_diag", self.handle, ptr(d))
        return d

    def energy_grad(self, m, U, B):
        m = f64(m).reshape(-1)
        U = f64(U)
        B = f64(B).reshape(3)
        E = np.zeros(1)
        g = np.empty(3 * self.N)
        call("mm_energy_grad", self.handle, ptr(m), ptr(U), ptr(B), ptr(E), ptr(g))
        return float(E[0]), g.reshape(-1, 3)

    def local_grad(self, v):
        v = f64(v).reshape(-1)
        g = np.empty(3 * self.N)
        call("mm_local_grad", self.handle, ptr(v), ptr(g))
        return g.reshape(-1, 3)

    def j_parallel(self, m, h):
        m = f64(m).reshape(-1)
        h = f64(h).reshape(3)
        J = np.zeros(1)
        call("mm_j_parallel", self.handle, ptr(m), ptr(h), ptr(J))
        return float(J[0])

    # ------------------------------------------------------------------ preconditioner
    def setup_precond(self, precond_type="amgcl", log=None, **amg_kwargs):
        """'jacobi', 'amgcl' (SPAI0 V-cycle) or 'amg' (damped-Jacobi V-cycle) on a host SA hierarchy.
        ``log`` (a print-like callable) receives one line per set-up stage."""
        if precond_type in ("jacobi", "none"):
            call("mm_precond_jacobi", self.handle)
            self.precond_type = "jacobi"
            return
        from .amg import build_hierarchy, fine_matrix, upload_hierarchy

        A = fine_matrix(self)
        if log is not None:
            log(f"[amg] fine matrix: {A.shape[0]} rows, {A.nnz} nonzeros")
        levels = build_hierarchy(A, smoother=precond_type, log=log, **amg_kwargs)
        upload_hierarchy(self, levels, precond_type)
        self.precond_type = precond_type
        self.amg_levels = levels

    def precond_apply(self, r):
        r = f64(r)
        z = np.empty(self.N)
        call("mm_precond_apply", self.handle, ptr(r), ptr(z))
        return z

    def pcg(self, b, x0, tol, maxiter=2000):
        """PCG on A x = b: (x, iterations, r.r)."""
        b = f64(b)
        x = f64(np.array(x0, dtype=np.float64, copy=True))
        it = C.c_int32(0)
        r2 = C.c_double(0.0)
        call("mm_pcg_solve", self.handle, ptr(b), ptr(x), float(tol), int(maxiter), C.byref(it), C.byref(r2))
        return x, int(it.value), float(r2.value)

    def solve_U(self, m, x0, tol, maxiter=2000):
        """(U, iterations, relative residual)."""
        m = f64(m).reshape(-1)
        U = f64(np.array(x0, dtype=np.float64, copy=True))
        it = C.c_int32(0)
        rr = C.c_double(0.0)
        call("mm_solve_U", self.handle, ptr(m), ptr(U), float(tol), int(maxiter), C.byref(it), C.byref(rr))
        return U, int(it.value), float(rr.value)

    def solve_U_segmented(self, m, x0, tol, maxiter=2000):
        """solve_U for every segment at once (one launch per kernel for all segments, per-segment
        PCG scalars): (U, iterations per segment, relative residual per segment). ``tol`` is a
        scalar or one value per segment."""
        if self.segments is None:
            raise RuntimeError("solve_U_segmented needs a context created with segments")
        nseg = self.segments.size - 1
        m = f64(m).reshape(-1)
        U = f64(np.array(x0, dtype=np.float64, copy=True))
        tol = f64(np.broadcast_to(np.asarray(tol, dtype=np.float64), (nseg,)))
        it = np.zeros(nseg, dtype=np.int32)
        rr = np.zeros(nseg)
        call("mm_solve_U_segmented", self.handle, ptr(m), ptr(U), ptr(tol), int(maxiter), ptr(it), ptr(rr))
        return U, it, rr

    # ------------------------------------------------------------------ minimiser
    def state_set(self, m, U=None):
        m = f64(m).reshape(-1)
        call("mm_state_set", self.handle, ptr(m), ptr(None if U is None else f64(U)))

    def state_get(self):
        m = np.empty(3 * self.N)
        U = np.empty(self.N)
        call("mm_state_get", self.handle, ptr(m), ptr(U))
        return m.reshape(-1, 3), U

    def minimize_resident(self, params: "_hip.MMParams", B):
        B = f64(B).reshape(3)
        info = _hip.MMInfo()
        call("mm_minimize_resident", self.handle, C.byref(params), ptr(B), C.byref(info))
        return _info_dict(info)

    def state_j_parallel(self, h):
        h = f64(h).reshape(3)
        J = np.zeros(1)
        call("mm_state_j_parallel", self.handle, ptr(h), ptr(J))
        return float(J[0])

    def set_solve_guess(self, k: int):
        """Initial guess of the minimiser's Poisson solves: 0 the previous solution (the
        reference's warm start); k in 1..8 the better (by true residual) of that and the
        combination of the last k solutions whose right-hand sides best match the new one."""
        call("mm_set_solve_guess", self.handle, int(k))
        self.solve_guess = int(k)

    def minimize_segmented(self, params: "_hip.MMParams", B) -> list[dict]:
        """Field step of every segment at once (pcohen_hs): one info dict per segment, as
        ``minimize_resident`` would return for the segment in a context of its own."""
        if self.segments is None:
            raise RuntimeError("minimize_segmented needs a context created with segments")
        B = f64(B).reshape(3)
        infos = (_hip.MMInfo * (self.segments.size - 1))()
        call("mm_minimize_segmented", self.handle, C.byref(params), ptr(B), infos)
        return [_info_dict(i) for i in infos]

    def state_j_parallel_segmented(self, h) -> np.ndarray:
        """J_parallel of each segment's resident state."""
        h = f64(h).reshape(3)
        J = np.zeros(self.segments.size - 1)
        call("mm_state_j_parallel_segmented", self.handle, ptr(h), ptr(J))
        return J

    def segment_volumes(self) -> np.ndarray:
        """Magnetic volume V_mag of each segment."""
        V = np.zeros(1 if self.segments is None else self.segments.size - 1)
        call("mm_seg_volumes", self.handle, ptr(V))
        return V

    def counters(self):
        out = np.zeros(6, dtype=np.int64)
        call("mm_counters", self.handle, ptr(out))
        return dict(zip(("apply", "rhs", "energy", "local", "vcycle", "pcg_iters"), out.tolist()))

    def apply_timing(self, enable: bool):
        """(total ms, launches) of the apply kernel since the last call; then enable/disable timing."""
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        call("mm_apply_timing", self.handle, int(bool(enable)), C.byref(ms), C.byref(n))
        return float(ms.value), int(n.value)

    def bench_copy(self, reps=16):
        """Mean ms of a plain copy of N doubles (16 N bytes), ``reps`` launches."""
        ms = C.c_double(0.0)
        call("mm_bench_copy", self.handle, int(reps), C.byref(ms))
        return float(ms.value)

    def bench_apply(self, geom="on_the_fly", reps=50):
        ms = C.c_double(0.0)
        call("mm_bench_apply", self.handle, GEOM_MODES[geom], int(reps), C.byref(ms))
        return float(ms.value)

    def bench_precond(self, reps=20, use_graph=False):
        """Mean ms of one preconditioner application on device-resident vectors (in the context's
        precision form)."""
        ms = C.c_double(0.0)
        call("mm_bench_precond", self.handle, int(bool(use_graph)), int(reps), C.byref(ms))
        return float(ms.value)

    def amg_lanes(self) -> list[dict]:
        """Lanes per row of each level's CSR products: [{"A": a, "P": p, "R": r}, ...]."""
        out = []
        for lev in range(len(self.amg_levels or [])):
            v = np.zeros(3, dtype=np.int32)
            call("mm_amg_get_lanes", self.handle, lev, v.ctypes.data_as(C.POINTER(C.c_int32)))
            out.append({"A": int(v[0]), "P": int(v[1]), "R": int(v[2])})
        return out

    def set_amg_lanes(self, level: int, which: str, lanes: int):
        """Lanes per row (1..64) of level ``level``'s ``which`` product ("A", "P" or "R")."""
        call("mm_amg_set_lanes", self.handle, int(level), {"A": 0, "P": 1, "R": 2}[which], int(lanes))


def _info_dict(info) -> dict:
    return {
        "iters": info.iters,
        "evals": info.evals,
        "demag_iters": info.demag_iters,
        "preco_iters": info.preco_iters,
        "converged": bool(info.converged),
        "E": info.E,
        "gnorm": info.gnorm,
        "time": info.time,
    }


def minimize_batched(ctxs, params: "_hip.MMParams", B, n_threads: int = 0) -> list[dict]:
    """One field step on each of several independent contexts (their resident states), driven
    by ``n_threads`` host threads (0: up to 8); ``B`` is one field (3,) for all or one per
    context (n, 3). Same results as calling ``minimize_resident`` on each context in turn."""
    n = len(ctxs)
    B = f64(np.broadcast_to(np.asarray(B, dtype=np.float64), (n, 3)))
    handles = (C.c_int64 * n)(*[c.handle for c in ctxs])
    infos = (_hip.MMInfo * n)()
    call("mm_minimize_batched", n, handles, C.byref(params), ptr(B), int(n_threads), infos)
    return [_info_dict(i) for i in infos]


def solve_U_batched(ctxs, ms, U0s, tol, maxiter: int = 2000, n_threads: int = 0):
    """solve_U on several independent contexts at once (per-system tolerance ``tol``, a scalar
    or one per context). Returns lists (U, iterations, relative residual)."""
    n = len(ctxs)
    ms = [f64(m).reshape(-1) for m in ms]
    Us = [f64(U).copy() for U in U0s]
    tol = f64(np.broadcast_to(np.asarray(tol, dtype=np.float64), (n,)))
    f64p = C.POINTER(C.c_double)
    mp = (f64p * n)(*[m.ctypes.data_as(f64p) for m in ms])
    Up = (f64p * n)(*[U.ctypes.data_as(f64p) for U in Us])
    it = np.zeros(n, dtype=np.int32)
    rr = np.zeros(n)
    call("mm_solve_U_batched", n, (C.c_int64 * n)(*[c.handle for c in ctxs]), mp, Up, ptr(tol), int(maxiter),
         int(n_threads), ptr(it), ptr(rr))
    return Us, it.tolist(), rr.tolist()


def make_params(method="pcohen_hs", *, max_iter=2000, tau_f=1e-8, eps_a=1e-12, tau0=1e-2, ls_eta1=0.1,
                ls_eta2=0.1, ls_C=2.0, ls_c=0.5, ls_max_evals=15, pc_iters=10, pc_auto=True, pc_force_eta=0.5,
                pc_force_alpha=0.5, pc_stagnation_nu=0.01, pc_reg=0.0, L=None, cg_tol=1e-8, phi_tol=None,
                cg_maxiter=2000, memory=10, **_ignored) -> "_hip.MMParams":
    """C-ABI parameter struct with the reference's defaults (src/loop.py:528-700, src/minimizers.py:3447-3452).

    ``phi_tol`` defaults to min(cg_tol, 0.1 tau_f); the line search uses at most 15
    evaluations per phase (src/minimizers.py:1035,1335).
    """
    p = _hip.MMParams()
    p.method = METHODS[method]
    p.max_iter = int(max_iter)
    p.tau_f = float(tau_f)
    p.eps_a = float(eps_a)
    p.tau0 = float(tau0)
    p.ls_eta1 = float(ls_eta1)
    p.ls_eta2 = float(ls_eta2)
    p.ls_C = float(ls_C)
    p.ls_c = float(ls_c)
    p.ls_max_evals = int(ls_max_evals)
    p.pc_iters = int(pc_iters)
    p.pc_auto = int(bool(pc_auto))
    p.pc_force_eta = float(pc_force_eta)
    p.pc_force_alpha = float(pc_force_alpha)
    p.pc_stagnation_nu = float(pc_stagnation_nu)
    p.pc_reg = float(pc_reg)
    p.L = int(L or 0)
    p.cg_tol = float(cg_tol)
    p.phi_tol = float(min(cg_tol, 0.1 * tau_f) if phi_tol is None else phi_tol)
    p.cg_maxiter = int(cg_maxiter)
    p.memory = int(memory)
    call("mm_params_check", C.byref(p))
    return p
