"""Oracle: geometry precompute, assembled operators, energy and gradient (test infrastructure only).

All formulas are the reference's; arrays are float64 NumPy, matrices SciPy CSR.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

MU0 = 4e-7 * np.pi

GRAD_HAT = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


# ----------------------------------------------------------------------------- geometry
def compute_volume_JinvT(knt, conn):
    """Orientation fix, volumes and J^-T (``src/loop.py:56-90``).

    Elements with det J < 0 get nodes 1 and 2 swapped; Ve = |det J|/6.
    """
    knt = np.asarray(knt, dtype=np.float64)
    conn = np.asarray(conn, dtype=np.int64)[:, :4].copy()
    x0 = knt[conn[:, 0]]
    J = np.stack([knt[conn[:, 1]] - x0, knt[conn[:, 2]] - x0, knt[conn[:, 3]] - x0], axis=2)
    det = np.linalg.det(J)
    neg = det < 0
    if np.any(neg):
        conn[neg, 1], conn[neg, 2] = conn[neg, 2].copy(), conn[neg, 1].copy()
        x0 = knt[conn[:, 0]]
        J = np.stack([knt[conn[:, 1]] - x0, knt[conn[:, 2]] - x0, knt[conn[:, 3]] - x0], axis=2)
        det = np.linalg.det(J)
    vol = np.abs(det) / 6.0
    JinvT = np.swapaxes(np.linalg.inv(J), 1, 2)
    return conn.astype(np.int32), vol, JinvT


def grad_phi_from_JinvT(JinvT):
    """grad phi_a = J^-T grad phi_hat_a (``src/loop.py:93-102``) -> (E, 4, 3)."""
    return np.einsum("eij,aj->eai", JinvT, GRAD_HAT)


def node_sum(conn, per_elem, N):
    """Lumped nodal sum of per-element quarter weights: out_i = sum_{e ni i} w_e / 4.

    This is ``compute_node_volumes`` (``src/fem_utils.py:174-212``) with weights Ve
    (node volumes) or Js_red*Ve (M_nodal, ``src/loop.py:929-939``).
    """
    out = np.zeros(N)
    np.add.at(out, np.asarray(conn)[:, :4].ravel(), np.repeat(0.25 * np.asarray(per_elem), 4))
    return out


def reduce_materials(A_si, K1_si, Js_si, mesh_unit=1e-9):
    """Unit normalisation of ``src/loop.py:163-164,862-877``.

    Returns (A_red, K1_red, Js_red, Js_ref, Kd_ref); A in J/m is first scaled by (1/mesh_unit)^2.
    """
    A_si = np.asarray(A_si, dtype=np.float64)
    K1_si = np.asarray(K1_si, dtype=np.float64)
    Js_si = np.asarray(Js_si, dtype=np.float64)
    A_mesh = A_si * (1.0 / mesh_unit) ** 2
    Js_ref = float(np.max(Js_si)) if np.max(Js_si) > 0 else 1.0
    Kd = Js_ref**2 / (2.0 * MU0)
    return A_mesh / Kd, K1_si / Kd, Js_si / Js_ref, Js_ref, Kd


# ----------------------------------------------------------------------------- assembly
def _coo_4x4(conn, Ke, N):
    conn = np.asarray(conn, dtype=np.int64)
    rows = np.repeat(conn, 4, axis=1).ravel()
    cols = np.tile(conn, (1, 4)).ravel()
    M = sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(N, N)).tocsr()
    M.sum_duplicates()
    return M


def poisson_matrix(conn, vol, grad_phi, boundary_mask=None, reg=1e-12, N=None):
    """Assembled Poisson matrix with Dirichlet rows+cols and regularised diagonal (``src/amg_utils.py:21-84``).

    A_ab^e = Ve grad phi_a . grad phi_b; boundary rows and columns are zeroed, reg is
    added on the diagonal, and boundary diagonals are set to exactly 1.
    """
    N = int(np.max(conn)) + 1 if N is None else N
    Ke = vol[:, None, None] * np.einsum("eai,ebi->eab", grad_phi, grad_phi)
    A = _coo_4x4(conn, Ke, N)
    if boundary_mask is not None:
        is_b = np.asarray(boundary_mask) == 0
        rows = np.repeat(np.arange(N), np.diff(A.indptr))
        A.data[is_b[rows]] = 0.0
        A.data[is_b[A.indices]] = 0.0
        A.eliminate_zeros()
    A = (A + reg * sp.eye(N, format="csr")).tocsr()
    if boundary_mask is not None:
        d = A.diagonal()
        d[np.asarray(boundary_mask) == 0] = 1.0
        A.setdiag(d)
    A.sort_indices()
    return A


def divergence_matrices(conn, vol, grad_phi, Js_lookup, mat_id, N=None):
    """Dx, Dy, Dz with (D_i)_ab^e = Js_e Ve/4 (grad phi_a)_i for every column b (``src/amg_utils.py:660-702``)."""
    N = int(np.max(conn)) + 1 if N is None else N
    f = np.asarray(Js_lookup)[np.asarray(mat_id) - 1] * vol / 4.0
    out = []
    for c in range(3):
        Ke = f[:, None, None] * np.repeat(grad_phi[:, :, c][:, :, None], 4, axis=2)
        out.append(_coo_4x4(conn, Ke, N))
    return tuple(out)


def divergence_interleaved(Dx, Dy, Dz):
    """D (N x 3N) acting on interleaved m (column 3j+c), and G = 2 D^T (3N x N, row 3i+c).

    The C++ layouts of ``src/loop.py:1133-1166``.
    """
    N = Dx.shape[0]
    D = sp.hstack([Dx, Dy, Dz]).tocsr()
    # permute blocked columns (c*N + j) to interleaved (3j + c)
    perm = np.empty(3 * N, dtype=np.int64)
    for c in range(3):
        perm[c * N:(c + 1) * N] = 3 * np.arange(N) + c
    Dc = D.tocoo()
    Di = sp.csr_matrix((Dc.data, (Dc.row, perm[Dc.col])), shape=(N, 3 * N))
    Di.sum_duplicates()
    G = (2.0 * Di.T).tocsr()
    return Di, G


def exchange_anisotropy_matrix(conn, vol, grad_phi, A_lookup, K1_lookup, k_easy_lookup, mat_id, N=None):
    """K_eff (3N x 3N, interleaved) = exchange (x) I3 + anisotropy (x) k k^T (``src/amg_utils.py:741-789``).

    Kex_e = 2 A Ve grad phi_a . grad phi_b; Kan_e = -2 K1 Ve/20 (1 + delta_ab).
    """
    N = int(np.max(conn)) + 1 if N is None else N
    gi = np.asarray(mat_id) - 1
    A_e = np.asarray(A_lookup)[gi]
    K1_e = np.asarray(K1_lookup)[gi]
    k_e = np.asarray(k_easy_lookup)[gi]
    Kex = 2.0 * A_e[:, None, None] * vol[:, None, None] * np.einsum("eai,ebi->eab", grad_phi, grad_phi)
    Kan = (-2.0 * K1_e * vol / 20.0)[:, None, None] * (np.ones((4, 4)) + np.eye(4))
    blk = Kex[:, :, :, None, None] * np.eye(3)[None, None, None] + Kan[:, :, :, None, None] * np.einsum(
        "eu,ev->euv", k_e, k_e
    )[:, None, None]
    conn = np.asarray(conn, dtype=np.int64)
    rn = np.repeat(conn, 4, axis=1).ravel()
    cn = np.tile(conn, (1, 4)).ravel()
    r = 3 * rn[:, None, None] + np.arange(3)[None, :, None]
    c = 3 * cn[:, None, None] + np.arange(3)[None, None, :]
    r, c = np.broadcast_arrays(r, c)
    K = sp.coo_matrix((blk.ravel(), (r.ravel(), c.ravel())), shape=(3 * N, 3 * N)).tocsr()
    K.sum_duplicates()
    return K


def exchange_diagonal(conn, vol, grad_phi, A_lookup, mat_id, V_mag, N=None):
    """diag(K_ex)/V_mag: sum_e 2 A Ve |grad phi_a|^2 / V_mag (``src/energy_kernels.py:492-566``)."""
    N = int(np.max(conn)) + 1 if N is None else N
    a = 2.0 * np.asarray(A_lookup)[np.asarray(mat_id) - 1] * vol
    contrib = a[:, None] * np.sum(grad_phi**2, axis=2)
    d = np.zeros(N)
    np.add.at(d, np.asarray(conn)[:, :4].ravel(), contrib.ravel())
    return d / V_mag


def spai0_diagonal(A):
    """SPAI0 M_ii = a_ii / sum_j a_ij^2 (``src/amg_utils.py:87-111``)."""
    A = sp.csr_matrix(A)
    A2 = sp.csr_matrix((A.data**2, A.indices, A.indptr), shape=A.shape)
    return A.diagonal() / (np.asarray(A2.sum(axis=1)).ravel() + 1e-30)


# ----------------------------------------------------------------------------- physics
class Problem:
    """Everything the reference precomputes for one mesh (``src/loop.py:856-939,1089-1176``).

    Attributes are NumPy/SciPy objects: conn, vol, grad_phi, mat_id, N, E, V_mag,
    M_nodal, boundary_mask, A (Poisson, masked + reg), D/G (interleaved), K_eff,
    inv_M_rel and inv_M_prec (``src/hysteresis_loop.py:340-353``).
    """

    def __init__(self, knt, ijk, A_lookup, K1_lookup, Js_lookup, k_easy_lookup, boundary_mask=None, reg=1e-12):
        ijk = np.asarray(ijk)
        self.knt = np.asarray(knt, dtype=np.float64)
        self.N = self.knt.shape[0]
        self.mat_id = ijk[:, 4].astype(np.int32) if ijk.shape[1] == 5 else np.ones(ijk.shape[0], np.int32)
        self.conn, self.vol, self.JinvT = compute_volume_JinvT(self.knt, ijk[:, :4])
        self.E = self.conn.shape[0]
        self.grad_phi = grad_phi_from_JinvT(self.JinvT)
        self.A_lookup = np.asarray(A_lookup, dtype=np.float64)
        self.K1_lookup = np.asarray(K1_lookup, dtype=np.float64)
        self.Js_lookup = np.asarray(Js_lookup, dtype=np.float64)
        self.k_easy_lookup = np.asarray(k_easy_lookup, dtype=np.float64).reshape(-1, 3)
        is_mag = np.isin(self.mat_id, np.where(self.Js_lookup > 0)[0] + 1)
        V_mag = float(np.sum(self.vol[is_mag]))
        self.V_mag = V_mag if V_mag != 0 else 1.0
        self.boundary_mask = None if boundary_mask is None else np.asarray(boundary_mask, dtype=np.float64)
        self.reg = reg
        self.node_vols = node_sum(self.conn, self.vol, self.N)
        self.M_nodal = node_sum(self.conn, self.vol * self.Js_lookup[self.mat_id - 1], self.N)
        self.A = poisson_matrix(self.conn, self.vol, self.grad_phi, self.boundary_mask, reg, self.N)
        Dx, Dy, Dz = divergence_matrices(self.conn, self.vol, self.grad_phi, self.Js_lookup, self.mat_id, self.N)
        self.D, self.G = divergence_interleaved(Dx, Dy, Dz)
        self.K_eff = exchange_anisotropy_matrix(
            self.conn, self.vol, self.grad_phi, self.A_lookup, self.K1_lookup, self.k_easy_lookup, self.mat_id, self.N
        )
        self.inv_M_rel = np.where(self.M_nodal > 1e-20, self.V_mag / np.where(self.M_nodal > 1e-20, self.M_nodal, 1.0), 0.0)
        d = exchange_diagonal(self.conn, self.vol, self.grad_phi, self.A_lookup, self.mat_id, self.V_mag, self.N)
        self.exchange_diag = d
        self.inv_M_prec = np.where(d > 1e-20, 1.0 / np.where(d > 1e-20, d, 1.0), 1.0)

    def rhs(self, m):
        """Assembled rhs_from_m, b = D m (``src/poisson_solve.py:173-186``).

        The driver's ``sparse_ops`` carries no ``boundary_mask`` (``src/hysteresis_loop.py:421-440``),
        so no mask is applied here; the Dirichlet rows live in A itself.
        """
        return self.D @ np.asarray(m, dtype=np.float64).reshape(-1)

    def apply_A(self, U):
        """Assembled apply y = A U (``src/poisson_solve.py:166-171``); boundary rows of A are identity rows."""
        return self.A @ U

    def energy_and_grad(self, m, U, B_ext, B_bias=None):
        """Assembled energy and gradient (``src/energy_kernels.py:147-180``).

        g = K_eff m + G U - 2 M_nodal B; E = 1/2 sum m.(g + g_z); both divided by V_mag.
        """
        m = np.asarray(m, dtype=np.float64)
        g = (self.K_eff @ m.reshape(-1)).reshape(-1, 3) + (self.G @ U).reshape(-1, 3)
        B_eff = np.asarray(B_ext, dtype=np.float64)[None, :]
        if B_bias is not None:
            B_eff = B_eff + B_bias
        gz = -2.0 * self.M_nodal[:, None] * B_eff
        g = g + gz
        E = 0.5 * np.sum(m * (g + gz))
        return E / self.V_mag, g / self.V_mag

    def energy_terms(self, m, U, B_ext):
        """Exchange, anisotropy, demag and Zeeman energies separately (each / V_mag)."""
        m = np.asarray(m, dtype=np.float64)
        N = self.N
        Kex = exchange_anisotropy_matrix(self.conn, self.vol, self.grad_phi, self.A_lookup,
                                         np.zeros_like(self.K1_lookup), self.k_easy_lookup, self.mat_id, N)
        gex = (Kex @ m.reshape(-1)).reshape(-1, 3)
        gan = (self.K_eff @ m.reshape(-1)).reshape(-1, 3) - gex
        gdm = (self.G @ U).reshape(-1, 3)
        gz = -2.0 * self.M_nodal[:, None] * np.asarray(B_ext)[None, :]
        s = 1.0 / self.V_mag
        return {"exchange": 0.5 * np.sum(m * gex) * s, "anisotropy": 0.5 * np.sum(m * gan) * s,
                "demag": 0.5 * np.sum(m * gdm) * s, "zeeman": np.sum(m * gz) * s}

    def local_grad(self, v):
        """K_eff v / V_mag (``src/energy_kernels.py:190-195``)."""
        return (self.K_eff @ np.asarray(v, dtype=np.float64).reshape(-1)).reshape(-1, 3) / self.V_mag

    # --------------------------------------------------------- matrix-free element loops
    def rhs_elementwise(self, m):
        """Matrix-free rhs_from_m (``src/poisson_solve.py:254-280``)."""
        m = np.asarray(m, dtype=np.float64)
        me = m[self.conn].sum(axis=1)  # (E,3)
        Js = self.Js_lookup[self.mat_id - 1]
        contrib = (self.vol * Js)[:, None] * 0.25 * np.einsum("eai,ei->ea", self.grad_phi, me)
        b = np.zeros(self.N)
        np.add.at(b, self.conn.ravel(), contrib.ravel())
        return b

    def energy_and_grad_elementwise(self, m, U, B_ext):
        """Matrix-free energy_and_grad element loop (``src/energy_kernels.py:340-459``)."""
        m = np.asarray(m, dtype=np.float64)
        gi = self.mat_id - 1
        a_ve = 2.0 * self.A_lookup[gi] * self.vol
        q_ve = 2.0 * self.K1_lookup[gi] * self.vol / 20.0
        j_ve = 2.0 * self.Js_lookup[gi] * self.vol / 4.0
        k = self.k_easy_lookup[gi]
        B = self.grad_phi  # (E,4,3)
        me = m[self.conn]  # (E,4,3)
        grad_m = np.einsum("eai,eac->eci", B, me)  # d m_c / d x_i
        ex = a_ve[:, None, None] * np.einsum("eai,eci->eac", B, grad_m)
        v = np.einsum("eac,ec->ea", me, k)
        an = (-q_ve[:, None] * (v.sum(axis=1)[:, None] + v))[:, :, None] * k[:, None, :]
        gU = np.einsum("eai,ea->ei", B, U[self.conn])
        dm = np.broadcast_to((j_ve[:, None] * gU)[:, None, :], ex.shape)
        contrib = ex + an + dm
        g = np.zeros((self.N, 3))
        np.add.at(g, self.conn.ravel(), contrib.reshape(-1, 3))
        gz = -2.0 * self.M_nodal[:, None] * np.asarray(B_ext)[None, :]
        g = g + gz
        E = 0.5 * np.sum(m * (g + gz))
        return E / self.V_mag, g / self.V_mag

    def J_parallel(self, m, h_dir):
        """Volume-averaged polarisation along h (``src/hysteresis_loop.py:153-190``), reduced units."""
        h = np.asarray(h_dir, dtype=np.float64)
        h = h / (np.linalg.norm(h) + 1e-30)
        m_avg = np.asarray(m)[self.conn].mean(axis=1)
        Js_e = self.Js_lookup[self.mat_id - 1]
        Vm = np.sum(np.where(Js_e > 0, self.vol, 0.0)) + 1e-30
        J = np.sum(self.vol[:, None] * Js_e[:, None] * m_avg, axis=0) / Vm
        return float(J @ h)
