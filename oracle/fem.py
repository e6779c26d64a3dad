"""CPU oracle: FP64 restatement of topomax's FEniCS elasticity hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product (``topomax_amd``) never imports it.

What it restates (reference files, relative to the topomax repository):

* mesh: ``FEM_src/solver.py:38-47`` -- dolfin ``RectangleMesh`` with the
  default 'right' diagonal; vertex v = iy*(nx+1)+ix at (a + ab*ix, c + cd*iy)
  with ab = (b-a)/nx; each cell gives triangles (v0,v1,v3) and (v0,v2,v3).
* SIMP: ``src/penalizers.py:29-45`` -- r(x) = m + x^p (1-m), m = 1e-6.
* Lame parameters: ``FEM_src/elasisity_problem.py:98-99`` --
  mu = E/(2(1+nu)), lambda = mu*nu/(0.5-nu) (used as-is in 2-D).
* bilinear form: ``FEM_src/elasisity_problem.py:112-118`` --
  a(u,v) = int r(xi) [lambda div u div v + 2 mu eps(u):eps(v)].
* loads: ``FEM_src/elasisity_problem.py:20-73,120-124`` -- body force and
  traction are dolfin ``UserExpression``s (default degree 2), i.e. the P2
  Lagrange interpolant of the pointwise rule, evaluated at each cell's vertices
  and edge midpoints 0.5*x_a + 0.5*x_b; the traction is integrated over the
  whole exterior boundary.
* Dirichlet: ``FEM_src/elasisity_problem.py:171-192`` +
  ``FEM_src/domains.py:18-38`` -- both components of every dof on the facets of
  each side whose coordinate is ``near`` (|x - t| < DOLFIN_EPS) the side.
  Solved by eliminating those dofs, which gives the same solution as
  ``bc.apply`` with zero boundary values (``FEM_src/pde_solver.py:125``).
* filter: ``FEM_src/filter.py:15-41`` -- (eps^2 grad xi, grad v) + (xi, v) =
  (f, v) on CG1, natural boundary condition, eps used unscaled.
* objective: ``FEM_src/elasisity_problem.py:152-166`` -- J = int u.f + int_ds u.t.
* gradient: ``FEM_src/elasisity_problem.py:134-150`` -- returns
  w = A^{-1} b with b_i = int -r'(xi) (lambda (div u)^2 + 2 mu eps:eps) phi_i.
* integrate: ``FEM_src/solver.py:81-84`` -- int of the CG1 interpolant.

The displacement element degree is a parameter: ``degree=2`` is the
reference's element (``FEM_src/elasisity_problem.py:194-196``) and is the one
checked against the reference's golden EMD run; ``degree=1`` is the same code
with the P1 element that the GPU path implements.

All integrals use a collapsed (Duffy) Gauss-Legendre rule of ``QUAD_N``^2 points
per triangle, exact for polynomial integrands up to degree 2*QUAD_N-2; with
integer p <= 8 every integrand here is a polynomial of degree <= 13, so the
oracle integrates exactly (FEniCS estimates degree 5 for the P2 stiffness, which
is also exact for p <= 3).  Sparse systems are solved directly with SuperLU
(``scipy.sparse.linalg.splu``), the same algorithm class as the reference's
MUMPS LU (``FEM_src/pde_solver.py:130-131``).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

DOLFIN_EPS = 3.0e-16
QUAD_N = 8
SIDES = ("Left", "Right", "Top", "Bottom")


# --------------------------------------------------------------------------- quadrature
def triangle_rule(n: int = QUAD_N):
    """Collapsed Gauss-Legendre rule on the reference triangle (0,0),(1,0),(0,1).

    Returns barycentric-free reference points (X, Y) and weights summing to 1/2.
    """
    g, w = np.polynomial.legendre.leggauss(n)
    s = 0.5 * (g + 1.0)
    ws = 0.5 * w
    X = np.repeat(s, n)
    T = np.tile(s, n)
    Y = T * (1.0 - X)
    W = np.repeat(ws, n) * np.tile(ws, n) * (1.0 - X)
    return X, Y, W


def line_rule(n: int = QUAD_N):
    g, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (g + 1.0), 0.5 * w


# --------------------------------------------------------------------------- basis functions
def p1_basis(X, Y):
    return np.stack([1.0 - X - Y, X, Y])


def p1_ref_grad():
    return np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]])


def p2_basis(X, Y):
    """P2 Lagrange basis, local order: 3 vertices, then midpoints of edges
    (1,2), (0,2), (0,1) (opposite vertex 0, 1, 2)."""
    l0, l1, l2 = 1.0 - X - Y, X, Y
    return np.stack([
        l0 * (2 * l0 - 1), l1 * (2 * l1 - 1), l2 * (2 * l2 - 1),
        4 * l1 * l2, 4 * l0 * l2, 4 * l0 * l1,
    ])


def p2_ref_grad(X, Y):
    """Gradients of the P2 basis w.r.t. (X, Y): shape (6, 2, nq)."""
    l0, l1, l2 = 1.0 - X - Y, X, Y
    d0 = np.array([-1.0, -1.0])
    d1 = np.array([1.0, 0.0])
    d2 = np.array([0.0, 1.0])
    out = np.empty((6, 2, X.size))
    out[0] = np.outer(d0, 4 * l0 - 1)
    out[1] = np.outer(d1, 4 * l1 - 1)
    out[2] = np.outer(d2, 4 * l2 - 1)
    out[3] = 4 * (np.outer(d1, l2) + np.outer(d2, l1))
    out[4] = 4 * (np.outer(d0, l2) + np.outer(d2, l0))
    out[5] = 4 * (np.outer(d0, l1) + np.outer(d1, l0))
    return out


# --------------------------------------------------------------------------- mesh
class StructuredMesh:
    """dolfin ``RectangleMesh((0,0),(W,H),nx,ny)`` with the 'right' diagonal.

    Vertex coordinates follow dolfin's RectangleMesh: x = a + ab*ix with
    ab = (b - a)/nx (same for y), evaluated in IEEE double exactly as written.
    """

    def __init__(self, width: float, height: float, nx: int, ny: int):
        self.width, self.height = float(width), float(height)
        self.nx, self.ny = int(nx), int(ny)
        ab = (self.width - 0.0) / self.nx
        cd = (self.height - 0.0) / self.ny
        self.xs = 0.0 + ab * np.arange(self.nx + 1, dtype=np.float64)
        self.ys = 0.0 + cd * np.arange(self.ny + 1, dtype=np.float64)
        self.n_vertices = (self.nx + 1) * (self.ny + 1)
        iy, ix = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nx + 1), indexing="ij")
        self.coords = np.stack([self.xs[ix.ravel()], self.ys[iy.ravel()]], axis=1)

        cy, cx = np.meshgrid(np.arange(self.ny), np.arange(self.nx), indexing="ij")
        v0 = (cy * (self.nx + 1) + cx).ravel()
        v1 = v0 + 1
        v2 = v0 + self.nx + 1
        v3 = v2 + 1
        # triangle order: for each cell (row-major), (v0,v1,v3) then (v0,v2,v3)
        tri = np.empty((2 * v0.size, 3), dtype=np.int64)
        tri[0::2] = np.stack([v0, v1, v3], axis=1)
        tri[1::2] = np.stack([v0, v2, v3], axis=1)
        self.triangles = tri
        self.n_triangles = tri.shape[0]

    @classmethod
    def from_design(cls, width: float, height: float, full_n: int):
        """Mesh the way ``Solver.__init__`` does (``src/solver.py:65-67``) and
        ``FEMSolver.prepare_domain`` (``FEM_src/solver.py:38-47``)."""
        n = int(full_n / min(width, height))
        return cls(width, height, int(width * n), int(height * n))

    # P2 node numbering: a (2nx+1) x (2ny+1) lattice in doubled grid indices
    def p2_nodes(self):
        nX, nY = 2 * self.nx + 1, 2 * self.ny + 1
        vert_grid = np.stack([self.triangles % (self.nx + 1) * 2,
                              self.triangles // (self.nx + 1) * 2], axis=2)  # (T, 3, 2)
        mids = []
        for a, b in ((1, 2), (0, 2), (0, 1)):
            mids.append((vert_grid[:, a] + vert_grid[:, b]) // 2)
        grid = np.concatenate([vert_grid, np.stack(mids, axis=1)], axis=1)  # (T, 6, 2)
        nodes = grid[..., 1] * nX + grid[..., 0]
        # coordinates of every P2 node as dolfin's evaluate_dofs computes them
        gx = np.empty(nX)
        gy = np.empty(nY)
        gx[0::2] = self.xs
        gy[0::2] = self.ys
        gx[1::2] = 0.5 * self.xs[:-1] + 0.5 * self.xs[1:]
        gy[1::2] = 0.5 * self.ys[:-1] + 0.5 * self.ys[1:]
        jj, ii = np.meshgrid(np.arange(nY), np.arange(nX), indexing="ij")
        coords = np.stack([gx[ii.ravel()], gy[jj.ravel()]], axis=1)
        return nodes, coords, nX, nY


def nested_dissection(nx: int, ny: int, leaf: int = 64) -> np.ndarray:
    """Nested-dissection order of the (nx+1) x (ny+1) vertex grid: halve the longer side of a
    box by one grid line, order the two halves first (recursively) and the line last; boxes of
    at most ``leaf`` vertices are ordered row by row.  For a 2D grid this keeps the Cholesky
    fill at O(n log n) (George 1973), which the generic orderings do not reach at N = 1000."""
    out = []
    stack = [(0, nx + 1, 0, ny + 1, False)]
    while stack:  # explicit stack: (box, separator pending)
        x0, x1, y0, y1, sep = stack.pop()
        w, h = x1 - x0, y1 - y0
        if w <= 0 or h <= 0:
            continue
        if sep:
            if w >= h:
                m = (x0 + x1) // 2
                out.append(np.arange(y0, y1) * (nx + 1) + m)
            else:
                m = (y0 + y1) // 2
                out.append(m * (nx + 1) + np.arange(x0, x1))
            continue
        if w * h <= leaf:
            iy, ix = np.mgrid[y0:y1, x0:x1]
            out.append((iy * (nx + 1) + ix).ravel())
            continue
        if w >= h:
            m = (x0 + x1) // 2
            stack += [(x0, x1, y0, y1, True), (m + 1, x1, y0, y1, False), (x0, m, y0, y1, False)]
        else:
            m = (y0 + y1) // 2
            stack += [(x0, x1, y0, y1, True), (x0, x1, m + 1, y1, False), (x0, x1, y0, m, False)]
    return np.concatenate(out)


def _splu(A, permc_spec: str, order=None):
    """SuperLU factorisation of a sparse SPD matrix; with ``permc_spec == "ND"`` the rows and
    columns are permuted by ``order`` first and factored in that order on the diagonal (no
    pivoting moves the order).  Returns solve(b)."""
    if permc_spec != "ND":
        lu = spla.splu(A.tocsc(), permc_spec=permc_spec)
        return lu.solve
    A = A.tocsr()
    lu = spla.splu(A[order][:, order].tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                   options=dict(SymmetricMode=True))

    def solve(b):
        x = np.empty_like(b)
        x[order] = lu.solve(b[order])
        return x

    return solve


def lame(young_modulus: float, poisson_ratio: float):
    """``FEM_src/elasisity_problem.py:98-99``."""
    mu = young_modulus / (2 * (1 + poisson_ratio))
    lmbda = mu * poisson_ratio / (0.5 - poisson_ratio)
    return lmbda, mu


def simp(x, p, m=1e-6):
    """``src/penalizers.py:36-40``: r = m + x^p (1 - m)."""
    return m + x ** p * (1 - m)


def simp_derivative(x, p, m=1e-6):
    """``src/penalizers.py:42-45``: r' = p x^(p-1) (1 - m)."""
    return p * x ** (p - 1) * (1 - m)


# --------------------------------------------------------------------------- loads
def body_force_value(force, x, y):
    """``FEM_src/elasisity_problem.py:26-33``: value if |x - c| < r (strict)."""
    (cx, cy), radius = force["center"], force["radius"]
    dist = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    inside = dist < radius
    out = np.zeros((x.size, 2))
    out[inside] = force["value"]
    return out


def traction_value(tractions, width, height, x, y):
    """``FEM_src/elasisity_problem.py:47-70``: exact side test + inclusive
    ``between`` with DOLFIN_EPS; overlapping tractions add."""
    out = np.zeros((x.size, 2))
    for t in tractions:
        side, center, length, value = t["side"], t["center"], t["length"], t["value"]
        lo, hi = center - length / 2, center + length / 2
        if side == "Left":
            sel = (x == 0.0) & (y >= lo - DOLFIN_EPS) & (y <= hi + DOLFIN_EPS)
        elif side == "Right":
            sel = (x == width) & (y >= lo - DOLFIN_EPS) & (y <= hi + DOLFIN_EPS)
        elif side == "Top":
            sel = (y == height) & (x >= lo - DOLFIN_EPS) & (x <= hi + DOLFIN_EPS)
        elif side == "Bottom":
            sel = (y == 0) & (x >= lo - DOLFIN_EPS) & (x <= hi + DOLFIN_EPS)
        else:
            raise ValueError(f"Malformed side: {side}")
        out[sel] += value
    return out


# --------------------------------------------------------------------------- problem
class ElasticityOracle:
    """FP64 CPU restatement of ``ElasticityProblem`` (``FEM_src/elasisity_problem.py:76-196``).

    Arrays are in vertex order (v = iy*(nx+1)+ix); displacements are
    node-interleaved (u[2k], u[2k+1]) over the P1 vertices or the P2 lattice.
    ``permc_spec`` is SuperLU's column ordering, or "ND" for the grid's nested-dissection
    order (P1 only; ``nested_dissection``).  The ordering changes only the round-off of the
    direct solves; "ND" keeps the fill of the N=1000 systems within a workstation's memory.
    """

    permc_spec = "COLAMD"

    def __init__(self, mesh: StructuredMesh, *, young_modulus: float, poisson_ratio: float,
                 filter_radius: float, fixed_sides, tractions=None, body_force=None,
                 degree: int = 1, minimum: float = 1e-6):
        if degree not in (1, 2):
            raise ValueError("degree must be 1 or 2")
        self.mesh = mesh
        self.degree = degree
        self.minimum = minimum
        self.E, self.nu = young_modulus, poisson_ratio
        self.lmbda, self.mu = lame(young_modulus, poisson_ratio)
        self.filter_radius = filter_radius
        self.fixed_sides = list(fixed_sides)
        self.tractions = list(tractions or [])
        self.body_force = body_force
        self.penalty = None

        self._geometry()
        self._displacement_space()
        self.load = self._assemble_load()
        self.free = np.setdiff1d(np.arange(self.n_dofs), self.fixed_dofs)
        self.filter_matrix, self.mass = self._assemble_filter()
        nd = (nested_dissection(self.mesh.nx, self.mesh.ny)
              if self.permc_spec == "ND" and degree == 1 else None)
        self._nd_order = nd
        self._filter_solve = _splu(self.filter_matrix, self.permc_spec if nd is not None
                                   else ("COLAMD" if self.permc_spec == "ND" else self.permc_spec),
                                   nd)
        self.filtered_rho = None
        self.u = None

    # ------------------------------------------------------------- setup
    def _geometry(self):
        m = self.mesh
        xy = m.coords[m.triangles]  # (T, 3, 2)
        self.tri_xy = xy
        J = np.stack([xy[:, 1] - xy[:, 0], xy[:, 2] - xy[:, 0]], axis=2)  # (T, 2, 2): cols
        self.detJ = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
        self.area = 0.5 * np.abs(self.detJ)
        inv = np.empty_like(J)
        inv[:, 0, 0] = J[:, 1, 1]
        inv[:, 1, 1] = J[:, 0, 0]
        inv[:, 0, 1] = -J[:, 0, 1]
        inv[:, 1, 0] = -J[:, 1, 0]
        inv /= self.detJ[:, None, None]
        self.invJ = inv  # maps reference gradient (row) -> physical: g_phys = g_ref @ invJ
        self.qX, self.qY, self.qW = triangle_rule()
        self.p1_at_q = p1_basis(self.qX, self.qY)  # (3, nq)
        self.p1_grad = np.einsum("ak,tkj->taj", p1_ref_grad(), inv)  # (T, 3, 2)

    def _displacement_space(self):
        m = self.mesh
        if self.degree == 1:
            self.cell_nodes = m.triangles
            self.node_coords = m.coords
            self.n_nodes = m.n_vertices
            self.node_grid = (m.nx + 1, m.ny + 1, 1)  # (nX, nY, lattice stride)
            self.basis_at_q = self.p1_at_q
            self.grad_at_q = np.broadcast_to(self.p1_grad[:, :, :, None],
                                             self.p1_grad.shape + (self.qX.size,))
        else:
            nodes, coords, nX, nY = m.p2_nodes()
            self.cell_nodes = nodes
            self.node_coords = coords
            self.n_nodes = nX * nY
            self.node_grid = (nX, nY, 2)
            self.basis_at_q = p2_basis(self.qX, self.qY)
            gref = p2_ref_grad(self.qX, self.qY)  # (6, 2, nq)
            self.grad_at_q = np.einsum("akq,tkj->tajq", gref, self.invJ)  # (T, 6, 2, nq)
        self.n_dofs = 2 * self.n_nodes
        self.fixed_dofs = self._dirichlet_dofs()

    def side_is_fixed_coordinate(self, side: str) -> bool:
        """``SidesDomain.inside`` (``FEM_src/domains.py:18-38``): the facets of a side are
        marked iff their coordinate is ``near`` the side (|x-t| < DOLFIN_EPS)."""
        m = self.mesh
        if side == "Left":
            return abs(m.xs[0] - 0.0) < DOLFIN_EPS
        if side == "Right":
            return abs(m.xs[-1] - m.width) < DOLFIN_EPS
        if side == "Top":
            return abs(m.ys[-1] - m.height) < DOLFIN_EPS
        if side == "Bottom":
            return abs(m.ys[0] - 0.0) < DOLFIN_EPS
        raise ValueError(f"Malformed side: {side}")

    def _dirichlet_dofs(self):
        nX, nY, _ = self.node_grid
        jj, ii = np.meshgrid(np.arange(nY), np.arange(nX), indexing="ij")
        ii, jj = ii.ravel(), jj.ravel()
        mask = np.zeros(self.n_nodes, dtype=bool)
        for side in self.fixed_sides:
            if not self.side_is_fixed_coordinate(side):
                continue
            if side == "Left":
                mask |= ii == 0
            elif side == "Right":
                mask |= ii == nX - 1
            elif side == "Top":
                mask |= jj == nY - 1
            elif side == "Bottom":
                mask |= jj == 0
        nodes = np.nonzero(mask)[0]
        self.fixed_node_mask = mask
        return np.sort(np.concatenate([2 * nodes, 2 * nodes + 1]))

    def _assemble_load(self):
        """F = int f.v dx + int_ds t.v (``FEM_src/elasisity_problem.py:120-124``)."""
        m = self.mesh
        F = np.zeros(self.n_dofs)
        p2_nodes, p2_coords, nX, nY = m.p2_nodes()
        if self.body_force is not None:
            vals = body_force_value(self.body_force, p2_coords[:, 0], p2_coords[:, 1])
            cell_vals = vals[p2_nodes]  # (T, 6, 2)
            fq = np.einsum("aq,tac->tqc", p2_basis(self.qX, self.qY), cell_vals)  # (T, nq, 2)
            contrib = np.einsum("aq,q,t,tqc->tac", self.basis_at_q, self.qW,
                                2 * self.area, fq)  # (T, nb, 2)
            self._scatter_vector(F, contrib)
        if self.tractions:
            self._assemble_traction(F, p2_coords, nX, nY)
        return F

    def _scatter_vector(self, F, contrib):
        nodes = self.cell_nodes
        np.add.at(F, 2 * nodes, contrib[..., 0])
        np.add.at(F, 2 * nodes + 1, contrib[..., 1])

    def _assemble_traction(self, F, p2_coords, nX, nY):
        m = self.mesh
        s, w = line_rule()
        # P2 edge interpolant: endpoints a, b and midpoint c
        e_basis = np.stack([(1 - s) * (1 - 2 * s), s * (2 * s - 1), 4 * s * (1 - s)])
        if self.degree == 1:
            t_basis = np.stack([1 - s, s])
        else:
            t_basis = e_basis
        nXn = self.node_grid[0]
        edges = []  # (lattice point a, lattice point b) in doubled-grid coordinates
        for ix in range(m.nx):
            edges.append(((2 * ix, 0), (2 * ix + 2, 0)))
            edges.append(((2 * ix, 2 * m.ny), (2 * ix + 2, 2 * m.ny)))
        for iy in range(m.ny):
            edges.append(((0, 2 * iy), (0, 2 * iy + 2)))
            edges.append(((2 * m.nx, 2 * iy), (2 * m.nx, 2 * iy + 2)))
        for (ax, ay), (bx, by) in edges:
            pts = [(ax, ay), (bx, by), ((ax + bx) // 2, (ay + by) // 2)]
            idx = np.array([py * nX + px for px, py in pts])
            tv = traction_value(self.tractions, m.width, m.height,
                                p2_coords[idx, 0], p2_coords[idx, 1])  # (3, 2)
            if not tv.any():
                continue
            pa, pb = p2_coords[idx[0]], p2_coords[idx[1]]
            length = math.hypot(pb[0] - pa[0], pb[1] - pa[1])
            tq = e_basis.T @ tv  # (nq, 2)
            vals = np.einsum("aq,q,qc->ac", t_basis, w * length, tq)
            if self.degree == 1:
                nodes = [(py // 2) * nXn + px // 2 for px, py in pts[:2]]
            else:
                nodes = [py * nXn + px for px, py in pts]
            for k, node in enumerate(nodes):
                F[2 * node] += vals[k, 0]
                F[2 * node + 1] += vals[k, 1]

    def _assemble_filter(self):
        """A = eps^2 K_lap + M on CG1 (``FEM_src/filter.py:26-36``), and M."""
        m = self.mesh
        g = self.p1_grad  # (T, 3, 2)
        Kl = np.einsum("tai,tbi,t->tab", g, g, self.area)
        Me = np.einsum("aq,bq,q,t->tab", self.p1_at_q, self.p1_at_q, self.qW, 2 * self.area)
        eps = self.filter_radius
        A = self._scalar_matrix(eps ** 2 * Kl + Me)
        M = self._scalar_matrix(Me)
        return A, M

    def _scalar_matrix(self, elem):
        tri = self.mesh.triangles
        rows = np.repeat(tri, 3, axis=1).ravel()
        cols = np.tile(tri, (1, 3)).ravel()
        n = self.mesh.n_vertices
        return sp.csr_matrix((elem.ravel(), (rows, cols)), shape=(n, n))

    # ------------------------------------------------------------- API
    def set_penalization(self, penalty: float):
        """``FEM_src/problem.py:29-35`` -> ``src/penalizers.py:11-12``."""
        self.penalty = penalty

    def _p(self):
        if self.penalty is None:
            raise ValueError("You must set penalization before calling penalizer")
        return self.penalty

    def filter_apply(self, rho):
        """xi = A^{-1} M rho (``HelmholtzFilter.apply`` with a CG1 input)."""
        return self._filter_solve(self.mass @ np.asarray(rho, dtype=np.float64))

    def filter_solve(self, b):
        """w = A^{-1} b (``HelmholtzFilter.apply`` with a form input: b already assembled)."""
        return self._filter_solve(np.asarray(b, dtype=np.float64))

    def _xi_at_q(self, xi, local=False):
        """P1 interpolant of nodal xi at the quadrature points, (T, nq).  With
        ``local`` the argument is already gathered per triangle, (T, 3)."""
        xt = xi if local else xi[self.mesh.triangles]
        return np.einsum("aq,ta->tq", self.p1_at_q, xt)

    def element_weights(self, xi):
        """Per-triangle mean of r(xi): (1/|T|) int_T r(xi_h), by quadrature."""
        xi = np.asarray(xi, dtype=np.float64)
        out = np.empty(self.mesh.n_triangles)
        for sl in self._chunks():
            r = simp(self._xi_at_q(xi[self.mesh.triangles[sl]], local=True), self._p(),
                     self.minimum)
            out[sl] = 2.0 * (r @ self.qW)
        return out

    def _chunks(self, size=65536):
        nt = self.mesh.n_triangles
        return [slice(s0, min(nt, s0 + size)) for s0 in range(0, nt, size)]

    def _element_matrices(self, grads, weights):
        """6x6 / 12x12 element matrices sum_q weights[t,q] * B_q^T D B_q.

        grads: (T, nb, 2, nq) basis gradients; weights: (T, nq) = quadrature weight
        * |J| * r(xi) at each point."""
        lam, mu = self.lmbda, self.mu
        GG = np.einsum("taiq,tbkq,tq->tabik", grads, grads, weights)
        tr = GG[..., 0, 0] + GG[..., 1, 1]
        nb = grads.shape[1]
        Ke = np.empty((grads.shape[0], nb, 2, nb, 2))
        for i in range(2):
            for k in range(2):
                Ke[:, :, i, :, k] = lam * GG[:, :, :, i, k] + mu * GG[:, :, :, k, i]
                if i == k:
                    Ke[:, :, i, :, k] += mu * tr
        return Ke.reshape(grads.shape[0], 2 * nb, 2 * nb)

    def stiffness(self, xi):
        """Assembled K(xi) over all dofs (no boundary conditions)."""
        xi = np.asarray(xi, dtype=np.float64)
        nt = self.mesh.n_triangles
        if self.degree == 1:
            # constant strain: K_e = (int_T r(xi_h)) B^T D B, the integral by quadrature
            wts = (self.element_weights(xi) * self.area)[:, None]
            Ke = self._element_matrices(self.p1_grad[..., None], wts)
        else:
            Ke = np.empty((nt, 12, 12))
            for s0 in range(0, nt, 4096):
                sl = slice(s0, min(nt, s0 + 4096))
                r = simp(self._xi_at_q(xi[self.mesh.triangles[sl]], local=True),
                         self._p(), self.minimum)
                wq = self.qW[None, :] * (2 * self.area[sl])[:, None] * r
                Ke[sl] = self._element_matrices(self.grad_at_q[sl], wq)
        nb = self.cell_nodes.shape[1]
        dofs = np.stack([2 * self.cell_nodes, 2 * self.cell_nodes + 1], axis=2).reshape(nt, 2 * nb)
        rows = np.repeat(dofs, 2 * nb, axis=1).ravel()
        cols = np.tile(dofs, (1, 2 * nb)).ravel()
        return sp.csr_matrix((Ke.ravel(), (rows, cols)), shape=(self.n_dofs, self.n_dofs))

    def masked_operator(self, xi):
        """P K P + (I - P): K with Dirichlet rows/cols replaced by the identity."""
        K = self.stiffness(xi).tocsr()
        keep = np.ones(self.n_dofs)
        keep[self.fixed_dofs] = 0.0
        Pm = sp.diags(keep)
        return (Pm @ K @ Pm + sp.diags(1.0 - keep)).tocsr()

    def forward(self, xi):
        """Solve K(xi) u = F with homogeneous Dirichlet dofs (``forward`` :168-169)."""
        K = self.stiffness(xi).tocsc()
        f = self.free
        u = np.zeros(self.n_dofs)
        order = None
        if self._nd_order is not None:  # the grid order of the free dofs, both components
            pos = np.full(self.n_dofs, -1)
            pos[f] = np.arange(f.size)
            order = pos[np.stack([2 * self._nd_order, 2 * self._nd_order + 1], 1).ravel()]
            order = order[order >= 0]
        spec = "ND" if order is not None else ("COLAMD" if self.permc_spec == "ND"
                                               else self.permc_spec)
        u[f] = _splu(K[f][:, f], spec, order)(self.load[f])
        return u

    def compliance(self, u):
        """J = int u.f dx + int_ds u.t = F.u (``calculate_objective`` :161-164)."""
        return float(self.load @ u)

    def strain_energy_density(self, u, sl=slice(None)):
        """sigma_0(u):eps(u) = lam (div u)^2 + 2 mu eps:eps at the quadrature points
        of the triangles ``sl``, shape (T, nq)."""
        G = self.grad_at_q[sl]
        nodes = self.cell_nodes[sl]
        ue = np.stack([u[2 * nodes], u[2 * nodes + 1]], axis=2)  # (T, nb, 2)
        grad_u = np.einsum("tac,tajq->tcjq", ue, G)  # du_c/dx_j
        div = grad_u[:, 0, 0] + grad_u[:, 1, 1]
        exy = 0.5 * (grad_u[:, 0, 1] + grad_u[:, 1, 0])
        epse = grad_u[:, 0, 0] ** 2 + grad_u[:, 1, 1] ** 2 + 2 * exy ** 2
        return self.lmbda * div ** 2 + 2 * self.mu * epse

    def sensitivity_rhs(self, xi, u):
        """b_i = int -r'(xi) sigma_0(u):eps(u) phi_i over CG1 (gradient form :146-149)."""
        xi = np.asarray(xi, dtype=np.float64)
        u = np.asarray(u, dtype=np.float64)
        b = np.zeros(self.mesh.n_vertices)
        for sl in self._chunks():
            tri = self.mesh.triangles[sl]
            dr = simp_derivative(self._xi_at_q(xi[tri], local=True), self._p(), self.minimum)
            g = -dr * self.strain_energy_density(u, sl)  # (T, nq)
            contrib = np.einsum("aq,q,t,tq->ta", self.p1_at_q, self.qW, 2 * self.area[sl], g)
            np.add.at(b, tri, contrib)
        return b

    def calculate_objective(self, rho):
        """``calculate_objective`` (:152-166)."""
        self._p()
        self.filtered_rho = self.filter_apply(rho)
        self.u = self.forward(self.filtered_rho)
        return self.compliance(self.u)

    def calculate_objective_gradient(self):
        """``calculate_objective_gradient`` (:134-150): returns w = A^{-1} b.

        dJ/drho (Euclidean, nodal) is ``mass @ w``."""
        if self.filtered_rho is None or self.u is None:
            raise ValueError("You must call calculate_objective "
                             "before calling calculate_objective_gradient")
        return self.filter_solve(self.sensitivity_rhs(self.filtered_rho, self.u))

    def lumped_mass(self):
        """m_i = int phi_i; ``integrate(values)`` = m . values (``FEM_src/solver.py:81-84``)."""
        return np.asarray(self.mass.sum(axis=1)).ravel()

    def integrate(self, values):
        tri = self.mesh.triangles
        return float(np.sum(self.area * np.asarray(values)[tri].sum(axis=1) / 3.0))


def problem_from_design(design: dict, full_n: int, degree: int = 1):
    """Build an oracle problem from a parsed design JSON (``designs/*.json`` schema,
    ``designs/definitions.py:170-223``)."""
    (key,) = design.keys()
    dom = design[key]["domain_parameters"]
    par = design[key]["problem_parameters"]
    mesh = StructuredMesh.from_design(dom["width"], dom["height"], full_n)
    body = None
    if par.get("body_force") is not None:
        bf = par["body_force"]
        body = {"center": tuple(bf["region"]["center"]), "radius": bf["region"]["radius"],
                "value": tuple(bf["value"])}
    tractions = None
    if par.get("tractions") is not None:
        tractions = [{"side": t["side"], "center": t["center"], "length": t["length"],
                      "value": tuple(t["value"])} for t in par["tractions"]]
    return ElasticityOracle(mesh, young_modulus=par["young_modulus"],
                            poisson_ratio=par["poisson_ratio"],
                            filter_radius=par["filter_radius"],
                            fixed_sides=par["fixed_sides"], tractions=tractions,
                            body_force=body, degree=degree)
