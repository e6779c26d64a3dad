"""CPU tests pinning the oracle (oracle/fem.py, oracle/emd.py).

Each check is chosen so that a plausible mistake in the oracle fails it:

* the reference's only elasticity known answer (triangle design, P2, EMD, 24 iterations),
  asserted by tests/test_elasticity_solver.py:54-55 of the reference;
* the reference's filter tests (tests/test_filter.py:25-60): epsilon = 0 is the identity and
  the manufactured solution converges with slope <= -2;
* closed forms: exact integral of r(xi) over a triangle via complete homogeneous polynomials,
  the lumped mass of the 'right' mesh, total load of a body force / traction;
* invariants: symmetric PSD stiffness with a 3-dimensional rigid-body null space, the linear
  patch test, F 1 = 1, mass conservation and M-self-adjointness of the filter, J = F.u = u.K u,
  monotonicity of J in rho;
* brute force: an independent loop-based P1 assembler on tiny meshes;
* the adjoint gradient M w against central finite differences of J.
"""

import json
import math

import numpy as np
import pytest

from oracle.emd import run_emd
from oracle.fem import ElasticityOracle, StructuredMesh, lame, problem_from_design, simp
from tests.conftest import design_path


def small_problem(nx=6, ny=4, degree=1, fixed=("Left",), tractions=None, body=None, eps=0.05,
                  width=1.5, height=1.0):
    mesh = StructuredMesh(width, height, nx, ny)
    return ElasticityOracle(mesh, young_modulus=1.3, poisson_ratio=0.3, filter_radius=eps,
                            fixed_sides=list(fixed), tractions=tractions, body_force=body,
                            degree=degree)


# ------------------------------------------------------------------ reference golden run
def test_triangle_golden_emd_p2(golden_dir):
    golden = json.load(open(f"{golden_dir}/triangle_golden.json"))
    design = json.load(open(design_path("triangle")))
    prob = problem_from_design(design, golden["N"], degree=2)
    dom = design["Elasticity"]["domain_parameters"]
    res = run_emd(prob, width=dom["width"], height=dom["height"],
                  volume_fraction=dom["volume_fraction"], step_size=dom["fem_step_size"],
                  penalties=dom["penalties"])[0]
    assert res["iteration"] == golden["iteration"]
    assert abs(res["objectives"][-1] - golden["objective"]) < 1e-14  # reference's own bound
    assert abs(res["objectives"][-1] - golden["objective"]) < 1e-12 * golden["objective"]
    # the golden design is in dolfin dof order: compare order-free
    assert np.abs(np.sort(res["rho"]) - np.array(golden["rho_sorted"])).max() < 1e-10


def test_golden_rejects_inclusive_body_force(golden_dir, monkeypatch):
    """The body force uses a strict |x - c| < r at the P2 points: an inclusive test changes
    the golden run completely, so the golden pins this membership rule."""
    import oracle.fem as F
    golden = json.load(open(f"{golden_dir}/triangle_golden.json"))

    def inclusive(force, x, y):
        (cx, cy), radius = force["center"], force["radius"]
        dist = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
        out = np.zeros((x.size, 2))
        out[dist <= radius + 1e-12] = force["value"]
        return out

    monkeypatch.setattr(F, "body_force_value", inclusive)
    design = json.load(open(design_path("triangle")))
    prob = F.problem_from_design(design, 10, degree=2)
    res = run_emd(prob, width=1.0, height=1.0, volume_fraction=0.5, step_size=25.0,
                  penalties=[3.0])[0]
    assert abs(res["objectives"][-1] - golden["objective"]) > 1e-3 * golden["objective"]


# ------------------------------------------------------------------ filter (reference tests)
def test_filter_identity_when_eps_zero():
    prob = small_problem(10, 10, eps=0.0, width=1.0, height=1.0)
    np.random.seed(198)
    rho = np.random.random(prob.mesh.n_vertices)
    xi = prob.filter_apply(rho)
    assert np.sqrt((xi - rho) @ (prob.mass @ (xi - rho))) < 1e-14


def _l2_error_vs(prob, xi, exact):
    X, Y, W = prob.qX, prob.qY, prob.qW
    xy = prob.tri_xy  # (T, 3, 2)
    px = xy[:, 0, 0, None] + (xy[:, 1, 0] - xy[:, 0, 0])[:, None] * X + (xy[:, 2, 0] - xy[:, 0, 0])[:, None] * Y
    py = xy[:, 0, 1, None] + (xy[:, 1, 1] - xy[:, 0, 1])[:, None] * X + (xy[:, 2, 1] - xy[:, 0, 1])[:, None] * Y
    xh = np.einsum("aq,ta->tq", prob.p1_at_q, xi[prob.mesh.triangles])
    err = (xh - exact(px, py)) ** 2
    return math.sqrt(float(np.sum(err @ W * 2 * prob.area)))


def test_filter_manufactured_solution_converges():
    eps = np.e / np.pi

    def rho_f(x, y):
        return (8 * eps * eps * np.pi ** 2 + 1) * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)

    def xi_f(x, y):
        return np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)

    Ns = list(range(10, 91, 10))
    errors = []
    for N in Ns:
        prob = small_problem(N, N, eps=eps, width=1.0, height=1.0, fixed=())
        c = prob.mesh.coords
        xi = prob.filter_apply(rho_f(c[:, 0], c[:, 1]))
        errors.append(_l2_error_vs(prob, xi, xi_f))
    fit = np.polynomial.Polynomial.fit(np.log(Ns), np.log(errors), 1)
    # the reference's get_convergance (tests/utils.py:11-33) returns fit.coef[1], the slope in
    # numpy's scaled window; the true log-log slope (about -1.96 for P1 in L2) is also checked
    assert fit.coef[1] <= -2
    assert fit.convert().coef[1] < -1.9


def test_filter_invariants():
    prob = small_problem(12, 7, eps=0.1)
    n = prob.mesh.n_vertices
    assert np.abs(prob.filter_apply(np.ones(n)) - 1).max() < 1e-12  # F 1 = 1
    rng = np.random.default_rng(1)
    a, b = rng.random(n), rng.random(n)
    assert abs(prob.integrate(prob.filter_apply(a)) - prob.integrate(a)) < 1e-13
    lhs = prob.filter_apply(a) @ (prob.mass @ b)
    rhs = a @ (prob.mass @ prob.filter_apply(b))
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


# ------------------------------------------------------------------ closed forms
def _hsym(vals, p):
    # complete homogeneous symmetric polynomial by brute force over compositions
    from itertools import combinations_with_replacement
    return sum(np.prod([vals[i] for i in c]) for c in combinations_with_replacement(range(3), p))


@pytest.mark.parametrize("p", range(1, 9))
def test_element_weight_closed_form(p):
    prob = small_problem(3, 2)
    prob.set_penalization(p)
    rng = np.random.default_rng(p)
    xi = rng.uniform(1e-3, 1, prob.mesh.n_vertices)
    w = prob.element_weights(xi)
    m = prob.minimum
    tri = prob.mesh.triangles
    closed = np.array([m + (1 - m) * 2.0 / ((p + 1) * (p + 2)) * _hsym(xi[t], p) for t in tri])
    assert np.abs(w - closed).max() < 1e-14


def test_lumped_mass_closed_form():
    nx, ny = 5, 3
    prob = small_problem(nx, ny, width=5 / 3, height=1.0)
    h2 = (1 / 3) ** 2
    m = prob.lumped_mass().reshape(ny + 1, nx + 1)
    expect = np.full((ny + 1, nx + 1), h2)
    expect[0, :] = expect[-1, :] = expect[:, 0] = expect[:, -1] = h2 / 2
    expect[0, 0] = expect[-1, -1] = h2 / 3   # corners (0,0) and (W,H): 2 triangles
    expect[0, -1] = expect[-1, 0] = h2 / 6   # corners (W,0) and (0,H): 1 triangle
    assert np.abs(m - expect).max() < 1e-15


def test_total_loads():
    body = {"center": (0.75, 0.5), "radius": 0.3, "value": (0.0, -2.0)}
    trac = [{"side": "Right", "center": 0.5, "length": 0.5, "value": (1.0, 0.5)}]
    for degree in (1, 2):
        prob = small_problem(12, 8, degree=degree, fixed=(), tractions=trac, body=body)
        # traction on y in [0.25, 0.75], h = 1/8: the P2 edge interpolant is t on the 4 covered
        # edges and t N_a on the two neighbouring edges, int N_a = h/6: total t (0.5 + 2 h/6)
        pt = small_problem(12, 8, degree=degree, fixed=(), tractions=trac)
        fx, fy = pt.load[0::2].sum(), pt.load[1::2].sum()
        covered = 0.5 + 2 * (1 / 8) / 6
        assert abs(fx - 1.0 * covered) < 1e-13 and abs(fy - 0.5 * covered) < 1e-13
        # body force: total = value * int of the P2 interpolant of the indicator
        pb = small_problem(12, 8, degree=degree, fixed=(), body=body)
        assert abs(pb.load[0::2].sum()) < 1e-14
        assert pb.load[1::2].sum() < 0


@pytest.mark.parametrize("degree", [1, 2])
@pytest.mark.parametrize("side,center,length", [("Right", 0.4, 0.3), ("Left", 0.1, 0.2),
                                                ("Top", 1.35, 0.3), ("Bottom", 0.7, 0.45)])
def test_traction_first_moments(degree, side, center, length):
    """The split of a traction over the nodes: sum_i F_i x_i and sum_i F_i y_i must equal
    int t_h x ds and int t_h y ds over the whole exterior boundary, where t_h is the P2 edge
    interpolant of the pointwise rule (FEM_src/elasisity_problem.py:39-73).  t_h times a linear
    function is a cubic on each boundary edge, so Simpson's rule over the edge's end points and
    midpoint integrates it exactly; the rule is evaluated here independently (exact side test,
    the interval widened by DOLFIN_EPS), on every boundary edge: a corner that the rule selects
    also loads the first edge of the neighbouring side (the Left and Top cases, whose intervals
    reach a corner).  A contribution booked on the wrong end node of an edge, or edges walked
    in the wrong direction, changes a moment while keeping the total: the partially covered
    edge at each end of an interval carries an asymmetry of +-t h / 6, which an interval with
    two partial ends cancels and one that reaches its side's end exposes."""
    nx, ny, width, height = 12, 8, 1.5, 1.0
    value = (0.7, -1.3)
    trac = [{"side": side, "center": center, "length": length, "value": value}]
    prob = small_problem(nx, ny, degree=degree, fixed=(), tractions=trac, width=width,
                         height=height)
    lo, hi = center - length / 2, center + length / 2
    eps = 3.0e-16

    def rule(x, y):
        on = {"Left": x == 0.0, "Right": x == width, "Top": y == height, "Bottom": y == 0.0}[side]
        t = y if side in ("Left", "Right") else x
        return 1.0 if on and lo - eps <= t <= hi + eps else 0.0

    hx, hy = width / nx, height / ny
    edges = [((i * hx, 0.0), ((i + 1) * hx, 0.0)) for i in range(nx)]
    edges += [((i * hx, height), ((i + 1) * hx, height)) for i in range(nx)]
    edges += [((0.0, j * hy), (0.0, (j + 1) * hy)) for j in range(ny)]
    edges += [((width, j * hy), (width, (j + 1) * hy)) for j in range(ny)]
    m0 = mx = my = 0.0
    for (ax, ay), (bx, by) in edges:
        L = abs(bx - ax) + abs(by - ay)
        cx, cy = 0.5 * ax + 0.5 * bx, 0.5 * ay + 0.5 * by
        ta, tc, tb = rule(ax, ay), rule(cx, cy), rule(bx, by)
        m0 += L / 6 * (ta + 4 * tc + tb)
        mx += L / 6 * (ta * ax + 4 * tc * cx + tb * bx)
        my += L / 6 * (ta * ay + 4 * tc * cy + tb * by)
    xs, ys = prob.node_coords[:, 0], prob.node_coords[:, 1]
    for comp in range(2):
        F = prob.load[comp::2]
        assert abs(F.sum() - value[comp] * m0) < 1e-13
        assert abs(F @ xs - value[comp] * mx) < 1e-13
        assert abs(F @ ys - value[comp] * my) < 1e-13


# ------------------------------------------------------------------ invariants
@pytest.mark.parametrize("degree", [1, 2])
def test_rigid_body_null_space(degree):
    prob = small_problem(4, 3, degree=degree, fixed=())
    prob.set_penalization(3)
    rng = np.random.default_rng(0)
    K = prob.stiffness(rng.uniform(0.1, 1, prob.mesh.n_vertices)).toarray()
    assert np.abs(K - K.T).max() < 1e-13 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert ev[0] > -1e-12 * ev[-1]
    assert np.sum(np.abs(ev) < 1e-10 * ev[-1]) == 3
    c = prob.node_coords
    for mode in (np.stack([np.ones(len(c)), np.zeros(len(c))], 1),
                 np.stack([np.zeros(len(c)), np.ones(len(c))], 1),
                 np.stack([-c[:, 1], c[:, 0]], 1)):
        assert np.abs(K @ mode.ravel()).max() < 1e-12


@pytest.mark.parametrize("degree", [1, 2])
def test_patch_test_linear_field(degree):
    prob = small_problem(6, 4, degree=degree, fixed=())
    prob.set_penalization(3)
    K = prob.stiffness(np.ones(prob.mesh.n_vertices))
    c = prob.node_coords
    u = np.stack([0.3 + 0.2 * c[:, 0] - 0.1 * c[:, 1], -0.2 + 0.05 * c[:, 0] + 0.4 * c[:, 1]], 1).ravel()
    f = K @ u
    nX, nY, _ = prob.node_grid
    jj, ii = np.meshgrid(np.arange(nY), np.arange(nX), indexing="ij")
    interior = ((ii > 0) & (ii < nX - 1) & (jj > 0) & (jj < nY - 1)).ravel()
    assert np.abs(f.reshape(-1, 2)[interior]).max() < 1e-12


def _naive_p1_stiffness(prob, xi, p):
    """Independent loop-based assembly (explicit B matrices per triangle)."""
    lam, mu = prob.lmbda, prob.mu
    D = np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]])
    n = prob.n_dofs
    K = np.zeros((n, n))
    m = prob.minimum
    for t in prob.mesh.triangles:
        (x0, y0), (x1, y1), (x2, y2) = prob.mesh.coords[t]
        area = 0.5 * abs((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
        det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
        b = np.array([y1 - y2, y2 - y0, y0 - y1]) / det
        cc = np.array([x2 - x1, x0 - x2, x1 - x0]) / det
        B = np.zeros((3, 6))
        for a in range(3):
            B[0, 2 * a] = b[a]
            B[1, 2 * a + 1] = cc[a]
            B[2, 2 * a] = cc[a]
            B[2, 2 * a + 1] = b[a]
        rbar = m + (1 - m) * 2.0 / ((p + 1) * (p + 2)) * _hsym(xi[t], p)
        Ke = rbar * area * B.T @ D @ B
        dofs = np.ravel([[2 * v, 2 * v + 1] for v in t])
        K[np.ix_(dofs, dofs)] += Ke
    return K


@pytest.mark.parametrize("nx,ny,p", [(2, 2, 3), (4, 3, 1), (3, 4, 5)])
def test_bruteforce_dense_p1(nx, ny, p):
    prob = small_problem(nx, ny, width=1.0 * nx / ny, height=1.0,
                         tractions=[{"side": "Right", "center": 0.5, "length": 0.6,
                                     "value": (0.0, -1.0)}])
    prob.set_penalization(p)
    xi = np.random.default_rng(3).uniform(0.05, 1, prob.mesh.n_vertices)
    Kd = _naive_p1_stiffness(prob, xi, p)
    Ks = prob.stiffness(xi).toarray()
    assert np.abs(Ks - Kd).max() < 1e-13 * np.abs(Kd).max()
    f = prob.free
    u = np.zeros(prob.n_dofs)
    u[f] = np.linalg.solve(Kd[np.ix_(f, f)], prob.load[f])
    ud = prob.forward(xi)
    assert np.linalg.norm(u - ud) < 1e-12 * np.linalg.norm(u)
    J = prob.compliance(ud)
    assert abs(J - ud @ (Ks @ ud)) < 1e-12 * J and J > 0


def test_compliance_decreases_with_density():
    prob = small_problem(12, 8, tractions=[{"side": "Right", "center": 0.5, "length": 0.2,
                                            "value": (0.0, -1.0)}])
    prob.set_penalization(3)
    rho = np.random.default_rng(5).uniform(0.2, 0.8, prob.mesh.n_vertices)
    J0 = prob.calculate_objective(rho)
    J1 = prob.calculate_objective(rho + 0.1)
    assert J1 < J0


# ------------------------------------------------------------------ gradient
@pytest.mark.parametrize("degree,p", [(1, 3), (1, 5), (2, 3)])
def test_gradient_matches_central_differences(degree, p):
    prob = small_problem(8, 5, degree=degree,
                         tractions=[{"side": "Right", "center": 0.5, "length": 0.4,
                                     "value": (0.3, -1.0)}],
                         body={"center": (0.9, 0.6), "radius": 0.25, "value": (0.0, -0.5)})
    prob.set_penalization(p)
    rng = np.random.default_rng(11)
    rho = rng.uniform(0.3, 0.9, prob.mesh.n_vertices)
    prob.calculate_objective(rho)
    g = prob.mass @ prob.calculate_objective_gradient()
    h = 1e-6
    for k in range(4):
        d = rng.standard_normal(rho.size)
        d /= np.linalg.norm(d)
        fd = (prob.calculate_objective(rho + h * d) - prob.calculate_objective(rho - h * d)) / (2 * h)
        assert abs(fd - g @ d) < 1e-6 * np.linalg.norm(g)


def test_gradient_requires_objective_first():
    prob = small_problem(4, 3)
    prob.set_penalization(3)
    with pytest.raises(ValueError):
        prob.calculate_objective_gradient()


def test_penalty_required():
    prob = small_problem(4, 3)
    with pytest.raises(ValueError):
        prob.calculate_objective(np.full(prob.mesh.n_vertices, 0.5))


def test_simp_reference_constants():
    lam, mu = lame(2.5, 0.25)
    assert mu == 2.5 / 2.5 and abs(lam - 1.0) < 1e-15  # lambda = mu nu / (0.5 - nu)
    assert simp(1.0, 3) == 1.0 and abs(simp(0.0, 3) - 1e-6) < 1e-22


def test_fixed_side_near_semantics():
    # vertex x = (W/nx)*ix: W=3, nx=60 gives exactly 3.0 at the right side
    mesh = StructuredMesh(3.0, 1.0, 60, 20)
    assert mesh.xs[-1] == 3.0
    prob = ElasticityOracle(mesh, young_modulus=1, poisson_ratio=0.3, filter_radius=0.05,
                            fixed_sides=["Left", "Right"])
    assert prob.side_is_fixed_coordinate("Right")
    assert len(prob.fixed_dofs) == 2 * 2 * 21


def test_oracle_direct_solver_residual():
    prob = small_problem(30, 10, width=3.0, height=1.0,
                         tractions=[{"side": "Right", "center": 0.5, "length": 0.1,
                                     "value": (0.0, -1.0)}])
    prob.set_penalization(3)
    xi = np.random.default_rng(2).uniform(1e-3, 1, prob.mesh.n_vertices)
    u = prob.forward(xi)
    A = prob.masked_operator(xi)
    b = prob.load.copy()
    b[prob.fixed_dofs] = 0
    assert np.linalg.norm(A @ u - b) < 1e-10 * np.linalg.norm(b)
