"""CPU oracle of topomax's entropic-mirror-descent loop (test infrastructure only).

Restates ``src/solver.py:26-40`` (expit/logit), ``:149-186`` (volume projection:
Newton on int expit(psi + c) = V, Brent fallback via ``smart_brentq``,
``src/utils.py:138-157``), ``:188-206`` (step, tolerance, step size) and
``:208-302`` (the loop and its exit rules), driving an oracle problem from
``oracle.fem``.  Used only to reproduce the reference's golden run
(``tests/test_elasticity_solver.py:30-55``).
"""

from __future__ import annotations

import numpy as np
from scipy import optimize


def expit(x):
    return 1.0 / (1.0 + np.exp(-x))


def expit_diff(x):
    e = expit(x)
    return e * (1 - e)


def logit(x):
    return np.log(x / (1.0 - x))


def smart_brentq(f, initial_radius, max_radius):
    r = initial_radius
    while True:
        if r > max_radius:
            raise ValueError("f(-max_radius) and f(max_radius) must have different signs!")
        try:
            return optimize.brentq(f, -r, r, full_output=True)
        except ValueError:
            r *= 2


def project(integrate, half_step, volume):
    def error(c):
        return integrate(expit(half_step + c)) - volume

    def error_derivative(c):
        return integrate(expit_diff(half_step + c))

    try:
        c, result = optimize.newton(error, 0, error_derivative, tol=1e-12, full_output=True)
        if result.converged:
            return half_step + c
    except RuntimeError:
        pass
    c, result = smart_brentq(error, 2, 2000)
    if not result.converged:
        raise ValueError("Projection failed to converge")
    return half_step + c


def run_emd(problem, *, width, height, volume_fraction, step_size, penalties,
            max_iterations=1000, rho0=None):
    """Run the EMD loop; returns a list of per-penalty dicts with the objective
    history, the exit condition, the final design and iteration count."""
    volume = width * height * volume_fraction
    rho = np.full(problem.mesh.n_vertices, volume_fraction) if rho0 is None else rho0.copy()
    psi = logit(rho)
    results = []

    def tolerance(k):
        return min(25 * (k + 1) * 1e-5, 1e-2)

    def step_size_at_iter(k):
        if len(penalties) > 1:
            return step_size * min((k + 1), 10)
        return step_size * (k + 1)

    for penalty in penalties:
        problem.set_penalization(penalty)
        objectives = [problem.calculate_objective(rho)]
        exit_condition = ""
        k = 0
        for k in range(max_iterations):
            previous_psi = psi.copy()
            try:
                grad = problem.calculate_objective_gradient()
                psi = project(problem.integrate, previous_psi - step_size_at_iter(k) * grad, volume)
            except ValueError as e:
                exit_condition = str(e)
                break
            rho = expit(psi)
            objectives.append(problem.calculate_objective(rho))
            if np.isnan(objectives[-1]):
                exit_condition = "Objective is NaN"
                break
            min_index = int(np.argmin(objectives))
            if objectives[-1] > 2 * objectives[min_index]:
                exit_condition = "Objective is increasing"
                break
            if k >= min_index + 50:
                exit_condition = "Objective is not decreasing"
                break
            difference = np.sqrt(problem.integrate((rho - expit(previous_psi)) ** 2))
            if difference < tolerance(k):
                exit_condition = "Convergence treshold reached"
                break
        else:
            exit_condition = "Iteration did not converge"
        results.append({"penalty": penalty, "objectives": objectives, "iteration": k + 1,
                        "exit_condition": exit_condition, "rho": rho.copy()})
    return results
