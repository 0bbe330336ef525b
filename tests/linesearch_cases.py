"""Line-search cases whose outcome follows in closed form.

With r0 = RHS - A u and w = A du, the trial residual of lineSearch
(src/FEM3D.jl:123-152) is r(alpha) = r0 - alpha w, so
||r(alpha)||^2 - ||r0||^2 = alpha (alpha ||w||^2 - 2 r0.w): a trial is
rejected (residue > residual_old) exactly when alpha > alpha* = 2 r0.w / ||w||^2.
Scaling du by c moves alpha* to alpha*/c, so a case can put alpha* between
two chosen powers of the step factor and the accepted step is known before
the search runs. The expected alpha is built by the same repeated
multiplication by 0.9 as the search (floating-point powers differ in the
last bit), and every case keeps alpha* at least 2% away from each trial.
"""
import numpy as np

FACTOR = 0.9


def trial_alphas(min_step):
    """1, 0.9, 0.9*0.9, ... up to the first value not above min_step."""
    a = [1.0]
    while a[-1] > min_step:
        a.append(a[-1] * FACTOR)
    return a


def make_case(S, rng, alpha_star, nv):
    """u, du, RHS with 2 r0.w/||w||^2 == alpha_star (S: scipy CSR matrix).
    alpha_star < 0 makes du an ascent direction."""
    u = rng.normal(size=nv)
    RHS = rng.normal(size=nv)
    du = rng.normal(size=nv)
    r0 = RHS - S @ u
    w = S @ du
    a0 = 2.0 * (r0 @ w) / (w @ w)
    if alpha_star < 0:
        if a0 > 0:
            du = -du
            a0 = -a0
        du *= a0 / alpha_star
    else:
        if a0 < 0:
            du = -du
            a0 = -a0
        du *= a0 / alpha_star
    return u, du, RHS


def expected(S, u, du, RHS, min_step):
    """(k, alpha, residue, u after the search) from the closed form: the
    first trial alpha_k <= alpha* ends the loop (or the first alpha_k <=
    min_step); u moves to u + alpha_k du unless alpha_k < min_step, and then
    k is None."""
    r0 = RHS - S @ u
    w = S @ du
    astar = 2.0 * (r0 @ w) / (w @ w)
    alphas = trial_alphas(min_step)
    for a in alphas:
        assert abs(a - astar) > 0.02 * a, "case too close to a trial step"
    k = next((i for i, a in enumerate(alphas) if a <= astar), len(alphas) - 1)
    a = alphas[k]
    residue = float(np.linalg.norm(r0 - a * w))
    if a < min_step:
        return None, a, residue, u.copy()
    return k, a, residue, u + a * du


# (alpha*, min_step, first accepted k or None for "u unchanged")
CASES = [
    (FACTOR ** 7.5, 1e-4, 8),        # cut back eight times
    (1.7, 1e-4, 0),                  # the full step is accepted
    (FACTOR ** 40.5, 1e-4, 41),      # deep backtracking, still above min_step
    (-0.3, 1e-4, None),              # ascent direction: runs to alpha < min_step, u stays
    (9.7e-5, 1e-4, None),            # residual drops at 0.9^88 < min_step: still no update
    (FACTOR ** 7.5, 0.5, None),      # coarse min_step: 0.9^7 = 0.478 is still rejected and ends the loop
]
