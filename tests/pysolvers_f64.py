"""Independent float64 restatement of the step-doubling integrators
(numpy, one system at a time): SIE take_step with a dense solve of
(I - hJ) x = f (sie_gold.c:38-79 with gje replaced by numpy.linalg.solve),
RK2 take_step (rk2_gold.c:18-46), and the adaptive controller (accept if
err <= 1, keep 2 y2 - y1, h *= clamp(0.9 / sqrt(err), 0.2, 4)).  The float64
cross-check of the oracle's double build, in the role of the reference's
pure-numpy pysolvers.  Test infrastructure only."""
import numpy as np


def take_step(method, rhs, jac, y, h, n):
    y = y.copy()
    hs = h / n
    for _ in range(n):
        f = rhs(y)
        if method == 0:
            A = np.eye(len(y)) - hs * jac(y)
            y = y + hs * np.linalg.solve(A, f)
        else:
            y = y + hs * f
    return y


def integrate_adaptive(method, rhs, jac, y0, tnow, tend, nfix, abs_tol, rel_tol, max_attempts=100000):
    y = np.array(y0, np.float64)
    h = (tend - tnow) / nfix
    t = tnow
    attempts = 0
    while t < tend and attempts < max_attempts:
        hh = min(tend - t, h)
        y1 = take_step(method, rhs, jac, y, hh, 1)
        y2 = take_step(method, rhs, jac, y, hh, 2)
        err = np.max(np.abs(y2 - y1) / (abs_tol + rel_tol * np.abs(y2)))
        attempts += 1
        if err <= 1.0:
            y = 2.0 * y2 - y1
            t = t + hh
        h = hh * min(4.0, max(0.2, 0.9 / np.sqrt(err))) if err > 0 else hh * 4.0
    return y, 3 * attempts
