"""The recurrence of the one-launch PCG iteration (k_pcg_fused), restated in
numpy, against the standard block-Jacobi PCG of linear_solver_pcg.hpp:112-155
on random SPD block systems.

Launch L >= 1 applies iteration L-1's update to x, r and s (s -= alpha w with
w = M^-1 q) and takes the new r.M^-1 r from dot products formed in launch
L-1: dn' = r.s - 2 alpha s.q + alpha^2 q.w. With M symmetric this is exactly
r'.s', so the iterates, the stop decisions and the iteration counts are the
standard ones; a dropped term or a wrong sign in the formula would change the
stop iteration or the solution by far more than rounding."""
import numpy as np
import pytest


def _system(rng, nb, bd=6, density=0.3, cond=1e4):
    n = nb * bd
    A = np.zeros((n, n))
    for i in range(nb):
        for j in range(i + 1, nb):
            if rng.random() < density:
                B = rng.normal(size=(bd, bd))
                A[i * bd:(i + 1) * bd, j * bd:(j + 1) * bd] = B
                A[j * bd:(j + 1) * bd, i * bd:(i + 1) * bd] = B.T
    # diagonally weighted so that the system is SPD with a spread spectrum
    scale = np.exp(rng.uniform(0, np.log(cond), n))
    A = A * np.sqrt(np.outer(scale, scale)) * 0.1
    A += np.diag(np.abs(A).sum(axis=1) + scale)
    Minv = np.zeros((nb, bd, bd))
    for i in range(nb):
        Minv[i] = np.linalg.inv(A[i * bd:(i + 1) * bd, i * bd:(i + 1) * bd])
    b = rng.normal(size=n)
    return A, Minv, b


def _precond(Minv, v):
    nb, bd, _ = Minv.shape
    return np.einsum("kij,kj->ki", Minv, v.reshape(nb, bd)).ravel()


def pcg_standard(A, Minv, b, tol, max_iter):
    """linear_solver_pcg.hpp:120-153: the stop test dn <= tol * dn0 on the
    directly summed r.M^-1 r after each update."""
    x = np.zeros_like(b)
    r = b.copy()
    s = _precond(Minv, r)
    dn = r @ s
    d0 = tol * dn
    d = np.zeros_like(b)
    dn_old = 1.0
    it = 0
    while it < max_iter and dn > d0:
        beta = 0.0 if it == 0 else dn / dn_old
        d = s + beta * d
        q = A @ d
        alpha = dn / (d @ q)
        x += alpha * d
        r -= alpha * q
        s = _precond(Minv, r)
        dn_old, dn = dn, r @ s
        it += 1
    return x, it, dn


def pcg_one_launch(A, Minv, b, tol, max_iter):
    """The launch sequence of k_pcg_fused."""
    x = np.zeros_like(b)
    r = b.copy()
    s = _precond(Minv, r)
    rs = r @ s  # k_pcg_init's partials
    dn = rs
    d0 = tol * dn
    it = 0
    if dn <= d0:
        return x, 0, dn, []
    d = s.copy()
    q = A @ d
    w = _precond(Minv, q)
    DQ, SQ, QW, RS = d @ q, s @ q, q @ w, r @ s
    formula_vs_direct = []
    while True:
        it += 1
        alpha = RS / DQ
        dn = (RS - 2.0 * alpha * SQ) + alpha * alpha * QW
        x += alpha * d
        r -= alpha * q
        s -= alpha * w
        formula_vs_direct.append((dn, r @ _precond(Minv, r)))
        if dn <= d0 or it == max_iter:
            return x, it, dn, formula_vs_direct
        beta = dn / RS
        d = s + beta * d
        q = A @ d
        w = _precond(Minv, q)
        DQ, SQ, QW, RS = d @ q, s @ q, q @ w, r @ s


@pytest.mark.parametrize("seed,tol", [(1, 1e-6), (2, 1e-10), (3, 1e-14), (4, 1e-8)])
def test_one_launch_recurrence_is_the_standard_pcg(seed, tol):
    rng = np.random.default_rng(seed)
    A, Minv, b = _system(rng, nb=40)
    n = b.size
    x_s, it_s, dn_s = pcg_standard(A, Minv, b, tol, n)
    x_f, it_f, dn_f, pairs = pcg_one_launch(A, Minv, b, tol, n)
    assert it_s > 3
    assert it_f == it_s
    assert np.linalg.norm(x_f - x_s) <= 1e-10 * np.linalg.norm(x_s)
    assert abs(dn_f - dn_s) <= 1e-8 * abs(dn_s) + 1e-300
    # the formula equals r'.M^-1 r' computed from the updated vectors
    for f, direct in pairs:
        assert abs(f - direct) <= 1e-9 * max(abs(direct), 1e-300) + 1e-12 * abs(pairs[0][1])
    # and solves the system to the tolerance asked for
    res = b - A @ x_f
    assert res @ _precond(Minv, res) <= 1.01 * tol * (b @ _precond(Minv, b))


def test_a_wrong_formula_is_caught():
    """A plausible slip (dropping the alpha^2 q.w term, or +2 alpha s.q)
    changes the stop iteration: the checks above would fail."""
    rng = np.random.default_rng(7)
    A, Minv, b = _system(rng, nb=30)
    _, it_s, _ = pcg_standard(A, Minv, b, 1e-8, b.size)
    x = np.zeros_like(b)
    r = b.copy()
    s = _precond(Minv, r)
    rs = r @ s
    d0 = 1e-8 * rs
    d = s.copy()
    q = A @ d
    w = _precond(Minv, q)
    DQ, SQ, RS = d @ q, s @ q, r @ s
    it = 0
    while True:
        it += 1
        alpha = RS / DQ
        dn = RS - 2.0 * alpha * SQ  # the alpha^2 q.w term dropped
        x += alpha * d
        r -= alpha * q
        s -= alpha * w
        if dn <= d0 or it == b.size:
            break
        d = s + (dn / RS) * d
        q = A @ d
        w = _precond(Minv, q)
        DQ, SQ, RS = d @ q, s @ q, r @ s
    assert it != it_s
