"""Pins of the CG oracle (oracle/cg.py, a restatement of base/cgstep.cc:11-54
and base/lin_solver.cc:104-131) against textbook methods it must equal in
exact arithmetic:

- CGLS (Hestenes-Stiefel CG on the normal equations): Claerbout's two-term
  plane search minimises |F m - d| over the same Krylov space, so the
  iterates agree step by step, and after n steps on n unknowns the model is
  the least-squares solution (numpy's lstsq);
- textbook CG on an SPD system A x = b: with F = A^(1/2) (symmetric square
  root) and d = A^(-1/2) b, the normal equations are A x = b, so the cgstep
  iterates are the CG iterates.

A wrong sign of alfa or beta, a transposed 2x2 solve or a dropped term fails
the step-by-step comparison at the second iteration or earlier.
"""
import numpy as np
import pytest

from oracle.cg import CGStep, simple_solver


def cgls(A, d, niter):
    """textbook CGLS, x0 = 0"""
    x = np.zeros(A.shape[1])
    r = d.copy()
    s = A.T @ r
    p = s.copy()
    gamma = s @ s
    xs = []
    for _ in range(niter):
        q = A @ p
        alpha = gamma / (q @ q)
        x = x + alpha * p
        r = r - alpha * q
        s = A.T @ r
        gnew = s @ s
        p = s + (gnew / gamma) * p
        gamma = gnew
        xs.append(x.copy())
    return xs


def cg(A, b, niter):
    """textbook CG for SPD A, x0 = 0"""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rs = r @ r
    xs = []
    for _ in range(niter):
        Ap = A @ p
        alpha = rs / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        rnew = r @ r
        p = r + (rnew / rs) * p
        rs = rnew
        xs.append(x.copy())
    return xs


def _iterates(F, d, niter):
    """the oracle solver's model after each iteration"""
    out = []
    for k in range(1, niter + 1):
        m, res = simple_solver(lambda v: F @ v, lambda r: F.T @ r, d, k, (F.shape[1],))
        out.append(m)
    return out, res


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cgstep_iterates_equal_cgls(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((40, 12))
    d = rng.standard_normal(40)
    ours, res = _iterates(A, d, 12)
    ref = cgls(A, d, 12)
    for k, (a, b) in enumerate(zip(ours, ref)):
        assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b), k
    x_ls = np.linalg.lstsq(A, d, rcond=None)[0]
    assert np.linalg.norm(ours[-1] - x_ls) <= 1e-9 * np.linalg.norm(x_ls)
    # reported residuals are |F m_k - d|^2
    for k, m in enumerate(ours):
        assert abs(res[k] - np.sum((A @ m - d) ** 2)) <= 1e-9 * res[k]


def test_cgstep_equals_textbook_cg_on_spd_system():
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((10, 10)))
    lam = np.linspace(1.0, 30.0, 10)
    A = (Q * lam) @ Q.T
    F = (Q * np.sqrt(lam)) @ Q.T           # symmetric square root of A
    b = rng.standard_normal(10)
    d = (Q * (1 / np.sqrt(lam))) @ Q.T @ b  # F^T d = b
    ours, _ = _iterates(F, d, 10)
    ref = cg(A, b, 10)
    for k, (a, r) in enumerate(zip(ours, ref)):
        assert np.linalg.norm(a - r) <= 1e-9 * np.linalg.norm(r), k
    assert np.linalg.norm(A @ ours[-1] - b) <= 1e-9 * np.linalg.norm(b)


def test_first_step_is_steepest_descent():
    # cgstep.cc:16-25: beta = 0, alfa = -(gg . rr) / (gg . gg)
    rng = np.random.default_rng(3)
    A = rng.standard_normal((20, 6))
    d = rng.standard_normal(20)
    m, res = simple_solver(lambda v: A @ v, lambda r: A.T @ r, d, 1, (6,))
    g = A.T @ (-d)
    gg = A @ g
    alfa = -(gg @ -d) / (gg @ gg)
    assert np.allclose(m, alfa * g, rtol=1e-12)
    assert np.isclose(res[0], np.sum((A @ m - d) ** 2), rtol=1e-12)


def test_float32_vectors_keep_their_dtype_and_fp64_scalars():
    # hypercube_float vectors: float storage, double dot and scale_add
    rng = np.random.default_rng(4)
    A = rng.standard_normal((30, 8))
    d = rng.standard_normal(30).astype(np.float32)
    A32 = A.astype(np.float32)
    m, res = simple_solver(lambda v: A32 @ v, lambda r: A32.T @ r, d, 8, (8,))
    assert m.dtype == np.float32
    x_ls = np.linalg.lstsq(A32.astype(np.float64), d.astype(np.float64), rcond=None)[0]
    assert np.linalg.norm(m - x_ls) <= 1e-4 * np.linalg.norm(x_ls)


def test_step_refuses_a_zero_gradient():
    st = CGStep(np.zeros(3), np.ones(4))
    ok, _ = st.step(np.zeros(3), np.zeros(3), np.ones(4), np.zeros(4))
    assert not ok
