"""Conjugate-gradient solver of the reference -- TEST INFRASTRUCTURE.

A numpy restatement of Claerbout's cgstep (base/cgstep.cc:11-54) driven by
simple_solver / lin_solver::solve (base/lin_solver.cc:6-15, 73-98 for the
initial residual, 104-131 for the iteration).  Scalars are FP64, as in the
reference; vectors keep the dtype they are given (float32 reproduces the
reference's hypercube_float vectors, whose scale_add evaluates
v * s1 + s2 * o in double and stores float, hypercube_float.cc:213-222, and
whose dot accumulates in double, hypercube_float.cc:120-130).

Only tests/ use this module, as the checker of the shipped solver
(sc17-born_amd/host/born_op.cc).  Pinned in tests/test_oracle_cg.py against
textbook CGLS on explicit least-squares problems and textbook CG on an SPD
system (through its symmetric square root).
"""
import numpy as np


def dot(a, b):
    """FP64 accumulation of an elementwise product (hypercube_float::dot)"""
    return float(np.dot(np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)))


def _scale_add(v, sc_me, o, sc_o):
    # v = v * sc_me + sc_o * o, evaluated in double, stored in v's dtype
    v[...] = (v.astype(np.float64) * sc_me + sc_o * o.astype(np.float64)).astype(v.dtype)


class CGStep:
    """cgstep (base/cgstep.h, cgstep.cc): s and ss are the previous model and
    data steps; the first step is a steepest-descent step (forget)."""

    small = 1e-20

    def __init__(self, m, rr):
        self.s = np.zeros_like(m)
        self.ss = np.zeros_like(rr)
        self.forget = True

    def step(self, m, g, rr, gg):
        """One step: updates m and rr in place; returns (ok, |rr|^2)."""
        if self.forget:
            beta = 0.0
            gg_gg = dot(gg, gg)
            if abs(gg_gg) < self.small:
                return False, None
            alfa = -dot(gg, rr) / gg_gg
        else:
            gg_gg = dot(gg, gg)
            ss_ss = dot(self.ss, self.ss)
            gg_ss = dot(gg, self.ss)
            if abs(gg_gg) < self.small or abs(ss_ss) < self.small:
                return False, None
            determ = gg_gg * ss_ss * (1.0 - gg_ss / gg_gg * gg_ss / ss_ss)
            gg_rr = -dot(gg, rr)
            ss_rr = -dot(self.ss, rr)
            alfa = (ss_ss * gg_rr - gg_ss * ss_rr) / determ
            beta = (-gg_ss * gg_rr + gg_gg * ss_rr) / determ
        _scale_add(self.s, beta, g, alfa)
        _scale_add(self.ss, beta, gg, alfa)
        m += self.s
        rr += self.ss
        self.forget = False
        return True, dot(rr, rr)


def simple_solver(forward, adjoint, data, niter, m_shape, m0=None):
    """simple_solver + lin_solver::solve without weights: rr = F m0 - d, then
    per iteration g = F^T rr, gg = F g and one cgstep.  forward(m) and
    adjoint(d) return new arrays (add = false).  Returns the model and the
    residual |rr|^2 after each iteration."""
    rr = -np.array(data, copy=True)
    if m0 is not None:
        rr += forward(m0).astype(rr.dtype)
    m = np.zeros(m_shape, dtype=rr.dtype)
    st = CGStep(m, rr)
    res = []
    for _ in range(niter):
        g = np.asarray(adjoint(rr), dtype=rr.dtype).reshape(m_shape)
        gg = np.asarray(forward(g), dtype=rr.dtype).reshape(rr.shape)
        ok, val = st.step(m, g, rr, gg)
        if not ok:
            break
        res.append(val)
    return m, res
