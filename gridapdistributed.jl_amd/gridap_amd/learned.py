"""Learned multilevel preconditioner and flexible PCG (device), over the C ABI.

    net = LearnedCoarseSolver(ctx, n, seed)     # n^3 network grid: problems of (2n)^3 Q1 cells
    net.attach(prob)                            # after every numeric assembly of prob
    loss = net.train_step(seed, lr)             # one Adam step on a sample from prob's matrix
    x, it, relres, ms, fellback = fpcg(prob, net, rhs, rtol)   # net=None: Jacobi-PCG
"""
import ctypes

import numpy as np

from . import _lib


class LearnedCoarseSolver:
    """BF16 convolutional multilevel network (channels 8/16/24, ~60k parameters)
    used as the coarse-grid guess of a multilevel preconditioner (Galerkin
    V-cycle below it); trained on device."""

    def __init__(self, ctx, n, seed=0):
        h = ctypes.c_void_p()
        _lib.call("gdx_nn_create", ctx.handle, int(n), int(seed), ctypes.byref(h))
        self.handle, self.n, self.ctx = h, int(n), ctx
        npar = ctypes.c_int64()
        _lib.call("gdx_nn_num_params", self.handle, ctypes.byref(npar))
        self.num_params = npar.value

    def get_params(self):
        th = np.empty(self.num_params, np.float32)
        _lib.call("gdx_nn_save", self.handle, th.ctypes.data)
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        if th.shape != (self.num_params,):
            raise ValueError(f"expected {self.num_params} parameters")
        _lib.call("gdx_nn_load", self.handle, th.ctypes.data)

    def forward(self, inp):
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        if inp.shape != (2, self.n, self.n, self.n):
            raise ValueError("input must be [2][n][n][n]")
        out = np.empty((1, self.n, self.n, self.n), np.float32)
        _lib.call("gdx_nn_forward", self.handle, inp.ctypes.data, out.ctypes.data)
        return out

    def backward(self, dout, want_dinp=False):
        dout = np.ascontiguousarray(dout, dtype=np.float32)
        g = np.empty(self.num_params, np.float32)
        dinp = np.empty((2, self.n, self.n, self.n), np.float32) if want_dinp else None
        _lib.call("gdx_nn_backward", self.handle, dout.ctypes.data, g.ctypes.data,
                  None if dinp is None else dinp.ctypes.data)
        return g, dinp

    def attach(self, prob):
        _lib.call("gdx_nn_attach", self.handle, prob.handle)

    def train_step(self, seed, lr=1e-3, smooth_passes=8):
        loss = ctypes.c_float()
        _lib.call("gdx_nn_train_step", self.handle, int(seed), float(lr), int(smooth_passes), ctypes.byref(loss))
        return loss.value

    def set_options(self, coarse_sweeps=8, use_network=True):
        """Coarse damped-Jacobi sweeps after the network; use_network=False is
        the same preconditioner with the coarse sweeps started from zero."""
        _lib.call("gdx_nn_set_options", self.handle, int(coarse_sweeps), 1 if use_network else 0)

    def set_multigrid(self, enable=True, nu=1):
        """V-cycle coarse correction (default) or the two-grid coarse sweeps;
        returns the number of grids of the V-cycle."""
        lv = ctypes.c_int32()
        _lib.call("gdx_nn_set_multigrid", self.handle, 1 if enable else 0, int(nu), ctypes.byref(lv))
        return lv.value

    def level_operator(self, level):
        """Galerkin operator of V-cycle level `level` (0 = network grid): stencil [27][m][m][m]."""
        m = ctypes.c_int32()
        _lib.call("gdx_nn_level_operator", self.handle, int(level), ctypes.byref(m), None)
        st = np.empty((27, m.value, m.value, m.value))
        _lib.call("gdx_nn_level_operator", self.handle, int(level), ctypes.byref(m), st.ctypes.data)
        return st

    def coarse_operator(self):
        """Galerkin coarse operator of the attached matrix, stencil [27][n][n][n]."""
        st = np.empty((27, self.n, self.n, self.n))
        _lib.call("gdx_nn_coarse_operator", self.handle, st.ctypes.data)
        return st

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        _lib.call("gdx_nn_apply", self.handle, r.ctypes.data, z.ctypes.data)
        return z

    def close(self):
        if self.handle:
            _lib.call("gdx_nn_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fpcg(prob, net=None, rhs=None, rtol=1e-8, maxiter=20000):
    """Flexible PCG with the learned preconditioner (net attached to prob), or
    Jacobi-PCG (net=None).  Returns (x, iterations, true relative residual, ms, fell_back);
    ms is the solve's wall time from rhs resident on the device to the solution
    there (the host copies of rhs and x are outside it)."""
    x = np.empty(prob.ngdofs)
    r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
    it, rr, ms, fb = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
    _lib.call("gdx_fpcg", prob.handle, None if net is None else net.handle, None if r is None else r.ctypes.data,
              float(rtol), int(maxiter), x.ctypes.data, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(ms),
              ctypes.byref(fb))
    return x, it.value, rr.value, ms.value, bool(fb.value)
