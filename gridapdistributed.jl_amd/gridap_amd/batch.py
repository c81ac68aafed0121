"""Batched numeric assembly of independent problems in one hipGraph launch.

    b = AssemblyBatch(ctx)
    for prob in problems: b.add(prob)        # GpuProblem after its symbolic phase
    b.autotune()                             # cells per workgroup per problem
    b.capture()
    b.launch(); b.sync()                     # every problem's values and RHS refilled
"""
import ctypes

import numpy as np

from . import _lib


class AssemblyBatch:
    def __init__(self, ctx):
        h = ctypes.c_void_p()
        _lib.call("gdx_batch_create", ctx.handle, ctypes.byref(h))
        self.handle, self.ctx, self.problems = h, ctx, []

    def add(self, prob):
        _lib.call("gdx_batch_add", self.handle, prob.handle)
        self.problems.append(prob)

    def autotune(self, reps=3):
        chosen = np.zeros(len(self.problems), np.int32)
        _lib.call("gdx_batch_autotune", self.handle, int(reps), chosen.ctypes.data)
        return chosen

    def capture(self):
        _lib.call("gdx_batch_capture", self.handle)

    def launch(self):
        _lib.call("gdx_batch_launch", self.handle)

    def sync(self):
        if self.problems:
            self.problems[0].sync()

    def close(self):
        if self.handle:
            _lib.call("gdx_batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
