"""Many independent scenes stepped together (cipc_batch_step): one implicit-Euler step of
every scene per call, host threads taking scenes from a shared queue, each scene's kernels
on its own stream.  Reports per-scene step latency."""
import ctypes

import numpy as np

from . import _lib
from .sim import Simulator, CipcError


class BatchStepper:
    def __init__(self, sims, threads=0):
        self.sims = list(sims)
        self.threads = int(threads)
        self._h = (ctypes.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.lib = _lib.load()

    def step(self):
        """One step of every scene; returns (list of per-scene stats dicts, per-scene ms)."""
        n = len(self.sims)
        st = (_lib.StepStats * n)()
        ms = np.zeros(n)
        rc = self.lib.cipc_batch_step(ctypes.cast(self._h, ctypes.POINTER(ctypes.c_void_p)), n, self.threads,
                                      ctypes.cast(st, ctypes.c_void_p), ms.ctypes.data_as(_lib.D))
        if rc != 0:
            bad = [i for i in range(n) if st[i].status != 0]
            raise CipcError(rc, "batch step failed for scenes %s" % bad[:8])
        return [s.as_dict() for s in st], ms
