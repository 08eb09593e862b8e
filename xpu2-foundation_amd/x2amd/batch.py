"""Batched render jobs (the x2batch_* entry points): many independent frames in one persistent kernel fed from a
device queue, replayed from a captured hipGraph, its shape chosen by the library's autotuner; or, for comparison,
one launch per job."""
import ctypes

import numpy as np

from . import abi


class Batch:
    def __init__(self, device=0):
        self.lib = abi.load()
        h = ctypes.c_void_p()
        abi.check(self.lib.x2batch_create(device, ctypes.byref(h)), "x2batch_create")
        self.h = h
        self.scenes = []  # the scenes' host arrays stay alive with the batch
        self.configs = []

    def close(self):
        if self.h:
            self.lib.x2batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, scene, config):
        job = ctypes.c_uint32()
        abi.check(self.lib.x2batch_add_job(self.h, ctypes.byref(scene.c), ctypes.byref(config), ctypes.byref(job)),
                  "x2batch_add_job")
        self.scenes.append(scene)
        self.configs.append(config)
        return job.value

    def tune(self):
        t = abi.BatchTuning()
        abi.check(self.lib.x2batch_tune(self.h, ctypes.byref(t)), "x2batch_tune")
        return t

    def set_tuning(self, block_threads, claim_items, waves_per_simd):
        t = abi.BatchTuning(block_threads, claim_items, waves_per_simd, 0, 0.0)
        abi.check(self.lib.x2batch_set_tuning(self.h, ctypes.byref(t)), "x2batch_set_tuning")

    def run(self, per_job_launches=False):
        """Renders every job once; returns (stats, per-job completion seconds from the run's start)."""
        n = len(self.configs)
        secs = (ctypes.c_double * max(1, n))()
        st = abi.BatchStats()
        abi.check(self.lib.x2batch_run(self.h, 1 if per_job_launches else 0, secs, ctypes.byref(st)), "x2batch_run")
        return st, np.array(secs[:n])

    def film(self, job):
        s, c = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_uint32)()
        abi.check(self.lib.x2batch_job_film(self.h, job, ctypes.byref(s), ctypes.byref(c)), "x2batch_job_film")
        cfg = self.configs[job]
        n = cfg.width * cfg.height
        return (np.ctypeslib.as_array(s, shape=(n, 4)).copy(), np.ctypeslib.as_array(c, shape=(n,)).copy())
