"""Python mirror of the batch C ABI: many fields of uneven shapes corrected
together (one launch per pass and axis length over all fields, per-field
convergence flags on the device, the iteration replayed as a HIP graph).

The reference has no batch interface; each field keeps the single-field
``ProjectionSolver`` API (initialize*, getters, extraction, certification).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FFCZ_F32, FFCZ_F64, call
from .solver import CertResult, ProjectionSolver, _np_dtype, as_pointer


class ProjectionBatch:
    def __init__(self, shapes, dtype=np.float32):
        self.dtype = _np_dtype(dtype)
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        if not self.shapes:
            raise ValueError("a batch needs at least one field")
        dims = (C.c_size_t * (3 * len(self.shapes)))(*[v for s in self.shapes for v in s])
        h = C.c_void_p()
        call("ffcz_batch_create", FFCZ_F32 if self.dtype == np.float32 else FFCZ_F64, len(self.shapes), dims,
             C.byref(h))
        self._h = h
        self.fields = []
        for i, s in enumerate(self.shapes):
            fh = C.c_void_p()
            call("ffcz_batch_field", self._h, i, C.byref(fh))
            self.fields.append(ProjectionSolver._borrowed(fh, s, self.dtype, self))

    def __len__(self):
        return len(self.fields)

    def __getitem__(self, i) -> ProjectionSolver:
        return self.fields[i]

    def solve(self, mode: int, spatial_errors, max_iterations: int, convergence_tolerance: float,
              use_graph: bool = True):
        """The projection loop of every field (no extraction); ``spatial_errors``
        holds one initial error x' - x per field (numpy or torch)."""
        if len(spatial_errors) != len(self.fields):
            raise ValueError("one initial error per field")
        keep = []
        devices = set()
        ptrs = (C.c_void_p * len(self.fields))()
        for i, (f, e) in enumerate(zip(self.fields, spatial_errors)):
            p, k = as_pointer(e, self.dtype, sync=False)
            if int(np.prod(k.shape)) != f.N:
                raise ValueError(f"initial error {i} has the wrong size")
            keep.append(k)
            ptrs[i] = p.value
            if getattr(k, "is_cuda", False):
                devices.add(k.device)
        if devices:
            import torch

            # the library works on its own non-blocking streams: one wait per device
            for d in devices:
                torch.cuda.current_stream(d).synchronize()
        call("ffcz_batch_solve", self._h, int(mode), ptrs, int(max_iterations), float(convergence_tolerance),
             int(bool(use_graph)))
        del keep

    def status(self):
        """(iterations[], violations[], converged[], iterations the batch ran)."""
        n = len(self.fields)
        it = (C.c_int * n)()
        vi = (C.c_double * n)()
        cv = (C.c_int * n)()
        launched = C.c_int()
        call("ffcz_batch_status", self._h, it, vi, cv, n, C.byref(launched))
        return list(it), list(vi), [bool(c) for c in cv], launched.value

    def extract_all(self):
        """extract_compact_representation of every field: one launch of each
        extraction kernel over all fields' edit arrays."""
        call("ffcz_batch_extract", self._h)

    def certify_all(self, quantized: bool):
        """FP64 certification of every field (one launch per pass and axis
        length) -> [CertResult]."""
        n = len(self.fields)
        out = (_lib.Cert * n)()
        call("ffcz_batch_certify", self._h, int(bool(quantized)), out, n)
        return [CertResult(bool(r.ok), r.spatial_violations, r.freq_violations, r.spatial_max_excess,
                           r.freq_max_excess) for r in out]

    def time_iteration(self, reps: int = 10) -> float:
        """Mean ms of one whole batched iteration after a completed solve
        (reset the edit accumulators afterwards)."""
        ms = C.c_float()
        call("ffcz_batch_time_iteration", self._h, int(reps), C.byref(ms))
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            for f in self.fields:
                f._h = None
            call("ffcz_batch_destroy", self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
