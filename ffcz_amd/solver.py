"""Python mirror of the reference ``ProjectionSolver<T>``
(projection_algorithm.cuh:238-333) over the C ABI.

Method names, argument meaning and error behaviour follow the reference: the
same ``initialize*`` / ``solve*`` / getter surface, ``solve*`` runs the
extraction at the end (projection_algorithm_impl.cuh:813,857,901), repeated
solves keep accumulating edits until ``reset_edit_accumulators``, and
``solve*`` before ``initialize*`` raises ``"Solver not initialized"``
(projection_algorithm_impl.cuh:778-781).  Inputs may be numpy arrays (host) or
torch tensors (host or GPU; GPU tensors are read in place).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (FFCZ_F32, FFCZ_F64, FFT_BF16X3, FFT_FP16X3, FFT_VALU, FFT_VARIANTS, MODE_ABS, MODE_DIRECT,  # noqa: F401
                   MODE_PARTIAL_PTW, MODE_PTW, QUANT_REF16, QUANT_STEP, call)


def _np_dtype(dtype) -> np.dtype:
    d = np.dtype(dtype)
    if d not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError("ProjectionSolver supports float32 and float64")
    return d


def as_pointer(a, dtype=None, sync=True):
    """(void* address, keep-alive object) for a numpy array or torch tensor.
    sync=False leaves the wait for torch's current stream on a device tensor
    to the caller (one wait per device for many tensors)."""
    if a is None:
        return None, None
    try:
        import torch

        if isinstance(a, torch.Tensor):
            if not a.is_contiguous():
                a = a.contiguous()
            if dtype is not None and a.dtype != getattr(torch, np.dtype(dtype).name):
                raise TypeError(f"tensor dtype {a.dtype} does not match {np.dtype(dtype)}")
            if a.is_cuda and sync:
                # the library works on its own non-blocking stream
                torch.cuda.current_stream(a.device).synchronize()
            return C.c_void_p(a.data_ptr()), a
    except ImportError:  # pragma: no cover
        pass
    arr = np.ascontiguousarray(a, dtype=dtype)
    return C.c_void_p(arr.ctypes.data), arr


@dataclass
class CertResult:
    ok: bool
    spatial_violations: int
    freq_violations: int
    spatial_max_excess: float
    freq_max_excess: float


class ProjectionSolver:
    """``ProjectionSolver<T>(Nx, Ny, Nz)`` on the current HIP device."""

    def __init__(self, Nx: int, Ny: int, Nz: int, dtype=np.float32):
        self.dtype = _np_dtype(dtype)
        self._sfx = "f32" if self.dtype == np.float32 else "f64"
        self.shape = (int(Nx), int(Ny), int(Nz))
        self.N = self.shape[0] * self.shape[1] * self.shape[2]
        self.F = self.shape[0] * self.shape[1] * (self.shape[2] // 2 + 1)
        self._owner = None
        h = C.c_void_p()
        call("ffcz_create", FFCZ_F32 if self.dtype == np.float32 else FFCZ_F64,
             self.shape[0], self.shape[1], self.shape[2], C.byref(h))
        self._h = h

    @classmethod
    def _borrowed(cls, handle, shape, dtype, owner):
        """A batch field: the handle belongs to ``owner`` (a ProjectionBatch)."""
        self = cls.__new__(cls)
        self.dtype = _np_dtype(dtype)
        self._sfx = "f32" if self.dtype == np.float32 else "f64"
        self.shape = tuple(int(v) for v in shape)
        self.N = self.shape[0] * self.shape[1] * self.shape[2]
        self.F = self.shape[0] * self.shape[1] * (self.shape[2] // 2 + 1)
        self._owner = owner
        self._h = handle
        return self

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            if getattr(self, "_owner", None) is None:
                call("ffcz_destroy", self._h)
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

    def _scalar(self, v):
        return float(self.dtype.type(v))

    def _field(self, a, name):
        p, keep = as_pointer(a, self.dtype)
        n = int(np.prod(keep.shape)) if hasattr(keep, "shape") else None
        if n is not None and n != self.N:
            raise ValueError(f"{name} has {n} elements, expected {self.N}")
        return p, keep

    # -- initialize* (projection_algorithm_impl.cuh:598-738) ----------------
    def initialize(self, orig_data, freq_indices, delta_k, spatial_epsilon, num_freq_constraints=None):
        x, kx = self._field(orig_data, "orig_data")
        idx, ki = as_pointer(np.asarray(freq_indices, dtype=np.uint64) if not _is_tensor(freq_indices) else freq_indices)
        dk, kd = as_pointer(delta_k, self.dtype)
        K = int(num_freq_constraints if num_freq_constraints is not None else len(ki))
        call(f"ffcz_init_direct_{self._sfx}", self._h, x, idx, dk, self._scalar(spatial_epsilon), K)

    def initialize_all_abs(self, orig_data, spatial_epsilon, delta):
        x, kx = self._field(orig_data, "orig_data")
        call(f"ffcz_init_all_abs_{self._sfx}", self._h, x, self._scalar(spatial_epsilon), self._scalar(delta))

    def initialize_all_ptw(self, orig_data, spatial_epsilon, delta_ptw):
        x, kx = self._field(orig_data, "orig_data")
        call(f"ffcz_init_all_ptw_{self._sfx}", self._h, x, self._scalar(spatial_epsilon), self._scalar(delta_ptw))

    def initialize_partial_ptw(self, orig_data, freq_indices, delta_ptw, spatial_epsilon, num_freq_constraints=None):
        x, kx = self._field(orig_data, "orig_data")
        idx, ki = as_pointer(np.asarray(freq_indices, dtype=np.uint64))
        K = int(num_freq_constraints if num_freq_constraints is not None else len(ki))
        call(f"ffcz_init_partial_ptw_{self._sfx}", self._h, x, idx, self._scalar(delta_ptw),
             self._scalar(spatial_epsilon), K)

    # -- solve* (projection_algorithm_impl.cuh:772-902) ---------------------
    def solve(self, spatial_error, max_iterations, convergence_tolerance):
        e, ke = self._field(spatial_error, "spatial_error")
        call(f"ffcz_solve_{self._sfx}", self._h, e, int(max_iterations), self._scalar(convergence_tolerance))

    def solve_all_abs(self, spatial_error, max_iterations, convergence_tolerance):
        e, ke = self._field(spatial_error, "spatial_error")
        call(f"ffcz_solve_all_abs_{self._sfx}", self._h, e, int(max_iterations), self._scalar(convergence_tolerance))

    def solve_all_ptw(self, spatial_error, max_iterations, convergence_tolerance):
        e, ke = self._field(spatial_error, "spatial_error")
        call(f"ffcz_solve_all_ptw_{self._sfx}", self._h, e, int(max_iterations), self._scalar(convergence_tolerance))

    def solve_partial_ptw(self, spatial_error, max_iterations, convergence_tolerance):
        """The solve loop with the partial ptw projection over the bins given to
        initialize_partial_ptw (projection_algorithm_impl.cuh:303-333,
        projection_algorithm.cu:261-333; the reference stores the list but has
        no solve that launches it)."""
        e, ke = self._field(spatial_error, "spatial_error")
        call(f"ffcz_solve_partial_ptw_{self._sfx}", self._h, e, int(max_iterations),
             self._scalar(convergence_tolerance))

    def solve_loop(self, mode: int, spatial_error, max_iterations, convergence_tolerance):
        """The projection loop without the trailing extraction."""
        e, ke = self._field(spatial_error, "spatial_error")
        call("ffcz_solve_loop", self._h, int(mode), e, int(max_iterations), float(convergence_tolerance))

    def extract_compact_representation(self):
        call("ffcz_extract", self._h)

    def reset_edit_accumulators(self):
        call("ffcz_reset_edit_accumulators", self._h)

    # -- status ------------------------------------------------------------
    @property
    def iteration_count(self) -> int:
        return self.status()[0]

    @property
    def convergence_metric(self) -> float:
        return self.status()[1]

    def status(self):
        it, v, c = C.c_int(), C.c_double(), C.c_int()
        call("ffcz_get_status", self._h, C.byref(it), C.byref(v), C.byref(c))
        return it.value, v.value, bool(c.value)

    # -- getters (projection_algorithm_impl.cuh:458-595, 904-910) ----------
    def _get(self, name, dtype, out=None):
        n = C.c_size_t()
        call(name, self._h, None, 0, C.byref(n))
        if out is None:
            out = np.empty(n.value, dtype=dtype)
        p, keep = as_pointer(out)
        call(name, self._h, p, n.value, C.byref(n))
        return out

    def get_results(self, out=None):
        return self._get("ffcz_get_results", self.dtype, out)

    def get_freq_flags(self):
        return self._get("ffcz_get_freq_flags", np.uint8).astype(bool)

    def get_spatial_flags(self):
        return self._get("ffcz_get_spatial_flags", np.uint8).astype(bool)

    def get_freq_flags_pack(self):
        return self._get("ffcz_get_freq_flags_pack", np.uint8)

    def get_spatial_flags_pack(self):
        return self._get("ffcz_get_spatial_flags_pack", np.uint8)

    def get_active_freq_quant_edit_distances(self):
        return self._get("ffcz_get_active_freq_quant", np.uint16)

    def get_active_spatial_quant_edit_distances(self):
        return self._get("ffcz_get_active_spatial_quant", np.uint16)

    def get_active_freq_edit_distances(self):
        return self._get("ffcz_get_active_freq_edits", self.dtype)

    def get_active_spatial_edit_distances(self):
        return self._get("ffcz_get_active_spatial_edits", self.dtype)

    def get_freq_edits_full(self, out=None):
        return self._get("ffcz_get_freq_edits_full", self.dtype, out)

    def get_spatial_edits_full(self, out=None):
        return self._get("ffcz_get_spatial_edits_full", self.dtype, out)

    def _range(self):
        v = [C.c_double() for _ in range(4)]
        call("ffcz_get_edit_range", self._h, *[C.byref(x) for x in v])
        return [self.dtype.type(x.value) for x in v]

    def get_min_freq_edit(self):
        return self._range()[0]

    def get_max_freq_edit(self):
        return self._range()[1]

    def get_min_spatial_edit(self):
        return self._range()[2]

    def get_max_spatial_edit(self):
        return self._range()[3]

    # -- additions: certified output ----------------------------------------
    def set_margins(self, spatial: float, freq: float):
        call("ffcz_set_margins", self._h, float(spatial), float(freq))

    def set_ptw_semantics(self, corrected: bool):
        """The bound of the ptw solves: the reference's arithmetic (False,
        projection_algorithm.cu:199-258) or the corrected one (True): the power
        ratio |X+E|^2/|X|^2 held within [1-d, 1+d] on every constrained bin, DC
        included, with Re and Im clamped relative to |X|."""
        call("ffcz_set_ptw_semantics", self._h, 1 if corrected else 0)

    def set_fft_precision(self, variant):
        """FFT variant of the iteration: FFT_VALU, FFT_BF16X3, FFT_FP16X3 or
        one of their names ("valu-fp32", "mfma-bf16x3", "mfma-fp16x3")."""
        if isinstance(variant, str):
            variant = FFT_VARIANTS[variant]
        call("ffcz_set_fft_precision", self._h, int(variant))

    TILE_VAR_NAMES = ("z_fwd", "z_inv", "y", "x", "cert_z", "cert_y", "cert_x")

    def autotune(self, reps: int = 3, use_cache: bool = True, timings: bool = False) -> dict:
        """Choose each pass kernel's tile variant by timing (after initialize*,
        before solving).  Returns {pass: variant}, plus "ms" when ``timings``."""
        v = (C.c_int * 7)()
        ms = (C.c_float * 9)() if timings else None
        call("ffcz_autotune", self._h, int(reps), int(bool(use_cache)), v, ms)
        out = dict(zip(self.TILE_VAR_NAMES, list(v)))
        if timings:
            out["ms"] = [round(float(t), 5) for t in ms]
        return out

    TUNE_SOURCES = ("none", "timed", "process", "disk")

    def tune_source(self) -> str:
        """Where the last autotune's plan came from: "none" (no autotune yet),
        "timed" (tuned by that call), "process" (an earlier tuning in this
        process) or "disk" (the plan file, see set_plan_cache)."""
        v = C.c_int()
        call("ffcz_get_tune_source", self._h, C.byref(v))
        return self.TUNE_SOURCES[v.value]

    def tile_vars(self) -> dict:
        v = (C.c_int * 7)()
        call("ffcz_get_tile_vars", self._h, v)
        return dict(zip(self.TILE_VAR_NAMES, list(v)))

    def set_tile_vars(self, vars_: dict):
        v = (C.c_int * 7)(*[int(vars_.get(k, 0)) for k in self.TILE_VAR_NAMES])
        call("ffcz_set_tile_vars", self._h, v)

    def set_cert_pair(self, paired: bool):
        """certify_both's forward z transforms in one launch for both fields
        (True) or one each (False); autotune chooses."""
        call("ffcz_set_cert_pair", self._h, int(bool(paired)))

    def fft_precision(self) -> int:
        v = C.c_int()
        call("ffcz_get_fft_precision", self._h, C.byref(v))
        return v.value

    def set_quantizer(self, kind: int, freq_step: float = 0.0, spatial_step: float = 0.0):
        call("ffcz_set_quantizer", self._h, int(kind), float(freq_step), float(spatial_step))

    def certify(self, quantized: bool) -> CertResult:
        r = _lib.Cert()
        call("ffcz_certify", self._h, int(bool(quantized)), C.byref(r))
        return CertResult(bool(r.ok), r.spatial_violations, r.freq_violations, r.spatial_max_excess,
                          r.freq_max_excess)

    def certify_both(self):
        """(unquantised, quantised) certifications as one paired FP64 sequence
        (after extract_compact_representation)."""
        r = (_lib.Cert * 2)()
        call("ffcz_certify_both", self._h, r)
        return tuple(CertResult(bool(c.ok), c.spatial_violations, c.freq_violations, c.spatial_max_excess,
                                c.freq_max_excess) for c in r)

    def certify_both_enqueue(self) -> None:
        """Queue certify_both's sequence without waiting (read it with
        certify_both_results; work queued in between runs after it)."""
        call("ffcz_certify_both_enqueue", self._h)

    def certify_both_results(self):
        """Wait for the queued certify_both_enqueue and return its
        (unquantised, quantised) results."""
        r = (_lib.Cert * 2)()
        call("ffcz_certify_both_results", self._h, r)
        return tuple(CertResult(bool(c.ok), c.spatial_violations, c.freq_violations, c.spatial_max_excess,
                                c.freq_max_excess) for c in r)

    def stream_handle(self) -> int:
        s = C.c_void_p()
        call("ffcz_get_stream", self._h, C.byref(s))
        return s.value or 0

    def set_borrow_e0(self, borrow: bool = True):
        """Read a device-resident initial error in place in the certifications
        after a solve instead of copying it (the caller keeps it unchanged
        until the next solve)."""
        call("ffcz_set_borrow_e0", self._h, int(bool(borrow)))

    def set_graph(self, enable: bool = True):
        """Replay each iteration after a solve's first as one HIP graph launch
        (the default), or launch every kernel eagerly."""
        call("ffcz_set_graph", self._h, int(bool(enable)))

    def set_pass_timing(self, enable: bool = True):
        """Record HIP events between the launches of every later solve's
        iterations (resets the averages)."""
        call("ffcz_set_pass_timing", self._h, int(bool(enable)))

    def pass_timing(self):
        """(mean ms of [z fwd, y fwd, x fused, y inv, z inv, finalize] over the
        iterations that did work, their count)."""
        ms = (C.c_float * 6)()
        n = C.c_long()
        call("ffcz_get_pass_timing", self._h, ms, C.byref(n))
        return list(ms), n.value

    def pass_counts(self):
        """Timed launches behind each of pass_timing's six means."""
        n = (C.c_long * 6)()
        call("ffcz_get_pass_counts", self._h, n)
        return list(n)

    def time_iteration(self, reps: int = 10):
        ms = (C.c_float * 6)()
        call("ffcz_time_iteration", self._h, int(reps), ms)
        return list(ms)


def _is_tensor(a) -> bool:
    try:
        import torch

        return isinstance(a, torch.Tensor)
    except ImportError:  # pragma: no cover
        return False


__all__ = ["ProjectionSolver", "CertResult", "as_pointer", "MODE_ABS", "MODE_DIRECT", "MODE_PTW", "MODE_PARTIAL_PTW",
           "QUANT_REF16", "QUANT_STEP"]


def set_plan_cache(path):
    """Path of the autotuner's plan file, shared by processes (None or "":
    plans stay in this process).  The default is FFCZ_PLAN_CACHE, else
    $XDG_CACHE_HOME/ffcz/plans.txt, else ~/.cache/ffcz/plans.txt."""
    call("ffcz_set_plan_cache", None if path is None else str(path).encode())


def plan_cache() -> str:
    """The plan file's path ("" when plans stay in the process)."""
    n = C.c_size_t()
    call("ffcz_get_plan_cache", None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    call("ffcz_get_plan_cache", buf, n.value + 1, C.byref(n))
    return buf.value.decode()
