"""MaBoSS on MI355X: Python face of the engine's C-ABI.

The reference (MaBoSS.GPU) is driven from its command line
(``MaBoSS.GPU [-o prefix] model.bnd model.cfg``, src/main.cpp:101-169).  This
module exposes the same pipeline -- parse, generate + JIT, simulate, write
``_probtraj.csv`` / ``_fp.csv`` -- as objects, for tests, benchmarks and the
differentiable form.  All computation happens in the HIP kernels behind
``libmaboss_gpu.so``; there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

from ._capi import (CLI_PATH, BatchInfo, BatchParams, CompileOpts, MbgError, ModelInfo,
                    ResultInfo, SimParams, Variant, check, lib)

__all__ = ["Model", "Program", "Result", "simulate", "MbgError", "run_cli"]

# Load the engine when the package is imported (after torch, see
# _capi.lib), so the process's HIP runtime is settled before any other
# import can pick one.  A missing library is reported by the first call that
# needs it.
try:
    lib()
except RuntimeError:
    pass


def _release(fn_name, h):
    """Free a handle; quiet during interpreter shutdown."""
    try:
        getattr(lib(), fn_name)(h)
    except Exception:
        pass

WINDOW_MODES = {"auto": -1, "dense": 0, "hash": 1, "lds": 2, "sort": 3}


class Model:
    """A parsed .bnd/.cfg model (the reference's ``driver``)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load(cls, bnd_path: str, cfg_path: str) -> "Model":
        h = ctypes.c_void_p()
        check(lib().mbg_model_load(os.fsencode(bnd_path), os.fsencode(cfg_path), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_strings(cls, bnd: str, cfg: str) -> "Model":
        h = ctypes.c_void_p()
        check(lib().mbg_model_load_strings(bnd.encode(), cfg.encode(), ctypes.byref(h)))
        return cls(h)

    def __del__(self, _release=_release):
        h = getattr(self, "_h", None)
        if h:
            _release("mbg_model_free", h)
            self._h = None

    @property
    def info(self) -> ModelInfo:
        out = ModelInfo()
        check(lib().mbg_model_get_info(self._h, ctypes.byref(out)))
        return out

    @property
    def node_names(self) -> list[str]:
        return [lib().mbg_model_node_name(self._h, i).decode() for i in range(self.info.n_nodes)]

    @property
    def internal(self) -> list[bool]:
        return [lib().mbg_model_node_internal(self._h, i) == 1 for i in range(self.info.n_nodes)]

    @property
    def observed(self) -> list[int]:
        return [i for i, x in enumerate(self.internal) if not x]

    @property
    def param_names(self) -> list[str]:
        return [lib().mbg_model_param_name(self._h, k).decode() for k in range(self.info.n_params)]

    @property
    def param_values(self) -> np.ndarray:
        return np.array([lib().mbg_model_param_value(self._h, k) for k in range(self.info.n_params)],
                        dtype=np.float32)

    def set_constant(self, name: str, value: float):
        check(lib().mbg_model_set_constant(self._h, name.encode(), float(value)))

    def set_variable(self, name: str, value: float):
        check(lib().mbg_model_set_variable(self._h, name.encode(), float(value)))

    def rate(self, node: int, state_words) -> float:
        """Rate of ``node`` in a state, by the host expression interpreter."""
        arr = (ctypes.c_uint32 * len(state_words))(*state_words)
        out = ctypes.c_float()
        check(lib().mbg_model_eval_rate(self._h, node, arr, ctypes.byref(out)))
        return out.value

    def generate_code(self, entropy_formula: int = 0, with_gradient: bool = False) -> str:
        opts = CompileOpts(entropy_formula=entropy_formula, with_gradient=int(with_gradient))
        p = ctypes.c_void_p()
        check(lib().mbg_generate_code(self._h, ctypes.byref(opts), ctypes.byref(p)))
        try:
            return ctypes.string_at(p).decode()
        finally:
            lib().mbg_free_string(p)

    def compile(self, entropy_formula: int = 0, no_cache: bool = False, lane_occupancy: int = 0,
                grad_occupancy: int = 0) -> "Program":
        return Program(self, entropy_formula=entropy_formula, no_cache=no_cache, lane_occupancy=lane_occupancy,
                       grad_occupancy=grad_occupancy)


class Program:
    """Generated + JIT-compiled kernels of a model (the reference's kernel_compiler)."""

    def __init__(self, model: Model, entropy_formula: int = 0, no_cache: bool = False, lane_occupancy: int = 0,
                 grad_occupancy: int = 0):
        opts = CompileOpts(entropy_formula=entropy_formula, no_cache=int(no_cache), lane_occupancy=lane_occupancy,
                           grad_occupancy=grad_occupancy)
        h = ctypes.c_void_p()
        check(lib().mbg_program_compile(model._h, ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self.model = model

    def __del__(self, _release=_release):
        h = getattr(self, "_h", None)
        if h:
            _release("mbg_program_free", h)
            self._h = None

    @property
    def compile_seconds(self) -> float:
        return lib().mbg_program_compile_seconds(self._h)

    def simulate(self, n_trajectories: int = -1, first_id: int = 0, seed: int = -1, variant: int = 0,
                 max_steps: int = 0, window_mode: str = "auto", final_mode: str = "auto",
                 window_capacity: int = 0, final_capacity: int = 0, fixed_capacity: int = 0,
                 block_size: int = 0, blocks_per_cu: int = 0, normalization: float = 0.0,
                 device_only: bool = False, kernel: str = "auto", rates=None, stream=None,
                 wait: bool = True, errors: bool = False, moment_states: int = 0,
                 refill_lanes: int = 0) -> "Result":
        """Simulate; ``rates`` (one float per ``Model.param_names``) replaces the model's
        parameter values at run time (lane kernel).  With ``wait=False`` the call returns
        once the run is enqueued (``mbg_simulate_async``): the statistics reach the host on
        the program's copy stream while the next run's kernels go ahead, and the result
        completes in :meth:`Result.wait` (which every accessor calls).  ``errors`` fills the
        ErrorTH / ErrorProba columns from exact per-trajectory second moments
        (``MBG_SIM_ERRORS``: a replay of the trajectories after the run); ``moment_states``
        (8 or 1) is the replay's per-lane state buffer, which changes no result."""
        p, keep = self._params(n_trajectories, first_id, seed, variant, max_steps, window_mode, final_mode,
                               window_capacity, final_capacity, fixed_capacity, block_size, blocks_per_cu,
                               normalization, device_only, kernel, rates, errors, moment_states)
        p.refill_lanes = refill_lanes
        h = ctypes.c_void_p()
        s = ctypes.c_void_p(stream) if stream else None
        entry = lib().mbg_simulate if wait else lib().mbg_simulate_async
        check(entry(self._h, ctypes.byref(p), s, ctypes.byref(h)))
        return Result(h, self.model)

    def _params(self, n_trajectories, first_id, seed, variant, max_steps, window_mode, final_mode, window_capacity,
                final_capacity, fixed_capacity, block_size, blocks_per_cu, normalization, device_only, kernel, rates,
                errors, moment_states):
        p = SimParams()
        lib().mbg_sim_params_default(ctypes.byref(p))
        p.n_trajectories = n_trajectories
        p.first_id = first_id
        p.seed = seed
        p.variant = variant
        p.max_steps = max_steps
        p.window_mode = WINDOW_MODES[window_mode]
        p.final_mode = WINDOW_MODES[final_mode]
        p.window_capacity = window_capacity
        p.final_capacity = final_capacity
        p.fixed_capacity = fixed_capacity
        p.block_size = block_size
        p.blocks_per_cu = blocks_per_cu
        p.normalization = normalization
        p.flags = (1 if device_only else 0) | (2 if errors else 0)
        p.moment_states = moment_states
        p.kernel = {"auto": -1, "lane": 0, "wave": 1}[kernel]
        keep = None
        if rates is not None:
            keep = np.ascontiguousarray(rates, dtype=np.float32)
            if keep.shape != (self.model.info.n_params,):
                raise ValueError(f"rates must have {self.model.info.n_params} entries")
            p.rates = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        return p, keep

    def simulate_checkpointed(self, path: str, chunk: int = 0, n_trajectories: int = -1, first_id: int = 0,
                              seed: int = -1, variant: int = 0, max_steps: int = 0, kernel: str = "auto", rates=None,
                              errors: bool = False, stream=None) -> "Result":
        """Simulate in chunks of ``chunk`` trajectory ids, saving the merged statistics to
        ``path`` after each (``mbg_simulate_checkpointed``); started again with the same
        ``path`` it resumes after the last saved id.  The result equals one
        :meth:`simulate` of the whole range bit for bit."""
        p, keep = self._params(n_trajectories, first_id, seed, variant, max_steps, "auto", "auto", 0, 0, 0, 0, 0,
                               0.0, False, kernel, rates, errors, 0)
        h = ctypes.c_void_p()
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().mbg_simulate_checkpointed(self._h, ctypes.byref(p), chunk, os.fsencode(path), s, ctypes.byref(h)))
        del keep
        return Result(h, self.model)

    def load_result(self, path: str) -> "Result":
        """A result saved with :meth:`Result.save` (``mbg_result_load``)."""
        h = ctypes.c_void_p()
        check(lib().mbg_result_load(self._h, os.fsencode(path), ctypes.byref(h)))
        return Result(h, self.model)

    def simulate_batched(self, variants, seed: int = -1, max_steps: int = 0, block_size: int = 0,
                         blocks_per_cu: int = 0, traj_per_lane: int = 0, use_graph: bool = True, lds_cells: int = 0,
                         device_only: bool = False, stream=None):
        """Run several variants of the model in one launch.

        ``variants`` is a list of dicts with ``n_trajectories`` and optional
        ``max_time``, ``rates`` (one per ``Model.param_names``) and
        ``mutations`` (one int per node: +1 on, -1 off, 0 free).  Returns
        ``(results, info)``: one :class:`Result` per variant and the launch's
        :class:`BatchInfo`.
        """
        n_nodes, n_params = self.model.info.n_nodes, self.model.info.n_params
        arr = (Variant * len(variants))()
        keep = []
        for v, spec in zip(arr, variants):
            v.n_trajectories = int(spec["n_trajectories"])
            v.max_time = float(spec.get("max_time", 0.0))
            if spec.get("rates") is not None:
                r = np.ascontiguousarray(spec["rates"], dtype=np.float32)
                if r.shape != (n_params,):
                    raise ValueError(f"rates must have {n_params} entries")
                keep.append(r)
                v.rates = r.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            if spec.get("mutations") is not None:
                mu = np.ascontiguousarray(spec["mutations"], dtype=np.int8)
                if mu.shape != (n_nodes,):
                    raise ValueError(f"mutations must have {n_nodes} entries")
                keep.append(mu)
                v.mutations = mu.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))
        bp = BatchParams()
        lib().mbg_batch_params_default(ctypes.byref(bp))
        bp.seed = seed
        bp.max_steps = max_steps
        bp.block_size = block_size
        bp.blocks_per_cu = blocks_per_cu
        bp.traj_per_lane = traj_per_lane
        bp.lds_cells = lds_cells
        bp.use_graph = int(use_graph)
        bp.flags = 1 if device_only else 0
        outs = (ctypes.c_void_p * len(variants))()
        info = BatchInfo()
        s = ctypes.c_void_p(stream) if stream else None
        check(lib().mbg_simulate_batched(self._h, ctypes.byref(bp), len(variants), arr, s,
                                         ctypes.cast(outs, ctypes.POINTER(ctypes.c_void_p)),
                                         ctypes.byref(info)))
        del keep
        return [Result(ctypes.c_void_p(h), self.model) for h in outs], info


@dataclass
class Window:
    states: np.ndarray  # uint64 observed-state indices
    raw: np.ndarray     # uint64 fixed-point sums
    probs: np.ndarray   # float64


class Result:
    """Statistics of one run: window probabilities, entropies, final states, fixed points."""

    def __init__(self, handle, model: Model):
        self._h = handle
        self.model = model

    def __del__(self, _release=_release):
        h = getattr(self, "_h", None)
        if h:
            _release("mbg_result_free", h)
            self._h = None

    @property
    def info(self) -> ResultInfo:
        out = ResultInfo()
        check(lib().mbg_result_get_info(self._h, ctypes.byref(out)))
        return out

    def wait(self) -> "Result":
        """Block until the statistics are on the host (runs started with ``wait=False``)."""
        check(lib().mbg_result_wait(self._h))
        return self

    def window(self, w: int) -> Window:
        n = lib().mbg_result_window_size(self._h, w)
        if n < 0:
            raise IndexError(w)
        states = np.zeros(n, dtype=np.uint64)
        raw = np.zeros(n, dtype=np.uint64)
        probs = np.zeros(n, dtype=np.float64)
        check(lib().mbg_result_window(self._h, w, states.ctypes.data, raw.ctypes.data))
        check(lib().mbg_result_window_probs(self._h, w, None, probs.ctypes.data))
        return Window(states, raw, probs)

    def windows(self) -> list[Window]:
        return [self.window(w) for w in range(self.info.windows)]

    def entropies(self) -> tuple[np.ndarray, np.ndarray]:
        nw = self.info.windows
        th = np.zeros(nw)
        h = np.zeros(nw)
        check(lib().mbg_result_entropies(self._h, th.ctypes.data, h.ctypes.data))
        return th, h

    def error_th(self) -> np.ndarray:
        """ErrorTH per window (zeros unless the run had ``errors=True``)."""
        out = np.zeros(self.info.windows)
        check(lib().mbg_result_error_th(self._h, out.ctypes.data))
        return out

    def window_errors(self, w: int) -> np.ndarray:
        """ErrorProba of window w's cells, in :meth:`window`'s order."""
        n = lib().mbg_result_window_size(self._h, w)
        if n < 0:
            raise IndexError(w)
        out = np.zeros(n)
        check(lib().mbg_result_window_errors(self._h, w, out.ctypes.data))
        return out

    def final_states(self) -> tuple[np.ndarray, np.ndarray]:
        n = lib().mbg_result_final_size(self._h)
        s = np.zeros(n, dtype=np.uint64)
        c = np.zeros(n, dtype=np.uint64)
        check(lib().mbg_result_final(self._h, s.ctypes.data, c.ctypes.data))
        return s, c

    def fixed_points(self) -> tuple[np.ndarray, np.ndarray]:
        n = lib().mbg_result_fixed_size(self._h)
        nw = self.info.n_words
        s = np.zeros((n, nw), dtype=np.uint32)
        c = np.zeros(n, dtype=np.uint64)
        check(lib().mbg_result_fixed(self._h, s.ctypes.data, c.ctypes.data))
        return s, c

    def write_csv(self, prefix: str, with_finals: bool = False):
        check(lib().mbg_result_write_csv(self._h, os.fsencode(prefix), int(with_finals)))

    def print(self, path: str | None = None):
        check(lib().mbg_result_print(self._h, os.fsencode(path) if path else None))

    def write_raw(self, path: str):
        check(lib().mbg_result_write_raw(self._h, os.fsencode(path)))

    def save(self, path: str):
        """The exact statistics and run identity to a checkpoint file (``mbg_result_save``)."""
        check(lib().mbg_result_save(self._h, os.fsencode(path)))

    def merge(self, other: "Result") -> "Result":
        """Add ``other`` (same run, adjacent trajectory ids) into this result."""
        check(lib().mbg_result_merge(self._h, other._h))
        return self


def simulate(bnd_path: str, cfg_path: str, **kw) -> Result:
    """Parse, compile and simulate in one call."""
    m = Model.load(bnd_path, cfg_path)
    return m.compile().simulate(**kw)


def run_cli(*args: str, timeout: float = 600) -> subprocess.CompletedProcess:
    """Run the MaBoSS.GPU command line built next to the library."""
    if not os.path.exists(CLI_PATH):
        raise RuntimeError(f"{CLI_PATH} is missing: build first")
    return subprocess.run([CLI_PATH, *args], capture_output=True, text=True, timeout=timeout)
