"""flock_hip: Python binding of libflock_hip.so, the gfx950 flock step.

Mirrors the reference's interface for the step (crates/flock-ref/src/state.rs
``Flock``/``.flk`` exchange, and the pass sequence of cuda/flock.cu:452-470):
state travels as ``.flk`` records (``WORDS`` = 15 float32 words per bird) in a
``numpy`` array of shape ``(n, 15)``; the device work all happens inside the
library, through its C ABI (include/flock_hip.h). There is no CPU fallback:
loading fails loudly when the library is not built, and device calls raise
``FlockHipError`` when no gfx950 GPU is present.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

WORDS = 15
# .flk record word indices (crates/flock-ref/src/state.rs:4-7)
PX, PY, PZ, VX, VY, VZ, ROLL, SPIN, RADIUS, SX, SY, SZ, NEXT, NB, CEN = range(WORDS)

FLAG_CANONICAL = 1
FLAG_UNFUSED = 2
FLAG_GRAPH = 4
FLAG_WAVE_AGG = 8
FLAG_NO_TILE = 16
FLAG_OCC = 32
FLAG_RESIDENT = 64
FLAG_SPATIAL = 128
FLAG_PERSISTENT = 256
FLAG_XCD = 512
FLAG_WINDOW = 1024
FLAG_FLIGHT = 2048
FLAG_LAZY = 4096
LAZY_PERIOD = 4  # steps per rebuild with FLAG_LAZY

OK = 0
ERR_INVALID, ERR_FLK, ERR_NOMEM, ERR_HIP, ERR_BUDGET, ERR_STATE, ERR_IO, ERR_NODEVICE = range(1, 9)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "_build", "libflock_hip.so")


class FlockHipError(RuntimeError):
    """A failed library call; ``status`` is the fh_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("cell", C.c_float),
        ("buckets", C.c_uint32),
        ("decide_block", C.c_uint32),
        ("birds_per_thread", C.c_uint32),
        ("flags", C.c_uint32),
        ("decide_group", C.c_uint32),
        ("bucket_shift", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class PassMs(C.Structure):
    _fields_ = [
        ("grid_histogram", C.c_double),
        ("grid_scan", C.c_double),
        ("grid_scatter", C.c_double),
        ("decide", C.c_double),
        ("integrate", C.c_double),
        ("step_total", C.c_double),
        ("wall_ms", C.c_double),
        ("timed_steps", C.c_uint32),
        ("due_last", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_P = C.c_void_p
_u32 = C.c_uint32
_SIGS = {
    "fh_version": ([], C.c_char_p),
    "fh_last_error": ([], C.c_char_p),
    "fh_constants": ([_P, C.c_int], C.c_int),
    "fh_buckets_for": ([_u32], _u32),
    "fh_params_default": ([_P], C.c_int32),
    "fh_device_info": ([C.c_char_p, C.c_int, _P, _P], C.c_int32),
    "fh_kernel_info": ([C.c_char_p, _P, _P, _P], C.c_int32),
    "fh_read_flk": ([C.c_char_p, _P, _P, _P, _u32], C.c_int32),
    "fh_write_flk": ([C.c_char_p, _u32, C.c_float, _P], C.c_int32),
    "fh_gen_reference": ([_u32, C.c_uint64, _P], C.c_int32),
    "fh_gen_cube": ([_u32, C.c_float, C.c_uint64, _P], C.c_int32),
    "fh_gen_clumps": ([_u32, _u32, C.c_float, C.c_float, C.c_uint64, _P], C.c_int32),
    "fh_gen_request": ([_u32, C.c_uint64, C.c_float, C.c_float, _P], C.c_int32),
    "fh_create": ([_u32, _P, _P], C.c_int32),
    "fh_destroy": ([_P], C.c_int32),
    "fh_get_params": ([_P, _P, _P], C.c_int32),
    "fh_upload_words": ([_P, _P, _u32, C.c_float], C.c_int32),
    "fh_download_words": ([_P, _P, _P], C.c_int32),
    "fh_hash_count": ([_P], C.c_int32),
    "fh_scan": ([_P], C.c_int32),
    "fh_scatter": ([_P], C.c_int32),
    "fh_decide": ([_P, C.c_float], C.c_int32),
    "fh_integrate": ([_P], C.c_int32),
    "fh_integrate_hash": ([_P], C.c_int32),
    "fh_step": ([_P, _u32, _u32, _P], C.c_int32),
    "fh_sync": ([_P], C.c_int32),
    "fh_get_time": ([_P, _P], C.c_int32),
    "fh_read_grid": ([_P, _P, _P, _P, _P], C.c_int32),
    "fh_neighbours": ([_P, _u32, C.c_int, _P, _u32, _P], C.c_int32),
    "fh_batch_create": ([_u32, _u32, _u32, _P, _P], C.c_int32),
    "fh_batch_destroy": ([_P], C.c_int32),
    "fh_batch_add": ([_P, _P, _u32, C.c_float, _P], C.c_int32),
    "fh_batch_remove": ([_P, _u32], C.c_int32),
    "fh_batch_limit": ([_P, _u32, _u32], C.c_int32),
    "fh_batch_download": ([_P, _u32, _P, _P], C.c_int32),
    "fh_batch_step": ([_P, _u32, _u32, _P], C.c_int32),
    "fh_batch_info": ([_P, _P, _P, _P, _P], C.c_int32),
    "fh_flock_bytes": ([_u32, _P], C.c_size_t),
    "fh_step_device_noalloc": ([_P, _u32, C.c_float, _u32, _P, _P, _P], C.c_int32),
}


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlockHipError(ERR_NODEVICE, f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != OK:
        raise FlockHipError(status, lib().fh_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def version() -> str:
    return lib().fh_version().decode()


def constants() -> np.ndarray:
    n = lib().fh_constants(None, 0)
    out = np.zeros(n, np.float32)
    lib().fh_constants(_ptr(out), n)
    return out


def buckets_for(n: int) -> int:
    return int(lib().fh_buckets_for(n))


def default_params() -> Params:
    p = Params()
    _check(lib().fh_params_default(C.addressof(p)))
    return p


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    hbm = C.c_size_t(0)
    _check(lib().fh_device_info(name, 256, C.addressof(cus), C.addressof(hbm)))
    return {"name": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value}


def kernel_info(kernel: str) -> dict:
    v, l, s = C.c_int(0), C.c_int(0), C.c_int(0)
    _check(lib().fh_kernel_info(kernel.encode(), C.addressof(v), C.addressof(l), C.addressof(s)))
    return {"vgprs": v.value, "lds_bytes": l.value, "scratch_bytes": s.value}


# ---- .flk files and generators (host only) ---------------------------------

def read_flk(path: str) -> tuple[np.ndarray, float]:
    """Flock::read (state.rs:135-170): returns (words[n, 15] float32, time)."""
    n, t = C.c_uint32(0), C.c_float(0)
    _check(lib().fh_read_flk(str(path).encode(), C.addressof(n), C.addressof(t), None, 0))
    w = np.zeros((n.value, WORDS), np.float32)
    _check(lib().fh_read_flk(str(path).encode(), C.addressof(n), C.addressof(t), _ptr(w), n.value))
    return w, float(t.value)


def write_flk(path: str, words: np.ndarray, time: float) -> None:
    """Flock::write (state.rs:104-132)."""
    w = np.ascontiguousarray(words, np.float32).reshape(-1, WORDS)
    _check(lib().fh_write_flk(str(path).encode(), w.shape[0], np.float32(time), _ptr(w)))


def gen_reference(n: int, seed: int = 1) -> np.ndarray:
    """Flock::scatter (state.rs:79-101): the reference benchmark start."""
    w = np.zeros((n, WORDS), np.float32)
    _check(lib().fh_gen_reference(n, seed, _ptr(w)))
    return w


def gen_cube(n: int, side: float, seed: int = 1) -> np.ndarray:
    """Uniform cube about the roost, isotropic headings, 10 +- 1 m/s."""
    w = np.zeros((n, WORDS), np.float32)
    _check(lib().fh_gen_cube(n, np.float32(side), seed, _ptr(w)))
    return w


def gen_clumps(n: int, clumps: int = 8, sigma: float = 5.0, cube: float = 1000.0, seed: int = 1) -> np.ndarray:
    """Gaussian clumps with aligned headings (10 degree cap) at 10 m/s."""
    w = np.zeros((n, WORDS), np.float32)
    _check(lib().fh_gen_clumps(n, clumps, np.float32(sigma), np.float32(cube), seed, _ptr(w)))
    return w


def gen_request(n: int, seed: int, birds_per_cell: float = 0.64, cell: float = 4.0) -> np.ndarray:
    """Start state of a served request (a cube at `birds_per_cell` birds per cell)."""
    w = np.zeros((n, WORDS), np.float32)
    _check(lib().fh_gen_request(n, seed, np.float32(birds_per_cell), np.float32(cell), _ptr(w)))
    return w


def cube_side_for(n: int, birds_per_cell: float = 0.64, cell: float = 4.0) -> float:
    """Cube edge giving `birds_per_cell` mean birds per cell: 100 m at 10k birds."""
    return float(cell * (n / birds_per_cell) ** (1.0 / 3.0))


# ---- one flock ------------------------------------------------------------

@dataclass
class Grid:
    keys: np.ndarray
    counts: np.ndarray
    offsets: np.ndarray | None
    sorted: np.ndarray | None


class Flock:
    """One flock on the GPU with the reference pass API (cuda/flock.cu:452-470)."""

    def __init__(self, n: int, *, cell: float = 0.0, buckets: int = 0, decide_block: int = 0,
                 birds_per_thread: int = 0, flags: int = 0, decide_group: int = 0, bucket_shift: int = 0):
        p = Params()
        p.cell, p.buckets, p.decide_block = cell, buckets, decide_block
        p.birds_per_thread, p.flags, p.decide_group = birds_per_thread, flags, decide_group
        p.bucket_shift = bucket_shift
        h = C.c_void_p()
        _check(lib().fh_create(n, C.addressof(p), C.addressof(h)))
        self._h = h
        self.n = n
        out = Params()
        b = C.c_uint32(0)
        _check(lib().fh_get_params(self._h, C.addressof(out), C.addressof(b)))
        self.params = out
        self.buckets = b.value

    @classmethod
    def from_words(cls, words: np.ndarray, time: float = 0.0, **kw) -> "Flock":
        w = np.ascontiguousarray(words, np.float32).reshape(-1, WORDS)
        f = cls(w.shape[0], **kw)
        f.upload(w, time)
        return f

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, words: np.ndarray, time: float = 0.0) -> None:
        w = np.ascontiguousarray(words, np.float32).reshape(-1, WORDS)
        _check(lib().fh_upload_words(self._h, _ptr(w), w.shape[0], np.float32(time)))

    def download(self) -> tuple[np.ndarray, float]:
        w = np.zeros((self.n, WORDS), np.float32)
        t = C.c_float(0)
        _check(lib().fh_download_words(self._h, _ptr(w), C.addressof(t)))
        return w, float(t.value)

    @property
    def time(self) -> float:
        t = C.c_float(0)
        _check(lib().fh_get_time(self._h, C.addressof(t)))
        return float(t.value)

    def hash_count(self) -> None:
        _check(lib().fh_hash_count(self._h))

    def scan(self) -> None:
        _check(lib().fh_scan(self._h))

    def scatter(self) -> None:
        _check(lib().fh_scatter(self._h))

    def decide(self, time: float) -> None:
        _check(lib().fh_decide(self._h, np.float32(time)))

    def integrate(self) -> None:
        _check(lib().fh_integrate(self._h))

    def integrate_hash(self) -> None:
        _check(lib().fh_integrate_hash(self._h))

    def sync(self) -> None:
        _check(lib().fh_sync(self._h))

    def step(self, steps: int = 1, warmup: int = 0, timing: bool = False) -> dict | None:
        out = PassMs() if timing else None
        _check(lib().fh_step(self._h, steps, warmup, C.addressof(out) if timing else None))
        return out.as_dict() if timing else None

    def read_grid(self, offsets: bool = True, sorted: bool = True) -> Grid:
        keys = np.zeros(self.n, np.uint32)
        counts = np.zeros(self.buckets, np.uint32)
        off = np.zeros(self.buckets + 1, np.uint32) if offsets else None
        srt = np.zeros(self.n, np.uint32) if sorted else None
        _check(lib().fh_read_grid(self._h, _ptr(keys), _ptr(counts), _ptr(off) if offsets else None,
                                  _ptr(srt) if sorted else None))
        return Grid(keys, counts, off, srt)

    def neighbours(self, bird: int, walk: int = 1, cap: int = 4096) -> np.ndarray:
        ids = np.zeros(cap, np.uint32)
        cnt = C.c_uint32(0)
        _check(lib().fh_neighbours(self._h, bird, walk, _ptr(ids), cap, C.addressof(cnt)))
        return ids[: min(cnt.value, cap)].copy() if cnt.value <= cap else ids.copy()


NOALLOC_MAX_BIRDS = 65536


def step_device_noalloc(d_words: int, n: int, time: float, steps: int, stream: int = 0, *, cell: float = 0.0,
                        buckets: int = 0, decide_block: int = 0, birds_per_thread: int = 0, flags: int = 0,
                        decide_group: int = 0, bucket_shift: int = 0) -> float:
    """fh_step_device_noalloc: steps the n .flk records at device address
    `d_words` in place (no allocation; launches on `stream`, not waited for);
    returns the new clock."""
    p = Params()
    p.cell, p.buckets, p.decide_block = cell, buckets, decide_block
    p.birds_per_thread, p.flags, p.decide_group = birds_per_thread, flags, decide_group
    p.bucket_shift = bucket_shift
    t = C.c_float(0)
    _check(lib().fh_step_device_noalloc(d_words, n, np.float32(time), steps, C.addressof(p), stream or None,
                                        C.addressof(t)))
    return float(t.value)


def flock_bytes(n: int, buckets: int = 0, bucket_shift: int = 0, flags: int = 0) -> int:
    """Device bytes one flock of n birds occupies in a batch (admission unit);
    with FLAG_RESIDENT it includes the resident layout's arrays."""
    p = Params()
    p.buckets, p.bucket_shift, p.flags = buckets, bucket_shift, flags
    return int(lib().fh_flock_bytes(n, C.addressof(p)))


class Batch:
    """Many independent flocks stepped together, one launch per stage."""

    def __init__(self, bird_capacity: int, bucket_capacity: int, flock_capacity: int, *, cell: float = 0.0,
                 buckets: int = 0, decide_block: int = 0, birds_per_thread: int = 0, flags: int = 0,
                 decide_group: int = 0, bucket_shift: int = 0):
        p = Params()
        p.cell, p.buckets, p.decide_block = cell, buckets, decide_block
        p.birds_per_thread, p.flags, p.decide_group = birds_per_thread, flags, decide_group
        p.bucket_shift = bucket_shift
        h = C.c_void_p()
        _check(lib().fh_batch_create(bird_capacity, bucket_capacity, flock_capacity, C.addressof(p),
                                     C.addressof(h)))
        self._h = h
        self.sizes: dict[int, int] = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, words: np.ndarray, time: float = 0.0) -> int:
        w = np.ascontiguousarray(words, np.float32).reshape(-1, WORDS)
        fid = C.c_uint32(0)
        _check(lib().fh_batch_add(self._h, _ptr(w), w.shape[0], np.float32(time), C.addressof(fid)))
        self.sizes[fid.value] = w.shape[0]
        return fid.value

    def remove(self, flock_id: int) -> None:
        _check(lib().fh_batch_remove(self._h, flock_id))
        self.sizes.pop(flock_id, None)

    def limit(self, flock_id: int, steps: int = 0xFFFFFFFF) -> None:
        """Resident batches: at most `steps` more steps for this flock, then frozen."""
        _check(lib().fh_batch_limit(self._h, flock_id, steps))

    def download(self, flock_id: int) -> tuple[np.ndarray, float]:
        w = np.zeros((self.sizes[flock_id], WORDS), np.float32)
        t = C.c_float(0)
        _check(lib().fh_batch_download(self._h, flock_id, _ptr(w), C.addressof(t)))
        return w, float(t.value)

    def step(self, steps: int = 1, warmup: int = 0, timing: bool = False) -> dict | None:
        out = PassMs() if timing else None
        _check(lib().fh_batch_step(self._h, steps, warmup, C.addressof(out) if timing else None))
        return out.as_dict() if timing else None

    def info(self) -> dict:
        lf, lb, br, db = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0), C.c_size_t(0)
        _check(lib().fh_batch_info(self._h, C.addressof(lf), C.addressof(lb), C.addressof(br), C.addressof(db)))
        return {"live_flocks": lf.value, "live_birds": lb.value, "bucket_range": br.value,
                "device_bytes": db.value}


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_double),
        ("max_active", C.c_uint32),
        ("bird_capacity", C.c_uint32),
        ("bucket_capacity", C.c_uint32),
        ("steps_per_iter", C.c_uint32),
        ("prep_threads", C.c_uint32),
        ("keep_states", C.c_uint32),
        ("birds_per_cell", C.c_float),
        ("cell_for_density", C.c_float),
        ("params", Params),
    ]


class Result(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("seed", C.c_uint64),
        ("birds", C.c_uint32),
        ("steps", C.c_uint32),
        ("status", C.c_int32),
        ("final_time", C.c_float),
        ("arrival_s", C.c_double),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
        ("checksum", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("iterations", C.c_uint64),
        ("steps", C.c_uint64),
        ("bird_steps", C.c_double),
        ("busy_s", C.c_double),
        ("wall_s", C.c_double),
        ("peak_bytes", C.c_double),
        ("peak_active", C.c_uint32),
        ("queued", C.c_uint32),
        ("active", C.c_uint32),
        ("admit_s", C.c_double),
        ("retire_s", C.c_double),
        ("gap_s", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS.update({
    "fh_server_config_default": ([_P], C.c_int32),
    "fh_server_create": ([_P, _P], C.c_int32),
    "fh_server_destroy": ([_P], C.c_int32),
    "fh_server_submit": ([_P, _u32, _u32, C.c_uint64, _P], C.c_int32),
    "fh_server_poll": ([_P, _P, _u32, _P], C.c_int32),
    "fh_server_drain": ([_P, C.c_double], C.c_int32),
    "fh_server_stats_get": ([_P, _P], C.c_int32),
    "fh_server_take_state": ([_P, C.c_uint64, _P, _u32, _P], C.c_int32),
    "fh_server_replay": ([_P, _P, _P, _P, _P, _u32, _P, _P], C.c_int32),
})


def server_config(**kw) -> ServerConfig:
    c = ServerConfig()
    _check(lib().fh_server_config_default(C.addressof(c)))
    for k, v in kw.items():
        if k in ("cell", "buckets", "decide_block", "birds_per_thread", "flags", "decide_group", "bucket_shift"):
            setattr(c.params, k, v)
        else:
            setattr(c, k, v)
    return c


class Server:
    """Continuous-batching request loop (a worker thread inside the library)."""

    def __init__(self, **kw):
        self.cfg = server_config(**kw)
        h = C.c_void_p()
        _check(lib().fh_server_create(C.addressof(self.cfg), C.addressof(h)))
        self._h = h
        self.sizes: dict[int, int] = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, birds: int, steps: int, seed: int) -> int:
        rid = C.c_uint64(0)
        _check(lib().fh_server_submit(self._h, birds, steps, seed, C.addressof(rid)))
        self.sizes[rid.value] = birds
        return rid.value

    def poll(self, cap: int = 1024) -> list[dict]:
        buf = (Result * cap)()
        n = C.c_uint32(0)
        _check(lib().fh_server_poll(self._h, C.addressof(buf), cap, C.addressof(n)))
        return [buf[i].as_dict() for i in range(n.value)]

    def drain(self, timeout_s: float = 600.0) -> None:
        _check(lib().fh_server_drain(self._h, timeout_s))

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().fh_server_stats_get(self._h, C.addressof(s)))
        return s.as_dict()

    def take_state(self, rid: int) -> tuple[np.ndarray, float]:
        w = np.zeros((self.sizes[rid], WORDS), np.float32)
        t = C.c_float(0)
        _check(lib().fh_server_take_state(self._h, rid, _ptr(w), w.shape[0], C.addressof(t)))
        return w, float(t.value)


def server_replay(birds, steps, seeds, arrival_s, **kw) -> tuple[list[dict], dict]:
    """Open-loop replay of a request stream; returns (results, stats)."""
    cfg = server_config(**kw)
    b = np.ascontiguousarray(birds, np.uint32)
    s = np.ascontiguousarray(steps, np.uint32)
    sd = np.ascontiguousarray(seeds, np.uint64)
    at = np.ascontiguousarray(arrival_s, np.float64)
    n = len(b)
    res = (Result * max(1, n))()
    st = ServerStats()
    _check(lib().fh_server_replay(C.addressof(cfg), _ptr(b), _ptr(s), _ptr(sd), _ptr(at), n, C.addressof(res),
                                  C.addressof(st)))
    return [res[i].as_dict() for i in range(n)], st.as_dict()


def _spread10(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def spatial_keys(words: np.ndarray, buckets: int, cell: float = 4.0) -> np.ndarray:
    """Bucket of each bird under FLAG_SPATIAL (the library's spatial hash:
    Morton order of the low 10 bits of each cell coordinate plus the
    reference hash of the 1024-cell block), for checking the grid."""
    w = np.ascontiguousarray(words, np.float32).reshape(-1, WORDS)
    inv = np.float32(1.0) / np.float32(cell)
    c = []
    for a in range(3):
        f = np.floor(w[:, a] * inv)
        f = np.where(np.isnan(f), 0.0, np.clip(f, -2147483648.0, 2147483647.0))
        c.append(f.astype(np.int64).astype(np.uint32).astype(np.uint64))
    m = _spread10(c[0]) | (_spread10(c[1]) << 1) | (_spread10(c[2]) << 2)
    hx, hy, hz = (ci >> 10 for ci in c)
    blk = ((hx * 73856093) ^ (hy * 19349663) ^ (hz * 83492791)) & 0xFFFFFFFF
    return (((m + blk) & 0xFFFFFFFF) & (buckets - 1)).astype(np.uint32)


def state_checksum(words: np.ndarray) -> int:
    """FNV-1a 64 over the 32-bit words of .flk records, as fh_result.checksum."""
    h = 1469598103934665603
    for word in np.ascontiguousarray(words, np.float32).view(np.uint32).ravel().tolist():
        h ^= word
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h
