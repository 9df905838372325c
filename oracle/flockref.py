"""ctypes binding of the CPU oracle (oracle/_build/libflockref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's CPU-baseline leg. The product (libflock_hip.so and
the flock_hip package) never imports it.

The oracle restates crates/flock-ref (see oracle/flockref.h for the file:line
map). State is exchanged as .flk records, numpy float32 arrays of shape (n, 15).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

WORDS = 15
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libflockref.so")
CLI_PATH = os.path.join(HERE, "_build", "flock-ref")

_lib = None
_P = C.c_void_p
_u32 = C.c_uint32
_SIGS = {
    "fr_constants": ([_P, C.c_int], C.c_int),
    "fr_radius_step": ([C.c_float, _u32], C.c_float),
    "fr_splitmix_next": ([_P], C.c_uint64),
    "fr_splitmix_unit": ([_P], C.c_float),
    "fr_scatter": ([_u32, C.c_uint64, _P], None),
    "fr_forward": ([_P, _P], None),
    "fr_side_of": ([_P, _P], None),
    "fr_body_up": ([_P, C.c_float, _P], None),
    "fr_decide": ([_P, _u32, _u32, C.c_float, _P], None),
    "fr_step_bird": ([_P, C.c_float], None),
    "fr_step": ([_P, _u32, _P, _u32, C.c_int], None),
    "fr_count_due": ([_P, _u32, C.c_float], _u32),
    "fr_stats": ([_P, _u32, _P], None),
    "fr_diff": ([_P, _P, _u32, C.c_float, C.c_float, _u32, C.c_float, _P, _P], C.c_int),
    "fr_statdiff": ([_P, _P, _u32, _P], C.c_int),
    "fr_read_flk": ([C.c_char_p, _P, _P, _P, _u32], C.c_int),
    "fr_write_flk": ([C.c_char_p, _u32, C.c_float, _P], C.c_int),
    "fr_buckets_for": ([_u32], _u32),
    "fr_cell_coord": ([C.c_float, C.c_float], C.c_int32),
    "fr_bucket_of": ([C.c_int32, C.c_int32, C.c_int32, _u32], _u32),
    "fr_cell_intersects_sphere": ([C.c_int32, C.c_int32, C.c_int32, C.c_float, _P, C.c_float], C.c_int),
    "fr_grid": ([_P, _u32, _u32, C.c_float, _P, _P, _P, _P], None),
    "fr_check_grid": ([_P, _u32, _P, _P, _u32, C.c_char_p, C.c_int], C.c_int),
    "fr_neighbours_brute": ([_P, _u32, _u32, C.c_int, _P, _u32], _u32),
    "fr_decide_grid": ([_P, _u32, _u32, C.c_float, C.c_float, _u32, _P, _P, _P], None),
    "fr_step_grid": ([_P, _u32, _P, _u32, C.c_float, _u32, C.c_int], None),
    "fr_decide_many": ([_P, _u32, _P, _u32, C.c_float, _P, C.c_int], None),
    "fr_decide_grid_many": ([_P, _u32, _P, _u32, C.c_float, C.c_float, _u32, _P, _P, _P, C.c_int], None),
}

Z0 = np.float32(100.0)


def build() -> None:
    """Compiles the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _w(words: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(words, np.float32)
    assert w.ndim == 2 and w.shape[1] == WORDS, w.shape
    return w


def constants() -> np.ndarray:
    n = lib().fr_constants(None, 0)
    out = np.zeros(n, np.float32)
    lib().fr_constants(out.ctypes.data, n)
    return out


def radius_step(r: float, n: int) -> np.float32:
    return np.float32(lib().fr_radius_step(np.float32(r), n))


def scatter(n: int, seed: int = 1) -> np.ndarray:
    w = np.zeros((n, WORDS), np.float32)
    lib().fr_scatter(n, seed, w.ctypes.data)
    return w


def vec(fn: str, v, *extra) -> np.ndarray:
    a = np.asarray(v, np.float32)
    out = np.zeros(3, np.float32)
    getattr(lib(), fn)(a.ctypes.data, *extra, out.ctypes.data)
    return out


def forward(v) -> np.ndarray:
    return vec("fr_forward", v)


def side_of(f) -> np.ndarray:
    return vec("fr_side_of", f)


def body_up(f, beta: float) -> np.ndarray:
    return vec("fr_body_up", f, np.float32(beta))


def decide(words: np.ndarray, i: int, z0: float = Z0) -> dict:
    """model::decide (model.rs:53-145), brute force."""
    w = _w(words)
    out = np.zeros(6, np.float32)
    lib().fr_decide(w.ctypes.data, w.shape[0], i, np.float32(z0), out.ctypes.data)
    return {"force": out[:3].copy(), "radius": out[3], "neighbours": int(out[4:5].view(np.uint32)[0]),
            "centrality": out[5]}


def step_bird(record: np.ndarray, dt: float = 0.005) -> np.ndarray:
    r = np.ascontiguousarray(record, np.float32).reshape(WORDS).copy()
    lib().fr_step_bird(r.ctypes.data, np.float32(dt))
    return r


def step(words: np.ndarray, time: float, steps: int, threads: int = 1) -> tuple[np.ndarray, float]:
    """`steps` x Flock::step (model.rs:182-200), brute-force neighbours."""
    w = _w(words).copy()
    t = C.c_float(time)
    lib().fr_step(w.ctypes.data, w.shape[0], C.addressof(t), steps, threads)
    return w, float(t.value)


def step_grid(words: np.ndarray, time: float, steps: int, cell: float = 4.0, buckets: int = 0,
              threads: int = 1) -> tuple[np.ndarray, float]:
    """Flock::step with the grid-walk decide of cuda/flock.cu:190-290."""
    w = _w(words).copy()
    b = buckets or buckets_for(w.shape[0])
    t = C.c_float(time)
    lib().fr_step_grid(w.ctypes.data, w.shape[0], C.addressof(t), steps, np.float32(cell), b, threads)
    return w, float(t.value)


def count_due(words: np.ndarray, t: float) -> int:
    w = _w(words)
    return int(lib().fr_count_due(w.ctypes.data, w.shape[0], np.float32(t)))


def stats(words: np.ndarray) -> np.ndarray:
    w = _w(words)
    out = np.zeros(4, np.float32)
    lib().fr_stats(w.ctypes.data, w.shape[0], out.ctypes.data)
    return out


def diff(a: np.ndarray, b: np.ndarray, pos_tol=1e-3, vel_tol=1e-3, allow_outliers=0, outlier_max=1e-2) -> dict:
    """flock-ref diff (main.rs:136-208)."""
    a, b = _w(a), _w(b)
    out = np.zeros(7, np.float32)
    cnt = np.zeros(2, np.uint32)
    ok = lib().fr_diff(a.ctypes.data, b.ctypes.data, a.shape[0], np.float32(pos_tol), np.float32(vel_tol),
                       allow_outliers, np.float32(outlier_max), out.ctypes.data, cnt.ctypes.data)
    keys = ["dpos", "dvel", "dradius", "dcentrality", "dsteering", "droll", "dspin"]
    d = {k: float(v) for k, v in zip(keys, out)}
    d.update(nb_mismatch=int(cnt[0]), outliers=int(cnt[1]), ok=bool(ok))
    return d


def statdiff(a: np.ndarray, b: np.ndarray) -> dict:
    """flock-ref statdiff (main.rs:104-133), tolerances [0.02, 0.1, 0.01, 0.005]."""
    a, b = _w(a), _w(b)
    out = np.zeros(8, np.float32)
    ok = lib().fr_statdiff(a.ctypes.data, b.ctypes.data, a.shape[0], out.ctypes.data)
    return {"a": out[:4].copy(), "b": out[4:].copy(), "ok": bool(ok)}


def read_flk(path: str) -> tuple[np.ndarray, float, int]:
    """Returns (words, time, status); status 0 ok, 1 open, 2 magic, 3 words, 4 truncated."""
    n, t = C.c_uint32(0), C.c_float(0)
    rc = lib().fr_read_flk(str(path).encode(), C.addressof(n), C.addressof(t), None, 0)
    if rc:
        return np.zeros((0, WORDS), np.float32), 0.0, rc
    w = np.zeros((n.value, WORDS), np.float32)
    rc = lib().fr_read_flk(str(path).encode(), C.addressof(n), C.addressof(t), w.ctypes.data, n.value)
    return w, float(t.value), rc


def write_flk(path: str, words: np.ndarray, time: float) -> int:
    w = _w(words)
    return lib().fr_write_flk(str(path).encode(), w.shape[0], np.float32(time), w.ctypes.data)


def buckets_for(n: int) -> int:
    return int(lib().fr_buckets_for(n))


def cell_coord(p: float, inv_cell: float = 0.25) -> int:
    return int(lib().fr_cell_coord(np.float32(p), np.float32(inv_cell)))


def bucket_of(cx: int, cy: int, cz: int, mask: int) -> int:
    return int(lib().fr_bucket_of(cx, cy, cz, mask))


def cell_intersects_sphere(c, p, r2: float, cell: float = 4.0) -> bool:
    pa = np.asarray(p, np.float32)
    return bool(lib().fr_cell_intersects_sphere(int(c[0]), int(c[1]), int(c[2]), np.float32(cell),
                                                pa.ctypes.data, np.float32(r2)))


def grid(words: np.ndarray, buckets: int = 0, cell: float = 4.0):
    """Stable counting sort: (keys[n], counts[B], offsets[B+1], sorted[n])."""
    w = _w(words)
    n = w.shape[0]
    b = buckets or buckets_for(n)
    keys = np.zeros(n, np.uint32)
    counts = np.zeros(b, np.uint32)
    offsets = np.zeros(b + 1, np.uint32)
    srt = np.zeros(n, np.uint32)
    lib().fr_grid(w.ctypes.data, n, b, np.float32(cell), keys.ctypes.data, counts.ctypes.data,
                  offsets.ctypes.data, srt.ctypes.data)
    return keys, counts, offsets, srt


def check_grid(bucket, offsets, sorted_, buckets: int) -> tuple[int, str]:
    """check_grid (cutile/src/main.rs:304-362): (0, "") when valid."""
    b = np.ascontiguousarray(bucket, np.uint32)
    o = np.ascontiguousarray(offsets, np.uint32)
    s = np.ascontiguousarray(sorted_, np.uint32)
    msg = C.create_string_buffer(256)
    rc = lib().fr_check_grid(b.ctypes.data, b.size, o.ctypes.data, s.ctypes.data, buckets, msg, 256)
    return rc, msg.value.decode()


def neighbours_brute(words: np.ndarray, i: int, walk: int = 1, cap: int = 1 << 16) -> np.ndarray:
    w = _w(words)
    out = np.zeros(cap, np.uint32)
    c = lib().fr_neighbours_brute(w.ctypes.data, w.shape[0], i, walk, out.ctypes.data, cap)
    return out[: min(c, cap)].copy()


def decide_grid(words: np.ndarray, i: int, offsets, sorted_, buckets: int, cell: float = 4.0,
                z0: float = Z0) -> dict:
    w = _w(words)
    out = np.zeros(6, np.float32)
    o = np.ascontiguousarray(offsets, np.uint32)
    s = np.ascontiguousarray(sorted_, np.uint32)
    lib().fr_decide_grid(w.ctypes.data, w.shape[0], i, np.float32(z0), np.float32(cell), buckets,
                         o.ctypes.data, s.ctypes.data, out.ctypes.data)
    return {"force": out[:3].copy(), "radius": out[3], "neighbours": int(out[4:5].view(np.uint32)[0]),
            "centrality": out[5]}


def decide_many(words: np.ndarray, birds, z0: float = Z0, threads: int = 1) -> dict:
    """fr_decide (brute force) for many birds on `threads` threads: arrays
    force[k, 3], radius, neighbours, centrality."""
    w = _w(words)
    idx = np.ascontiguousarray(birds, np.uint32)
    out = np.zeros((idx.size, 6), np.float32)
    lib().fr_decide_many(w.ctypes.data, w.shape[0], idx.ctypes.data, idx.size, np.float32(z0), out.ctypes.data,
                         threads)
    return {"force": out[:, :3].copy(), "radius": out[:, 3].copy(),
            "neighbours": out[:, 4].copy().view(np.uint32), "centrality": out[:, 5].copy()}


def decide_grid_many(words: np.ndarray, birds, offsets, sorted_, buckets: int, cell: float = 4.0,
                     z0: float = Z0, threads: int = 1) -> dict:
    """fr_decide_grid for many birds: arrays force[k, 3], radius, neighbours, centrality."""
    w = _w(words)
    idx = np.ascontiguousarray(birds, np.uint32)
    out = np.zeros((idx.size, 6), np.float32)
    o = np.ascontiguousarray(offsets, np.uint32)
    s = np.ascontiguousarray(sorted_, np.uint32)
    lib().fr_decide_grid_many(w.ctypes.data, w.shape[0], idx.ctypes.data, idx.size, np.float32(z0),
                              np.float32(cell), buckets, o.ctypes.data, s.ctypes.data, out.ctypes.data, threads)
    return {"force": out[:, :3].copy(), "radius": out[:, 3].copy(),
            "neighbours": out[:, 4].copy().view(np.uint32), "centrality": out[:, 5].copy()}
