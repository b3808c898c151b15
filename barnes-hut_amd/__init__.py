"""ctypes binding of libbh.so, the MI355X Barnes-Hut engine (C ABI: include/bh.h).

The package directory name carries a hyphen, so load it by path::

    import importlib.util, pathlib
    spec = importlib.util.spec_from_file_location(
        "bh_amd", pathlib.Path("barnes-hut_amd/__init__.py"),
        submodule_search_locations=[])
    bh = importlib.util.module_from_spec(spec); spec.loader.exec_module(bh)

`Engine` mirrors the reference's `--fc tree` step (src/main.cpp:191-281):
`compute_accelerations` is the drop-in for bbox -> keys -> sort -> build ->
exchange -> computeAccelerations (src/main.cpp:212-263) and `step` runs the
drift-first kick-drift-kick loop.  Arrays are numpy (n, 3) float64 in the
layout of `struct Vec3` (include/body.h:4-6).  Errors raise `BHError` with the
library's message; nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libbh.so"

OK = 0
ERR_ARG, ERR_NOMEM, ERR_HIP, ERR_STATE, ERR_IO, ERR_UNSUPPORTED, ERR_OVERFLOW = -1, -2, -3, -4, -5, -6, -7
FC_TREE, FC_LET, FC_LET_DIRECT = 0, 1, 2
FORM_PLAIN, FORM_MIXED, FORM_LEARNED = 0, 1, 2
SCHED_KEY_ORDER, SCHED_PREV_COST, SCHED_MODEL = 0, 1, 2
SORT_NEAR, SORT_FULL = 0, 1
GEN_BELT, GEN_KUIPER, GEN_TROJAN = 0, 1, 2
G = 1.48812e-34
MAX_DEPTH = 21
CENTRAL_MASS = 1.989e30


class BHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bh error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("device", C.c_int), ("fc_policy", C.c_int), ("form", C.c_int),
                ("theta", C.c_double), ("G", C.c_double), ("soft2", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("gmax", C.c_int), ("efac", C.c_double), ("occ", C.c_int), ("walk_blocks", C.c_int),
                ("best_ms", C.c_double), ("default_ms", C.c_double), ("n_tried", C.c_int)]


_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double

# (name, restype, argtypes) of every entry point declared in include/bh.h
SIGNATURES = [
    ("bh_config_default", None, [C.POINTER(Config)]),
    ("bh_create", C.c_int, [C.POINTER(_P), C.POINTER(Config)]),
    ("bh_destroy", None, [_P]),
    ("bh_last_error", C.c_char_p, [_P]),
    ("bh_set_bodies", C.c_int, [_P, _I64, _P, _P, _P, _P]),
    ("bh_num_bodies", _I64, [_P]),
    ("bh_get_bodies", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("bh_compute_accelerations", C.c_int, [_P, C.c_int, _D, _D, _D, _P]),
    ("bh_step", C.c_int, [_P, _I64, _D]),
    ("bh_time", _D, [_P]),
    ("bh_count_steps", _I64, [_D, _D]),
    ("bh_parse_time", C.c_int, [C.c_char_p, C.POINTER(_D)]),
    ("bh_export_keys", C.c_int, [_P, _P]),
    ("bh_export_bbox", C.c_int, [_P, _P]),
    ("bh_tree_size", C.c_int, [_P, C.POINTER(_I64), C.POINTER(_I64)]),
    ("bh_export_topology", C.c_int, [_P, _P, _P, _P, _P]),
    ("bh_set_counting", C.c_int, [_P, C.c_int]),
    ("bh_get_counters", C.c_int, [_P, _P, _P, _P]),
    ("bh_debug_raise_error", C.c_int, [_P, C.c_uint32]),
    ("bh_get_phase_ms", C.c_int, [_P, _P]),
    ("bh_get_step_phase_ms", C.c_int, [_P, _P, C.POINTER(_I64)]),
    ("bh_direct_sum", C.c_int, [_P, _I64, _P, _D, _D, _P]),
    ("bh_get_groups", C.c_int, [_P, C.POINTER(_I64), _I64, _P, _P, _P]),
    ("bh_set_schedule", C.c_int, [_P, C.c_int]),
    ("bh_get_group_timeline", C.c_int, [_P, _I64, _P, _P, _P]),
    ("bh_get_group_work", C.c_int, [_P, _I64, _P]),
    ("bh_get_group_costs", C.c_int, [_P, _I64, _P]),
    ("bh_model_configure", C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_float]),
    ("bh_model_reset", C.c_int, [_P, C.c_uint64]),
    ("bh_model_get_params", C.c_int, [_P, _P]),
    ("bh_model_set_params", C.c_int, [_P, _P]),
    ("bh_model_forward", C.c_int, [_P, _I64, _P, _P]),
    ("bh_model_gradients", C.c_int, [_P, _I64, _P, _P, _P, C.POINTER(C.c_float)]),
    ("bh_model_train", C.c_int, [_P, _I64, _P, _P, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    ("bh_model_status", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(C.c_float)]),
    ("bh_get_group_features", C.c_int, [_P, _I64, _P]),
    ("bh_generate", C.c_int, [C.c_int, C.c_uint64, _I64, _P, _P, _P]),
    ("bh_generate_ensemble_sizes", C.c_int, [C.c_uint64, _I64, _I64, _I64, _P]),
    ("bh_write_csv", C.c_int, [C.c_char_p, _I64, _P, _P, _P]),
    ("bh_read_csv", C.c_int, [C.c_char_p, _I64, _I64, C.POINTER(_I64), _P, _P, _P, _P]),
    ("bh_set_graph", C.c_int, [_P, C.c_int]),
    ("bh_get_graph_replays", C.c_int, [_P, C.POINTER(_I64)]),
    ("bh_set_systems", C.c_int, [_P, _I64, _P]),
    ("bh_get_systems", C.c_int, [_P, C.POINTER(_I64), _I64, _P]),
    ("bh_set_group_params", C.c_int, [_P, C.c_int, _D, C.c_int, C.c_int]),
    ("bh_get_group_params", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(_D), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("bh_autotune", C.c_int, [_P, C.c_int, C.c_int, _P]),
    ("bh_set_relative_opening", C.c_int, [_P, _D]),
    ("bh_set_schedule_split", C.c_int, [_P, C.c_int]),
    ("bh_set_sort_mode", C.c_int, [_P, C.c_int]),
    ("bh_set_moment_cutoff", C.c_int, [_P, C.c_int]),
    ("bh_get_sort_stats", C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(_I64)]),
    ("bh_get_relative_opening", C.c_int, [_P, C.POINTER(_D)]),
    ("bh_compute_accelerations_static", C.c_int, [_I64, _P, _P, _D, _D, _D, _P, _P]),
    ("bh_static_last_error", C.c_char_p, []),
]

_lib = None


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libbh.so (built in-tree by `make -C barnes-hut_amd`)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise BHError(ERR_STATE, f"{p} is not built (run __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(lib, rc: int, ctx=None):
    if rc != OK:
        msg = lib.bh_last_error(ctx)
        raise BHError(rc, msg.decode() if msg else "")


def generate(kind: int, seed: int, n: int):
    """Synthetic system: returns pos (n,3), vel (n,3), mass (n,)."""
    lib = load()
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    rc = lib.bh_generate(kind, seed, n, _ptr(pos), _ptr(vel), _ptr(mass))
    if rc != OK:
        raise BHError(rc, f"bh_generate(kind={kind}, seed={seed}, n={n})")
    return pos, vel, mass


def ensemble_sizes(seed: int, count: int, nmin: int, nmax: int) -> np.ndarray:
    lib = load()
    out = np.zeros(count, np.int64)
    rc = lib.bh_generate_ensemble_sizes(seed, count, nmin, nmax, _ptr(out))
    if rc != OK:
        raise BHError(rc, "bh_generate_ensemble_sizes")
    return out


def write_csv(path, pos, vel, mass):
    lib = load()
    pos, vel, mass = (np.ascontiguousarray(a, np.float64) for a in (pos, vel, mass))
    rc = lib.bh_write_csv(str(path).encode(), len(mass), _ptr(pos), _ptr(vel), _ptr(mass))
    if rc != OK:
        raise BHError(rc, f"bh_write_csv({path})")


def read_csv(path, body_limit: int = -1):
    """Returns ids, mass, pos, vel like readCSV (src/io.cpp:16-81)."""
    lib = load()
    n = _I64(0)
    rc = lib.bh_read_csv(str(path).encode(), body_limit, 0, C.byref(n), None, None, None, None)
    if rc != OK:
        raise BHError(rc, f"bh_read_csv({path})")
    cnt = n.value
    pos = np.zeros((cnt, 3))
    vel = np.zeros((cnt, 3))
    mass = np.zeros(cnt)
    ids = np.zeros(cnt, np.uint64)
    rc = lib.bh_read_csv(str(path).encode(), body_limit, cnt, C.byref(n), _ptr(pos), _ptr(vel), _ptr(mass), _ptr(ids))
    if rc != OK:
        raise BHError(rc, f"bh_read_csv({path})")
    return ids, mass, pos, vel


def count_steps(dt: float, t_end: float) -> int:
    return load().bh_count_steps(dt, t_end)


def parse_time(text: str) -> float:
    out = _D(0.0)
    rc = load().bh_parse_time(text.encode(), C.byref(out))
    if rc != OK:
        raise BHError(rc, f"cannot parse time {text!r}")
    return out.value


class Engine:
    """One device context (bh_ctx)."""

    def __init__(self, device: int = 0, form: int = FORM_PLAIN, theta: float = 1.05,
                 G_: float = G, soft2: float = 0.0, fc_policy: int = FC_TREE):
        self.lib = load()
        cfg = Config()
        self.lib.bh_config_default(C.byref(cfg))
        cfg.device, cfg.form, cfg.theta, cfg.G, cfg.soft2, cfg.fc_policy = device, form, theta, G_, soft2, fc_policy
        h = _P()
        rc = self.lib.bh_create(C.byref(h), C.byref(cfg))
        if rc != OK:
            raise BHError(rc, self.lib.bh_last_error(None).decode())
        self.h = h
        self.n = 0
        self.form = form

    def close(self):
        if getattr(self, "h", None):
            self.lib.bh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        _check(self.lib, rc, self.h)

    def set_bodies(self, pos, vel, mass, ids=None):
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        mass = np.ascontiguousarray(mass, np.float64)
        idp = None if ids is None else np.ascontiguousarray(ids, np.uint64)
        self.n = len(mass)
        self._ck(self.lib.bh_set_bodies(self.h, self.n, _ptr(pos), _ptr(vel), _ptr(mass), _ptr(idp)))

    def get_bodies(self):
        n = self.n
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        mass, ids = np.zeros(n), np.zeros(n, np.uint64)
        self._ck(self.lib.bh_get_bodies(self.h, _ptr(pos), _ptr(vel), _ptr(mass), _ptr(ids), _ptr(acc)))
        return dict(pos=pos, vel=vel, mass=mass, ids=ids, acc=acc)

    def compute_accelerations(self, form: int = FORM_PLAIN, theta: float = 1.05, G_: float = G,
                              soft2: float = 0.0) -> np.ndarray:
        acc = np.zeros((self.n, 3))
        self._ck(self.lib.bh_compute_accelerations(self.h, form, theta, G_, soft2, _ptr(acc)))
        return acc

    def step(self, nsteps: int, dt: float):
        self._ck(self.lib.bh_step(self.h, nsteps, dt))

    @property
    def time(self) -> float:
        return self.lib.bh_time(self.h)

    def keys(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint64)
        self._ck(self.lib.bh_export_keys(self.h, _ptr(out)))
        return out

    def bbox(self) -> np.ndarray:
        out = np.zeros(6)
        self._ck(self.lib.bh_export_bbox(self.h, _ptr(out)))
        return out

    def tree_size(self):
        a, b = _I64(0), _I64(0)
        self._ck(self.lib.bh_tree_size(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def topology(self):
        nf, _ = self.tree_size()
        prefix = np.zeros(nf, np.uint64)
        depth = np.zeros(nf, np.uint8)
        mass = np.zeros(nf)
        com = np.zeros((nf, 3))
        self._ck(self.lib.bh_export_topology(self.h, _ptr(prefix), _ptr(depth), _ptr(mass), _ptr(com)))
        return prefix, depth, mass, com

    def set_counting(self, on: bool):
        self._ck(self.lib.bh_set_counting(self.h, 1 if on else 0))

    def counters(self):
        """(accepted, opened full-chain, tested compressed) per body."""
        a = np.zeros(self.n, np.int64)
        o = np.zeros(self.n, np.int64)
        t = np.zeros(self.n, np.int64)
        self._ck(self.lib.bh_get_counters(self.h, _ptr(a), _ptr(o), _ptr(t)))
        return a, o, t

    def debug_raise_error(self, bits: int):
        """Test hook: raise device guard bits on the context stream."""
        self._ck(self.lib.bh_debug_raise_error(self.h, bits))

    def phase_ms(self) -> np.ndarray:
        out = np.zeros(5)
        self._ck(self.lib.bh_get_phase_ms(self.h, _ptr(out)))
        return out

    def step_phase_ms(self):
        out = np.zeros(5)
        k = _I64(0)
        self._ck(self.lib.bh_get_step_phase_ms(self.h, _ptr(out), C.byref(k)))
        return out, k.value

    def set_schedule(self, mode: int):
        self._ck(self.lib.bh_set_schedule(self.h, mode))

    def set_schedule_split(self, chunks_per_worker: int):
        self._ck(self.lib.bh_set_schedule_split(self.h, chunks_per_worker))

    def set_moment_cutoff(self, cutoff: int):
        """Depth from which one workgroup sums the tree moments in a single
        launch (-1: chosen when a step graph is captured; 21: none)."""
        self._ck(self.lib.bh_set_moment_cutoff(self.h, cutoff))

    def set_sort_mode(self, mode: int):
        """SORT_NEAR (default): sort only the keys out of order since the
        previous step and merge; SORT_FULL: full radix sort every step."""
        self._ck(self.lib.bh_set_sort_mode(self.h, mode))

    def sort_stats(self):
        """(took_near_path, keys_radix_sorted) of the last key sort."""
        near = C.c_int(0)
        k = _I64(0)
        self._ck(self.lib.bh_get_sort_stats(self.h, C.byref(near), C.byref(k)))
        return bool(near.value), int(k.value)

    # ------------------------------------------------ graphs, ensembles, tuning
    def set_graph(self, on: bool):
        self._ck(self.lib.bh_set_graph(self.h, 1 if on else 0))

    def graph_replays(self) -> int:
        r = _I64(0)
        self._ck(self.lib.bh_get_graph_replays(self.h, C.byref(r)))
        return r.value

    def set_systems(self, counts):
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        self._ck(self.lib.bh_set_systems(self.h, len(counts), _ptr(counts)))

    def systems(self) -> np.ndarray:
        """Body offsets of the systems (S + 1 entries)."""
        k = _I64(0)
        self._ck(self.lib.bh_get_systems(self.h, C.byref(k), 0, None))
        out = np.zeros(k.value + 1, np.int64)
        self._ck(self.lib.bh_get_systems(self.h, C.byref(k), len(out), _ptr(out)))
        return out

    def set_group_params(self, gmax=64, efac=2.0, occ=32, walk_blocks=0):
        self._ck(self.lib.bh_set_group_params(self.h, gmax, efac, occ, walk_blocks))

    def set_relative_opening(self, alpha: float):
        self._ck(self.lib.bh_set_relative_opening(self.h, alpha))

    def relative_opening(self) -> float:
        a = _D(0)
        self._ck(self.lib.bh_get_relative_opening(self.h, C.byref(a)))
        return a.value

    def group_params(self) -> dict:
        g, o, w = C.c_int(0), C.c_int(0), C.c_int(0)
        e = _D(0)
        self._ck(self.lib.bh_get_group_params(self.h, C.byref(g), C.byref(e), C.byref(o), C.byref(w)))
        return {"gmax": g.value, "efac": e.value, "occ": o.value, "walk_blocks": w.value}

    def autotune(self, form: int = None, reps: int = 3) -> dict:
        r = TuneResult()
        self._ck(self.lib.bh_autotune(self.h, self.form if form is None else form, reps, C.byref(r)))
        return {"gmax": r.gmax, "efac": r.efac, "occ": r.occ, "walk_blocks": r.walk_blocks,
                "best_ms": r.best_ms, "default_ms": r.default_ms, "n_tried": r.n_tried}

    # ------------------------------------------------ learned cost model
    MODEL_PARAMS = 1377

    def model_configure(self, train_steps=50, iters_per_step=16, batch=2048, lr=3e-3):
        self._ck(self.lib.bh_model_configure(self.h, train_steps, iters_per_step, batch, lr))

    def model_reset(self, seed: int):
        self._ck(self.lib.bh_model_reset(self.h, seed))

    def model_get_params(self) -> np.ndarray:
        out = np.zeros(self.MODEL_PARAMS, np.float32)
        self._ck(self.lib.bh_model_get_params(self.h, _ptr(out)))
        return out

    def model_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._ck(self.lib.bh_model_set_params(self.h, _ptr(p)))

    def model_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(len(x), np.float32)
        self._ck(self.lib.bh_model_forward(self.h, len(x), _ptr(x), _ptr(out)))
        return out

    def model_gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.MODEL_PARAMS, np.float32)
        loss = C.c_float(0)
        self._ck(self.lib.bh_model_gradients(self.h, len(x), _ptr(x), _ptr(t), _ptr(g), C.byref(loss)))
        return g, loss.value

    def model_train(self, x, t, iters=100, batch=1024, lr=3e-3) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float(0)
        self._ck(self.lib.bh_model_train(self.h, len(x), _ptr(x), _ptr(t), iters, batch, lr, C.byref(loss)))
        return loss.value

    def model_status(self):
        r, s, u, l = C.c_int(0), _I64(0), _I64(0), C.c_float(0)
        self._ck(self.lib.bh_model_status(self.h, C.byref(r), C.byref(s), C.byref(u), C.byref(l)))
        return bool(r.value), s.value, u.value, l.value

    def group_features(self) -> np.ndarray:
        k = _I64(0)
        self._ck(self.lib.bh_get_groups(self.h, C.byref(k), 0, None, None, None))
        out = np.zeros((k.value, 8), np.float32)
        self._ck(self.lib.bh_get_group_features(self.h, k.value, _ptr(out)))
        return out

    def group_timeline(self):
        """(t_start, t_end) in 100 MHz device ticks and worker per group."""
        k = _I64(0)
        self._ck(self.lib.bh_get_groups(self.h, C.byref(k), 0, None, None, None))
        t0 = np.zeros(k.value, np.uint64)
        t1 = np.zeros(k.value, np.uint64)
        w = np.zeros(k.value, np.uint32)
        self._ck(self.lib.bh_get_group_timeline(self.h, k.value, _ptr(t0), _ptr(t1), _ptr(w)))
        return t0, t1, w

    def group_work(self):
        """Walk work per group of the last mixed walk: list cost + visits / 2."""
        k = _I64(0)
        self._ck(self.lib.bh_get_groups(self.h, C.byref(k), 0, None, None, None))
        work = np.zeros(k.value, np.uint32)
        self._ck(self.lib.bh_get_group_work(self.h, k.value, _ptr(work)))
        return work

    def group_costs(self):
        """Estimated cost per group that the last walk's heaviest-first order used."""
        k = _I64(0)
        self._ck(self.lib.bh_get_groups(self.h, C.byref(k), 0, None, None, None))
        cost = np.zeros(k.value, np.float32)
        self._ck(self.lib.bh_get_group_costs(self.h, k.value, _ptr(cost)))
        return cost

    def groups(self):
        """Walk groups of the last mixed evaluation: first_body (G+1),
        interactions (G), visited (G)."""
        k = _I64(0)
        self._ck(self.lib.bh_get_groups(self.h, C.byref(k), 0, None, None, None))
        g = k.value
        first = np.zeros(g + 1, np.uint32)
        inter = np.zeros(g, np.uint32)
        visit = np.zeros(g, np.uint32)
        self._ck(self.lib.bh_get_groups(self.h, C.byref(k), g, _ptr(first), _ptr(inter), _ptr(visit)))
        return first, inter, visit

    def direct_sum(self, idx, G_: float = G, soft2: float = 0.0) -> np.ndarray:
        """FP64 direct-sum accelerations (device) for bodies idx of the current order."""
        idx = np.ascontiguousarray(idx, np.int64)
        out = np.zeros((len(idx), 3))
        self._ck(self.lib.bh_direct_sum(self.h, len(idx), _ptr(idx), G_, soft2, _ptr(out)))
        return out
