"""ctypes wrapper of the oracle libraries (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker.  The product never imports it.

  liboracle.so           C restatement of the reference path (bh_oracle.c)
  _ref/libref_strict.so  the reference's own sources, strict flags
  _ref/libref_fast.so    the reference's own sources, Release-like flags
The _ref libraries exist only where they were built (the development
container, and the GPU host when the built files travel with the tree).
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
_P, _I64, _D = C.c_void_p, C.c_int64, C.c_double
G = 1.48812e-34

_lib = None
_refs: dict = {}


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        p = HERE / "liboracle.so"
        if not p.exists():
            raise RuntimeError(f"{p} not built (make -C oracle)")
        L = C.CDLL(str(p))
        L.or_bbox.argtypes = [_I64, _P, _P]
        L.or_morton_keys.argtypes = [_I64, _P, _P, _P]
        L.or_sort.argtypes = [_I64, _P, _P, _P, _P, _P, _P]
        L.or_tree_build.argtypes = [_I64, _P, _P, _P]
        L.or_tree_build.restype = _P
        L.or_tree_free.argtypes = [_P]
        L.or_tree_free.restype = None
        L.or_tree_count.argtypes = [_P]
        L.or_tree_count.restype = _I64
        L.or_tree_export.argtypes = [_P, _P, _P, _P, _P]
        L.or_walk.argtypes = [_P, _I64, _P, _P, _P, _D, _D, _D, _P, _P, _P]
        L.or_direct.argtypes = [_I64, _P, _P, _D, _D, _I64, _P, _P]
        L.or_simulate.argtypes = [_I64, _P, _P, _P, _P, _P, _I64, _D, _D, _D, _D, _D]
        L.or_count_steps.argtypes = [_D, _D]
        L.or_count_steps.restype = _I64
        L.or_spread21.argtypes = [C.c_uint32]
        L.or_spread21.restype = C.c_uint64
        L.or_compact21.argtypes = [C.c_uint64]
        L.or_compact21.restype = C.c_uint32
        L.or_morton63.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_morton63.restype = C.c_uint64
        L.or_prefix.argtypes = [C.c_uint64, C.c_int]
        L.or_prefix.restype = C.c_uint64
        _lib = L
    return _lib


def ref_available(kind: str = "strict") -> bool:
    return (HERE / "_ref" / f"libref_{kind}.so").exists()


def ref(kind: str = "strict") -> C.CDLL:
    if kind not in _refs:
        p = HERE / "_ref" / f"libref_{kind}.so"
        if not p.exists():
            raise RuntimeError(f"{p} not built (make -C oracle ref, needs the reference checkout)")
        L = C.CDLL(str(p))
        L.ref_force.argtypes = [_I64, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P]
        L.ref_tree_count.restype = _I64
        L.ref_tree_export.argtypes = [_P, _P, _P, _P]
        L.ref_simulate.argtypes = [_I64, _P, _P, _P, _P, _P, _I64, _D, _D, _D, _D, _D, _P]
        _refs[kind] = L
    return _refs[kind]


def bbox(pos):
    pos = np.ascontiguousarray(pos, np.float64)
    bb = np.zeros(6)
    lib().or_bbox(len(pos), _ptr(pos), _ptr(bb))
    return bb


def morton_keys(pos, bb):
    pos = np.ascontiguousarray(pos, np.float64)
    k = np.zeros(len(pos), np.uint64)
    lib().or_morton_keys(len(pos), _ptr(pos), _ptr(np.ascontiguousarray(bb, np.float64)), _ptr(k))
    return k


class Tree:
    def __init__(self, keys, pos, mass):
        self.L = lib()
        self.h = self.L.or_tree_build(len(mass), _ptr(keys), _ptr(pos), _ptr(mass))
        if not self.h:
            raise RuntimeError("or_tree_build failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_tree_free(self.h)
            self.h = None

    def export(self):
        n = self.L.or_tree_count(self.h)
        prefix = np.zeros(n, np.uint64)
        depth = np.zeros(n, np.uint8)
        mass = np.zeros(n)
        com = np.zeros((n, 3))
        self.L.or_tree_export(self.h, _ptr(prefix), _ptr(depth), _ptr(mass), _ptr(com))
        return prefix, depth, mass, com

    def walk(self, keys, pos, bb, theta, G_=G, soft2=0.0, counts=False):
        n = len(keys)
        acc = np.zeros((n, 3))
        na = np.zeros(n, np.int64) if counts else None
        no = np.zeros(n, np.int64) if counts else None
        rc = self.L.or_walk(self.h, n, _ptr(keys), _ptr(pos), _ptr(np.ascontiguousarray(bb, np.float64)),
                            theta, G_, soft2, _ptr(acc), _ptr(na), _ptr(no))
        if rc != 0:
            raise RuntimeError(f"or_walk rc={rc}")
        return (acc, na, no) if counts else acc


def force_path(pos, mass, vel=None, ids=None, theta=1.05, G_=G, soft2=0.0, counts=False, tree=False):
    """bbox -> keys -> stable sort -> tree -> walk on copies of the inputs
    (src/main.cpp:212-263).  Returns a dict in sorted order."""
    n = len(mass)
    pos = np.array(pos, np.float64, order="C")
    mass = np.array(mass, np.float64)
    vel = np.zeros((n, 3)) if vel is None else np.array(vel, np.float64, order="C")
    ids = np.arange(n, dtype=np.uint64) if ids is None else np.array(ids, np.uint64)
    bb = bbox(pos)
    keys = morton_keys(pos, bb)
    perm = np.zeros(n, np.int64)
    lib().or_sort(n, _ptr(keys), _ptr(pos), _ptr(mass), _ptr(vel), _ptr(ids), _ptr(perm))
    tr = Tree(keys, pos, mass)
    out = dict(bb=bb, keys=keys, perm=perm, pos=pos, mass=mass, vel=vel, ids=ids)
    r = tr.walk(keys, pos, bb, theta, G_, soft2, counts)
    if counts:
        out["acc"], out["n_accept"], out["n_open"] = r
    else:
        out["acc"] = r
    if tree:
        out["tree"] = tr.export()
    return out


def direct(pos, mass, idx, G_=G, soft2=0.0):
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    idx = np.ascontiguousarray(idx, np.int64)
    acc = np.zeros((len(idx), 3))
    lib().or_direct(len(mass), _ptr(pos), _ptr(mass), G_, soft2, len(idx), _ptr(idx), _ptr(acc))
    return acc


def simulate(pos, vel, mass, nsteps, dt, theta=1.05, G_=G, soft2=0.0, t0=0.0, ids=None, acc=None):
    n = len(mass)
    st = dict(pos=np.array(pos, np.float64, order="C"), vel=np.array(vel, np.float64, order="C"),
              mass=np.array(mass, np.float64),
              ids=np.arange(n, dtype=np.uint64) if ids is None else np.array(ids, np.uint64),
              acc=np.zeros((n, 3)) if acc is None else np.array(acc, np.float64, order="C"))
    rc = lib().or_simulate(n, _ptr(st["pos"]), _ptr(st["vel"]), _ptr(st["mass"]), _ptr(st["ids"]), _ptr(st["acc"]),
                           nsteps, t0, dt, theta, G_, soft2)
    if rc != 0:
        raise RuntimeError(f"or_simulate rc={rc}")
    return st


def ref_force_path(pos, mass, theta=1.05, G_=G, soft2=0.0, kind="strict", tree=False):
    """The reference's own functions (oracle/_ref)."""
    L = ref(kind)
    n = len(mass)
    p = np.array(pos, np.float64, order="C")
    m = np.array(mass, np.float64)
    ids = np.arange(n, dtype=np.uint64)
    keys = np.zeros(n, np.uint64)
    acc = np.zeros((n, 3))
    bb = np.zeros(6)
    perm = np.zeros(n, np.int64)
    rc = L.ref_force(n, _ptr(p), _ptr(m), _ptr(ids), theta, G_, soft2, _ptr(keys), _ptr(acc), _ptr(bb), _ptr(perm))
    if rc != 0:
        raise RuntimeError(f"ref_force rc={rc}")
    out = dict(bb=bb, keys=keys, perm=perm, pos=p, mass=m, ids=ids, acc=acc)
    if tree:
        nt = L.ref_tree_count()
        prefix = np.zeros(nt, np.uint64)
        depth = np.zeros(nt, np.uint8)
        tm = np.zeros(nt)
        com = np.zeros((nt, 3))
        L.ref_tree_export(_ptr(prefix), _ptr(depth), _ptr(tm), _ptr(com))
        out["tree"] = (prefix, depth, tm, com)
    return out


def ref_simulate(pos, vel, mass, nsteps, dt, theta=1.05, G_=G, soft2=0.0, t0=0.0, kind="strict", acc=None):
    L = ref(kind)
    n = len(mass)
    st = dict(pos=np.array(pos, np.float64, order="C"), vel=np.array(vel, np.float64, order="C"),
              mass=np.array(mass, np.float64), ids=np.arange(n, dtype=np.uint64),
              acc=np.zeros((n, 3)) if acc is None else np.array(acc, np.float64, order="C"))
    ph = np.zeros(6)
    rc = L.ref_simulate(n, _ptr(st["pos"]), _ptr(st["vel"]), _ptr(st["mass"]), _ptr(st["ids"]), _ptr(st["acc"]),
                        nsteps, t0, dt, theta, G_, soft2, _ptr(ph))
    if rc != 0:
        raise RuntimeError(f"ref_simulate rc={rc}")
    st["phase_ms"] = ph
    return st
