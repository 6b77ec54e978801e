"""Parity oracle for the raster-scan generalised geodesic distance transform.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, and only as the
checker or as the timed CPU baseline.  The product (``fastgeodis_amd``) never
imports it and has no CPU fallback.

Contents
--------
* ``geodesic2d`` / ``geodesic3d`` / ``signed_*`` / ``gsf*``: numpy-in,
  numpy-out wrappers over ``fg_oracle.c`` (the C restatement of
  ``FastGeodis/fastgeodis_cpu.cpp`` and the compositions in
  ``FastGeodis/fastgeodis.cpp:241-258, 319-343``).  Arguments follow the
  ``FastGeodisCpp`` module: explicit ``l_grad`` and ``l_eucl``.
* ``py_*``: the same entry points with the Python wrappers' argument handling
  (``FastGeodis/__init__.py:63-65, 96-98, 127-129, 160-162, 553, 585``):
  ``l_grad = fl32(lamb)`` and ``l_eucl = fl32(1.0 - lamb)`` computed in double.
* ``reference_module()``: the reference's own CPU extension compiled from its
  sources into ``oracle/_ref`` by ``oracle/build.py`` (None when absent).
* ``dijkstra``: float32 shortest-path brute force for tiny grids (see
  ``oracle/dijkstra.py``).

Pinning (what checks each function): ``geodesic2d/3d``, ``signed_*`` and
``gsf*`` are compared bit-for-bit with the reference build on seeded inputs
and with committed golden fixtures (``tests/golden``); the converged transform
is compared bit-for-bit with the Dijkstra brute force; closed forms (iter 0,
zero/one masks, lamb=1 constant image, Euclidean chamfer) are checked in
``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libfg_oracle.so")
_REF_DIR = os.path.join(_HERE, "_ref")

_lib = None
_f32p = ctypes.POINTER(ctypes.c_float)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.isfile(_LIB_PATH):
        from . import build as _build

        _build.build_oracle()
    lib = ctypes.CDLL(_LIB_PATH)
    i, f, p = ctypes.c_int, ctypes.c_float, _f32p
    lib.fgo_geodesic2d.argtypes = [p, p, p, i, i, i, i, f, f, f, i]
    lib.fgo_geodesic3d.argtypes = [p, p, p, i, i, i, i, i, p, f, f, f, i]
    lib.fgo_signed_geodesic2d.argtypes = [p, p, p, i, i, i, i, f, f, f, i]
    lib.fgo_signed_geodesic3d.argtypes = [p, p, p, i, i, i, i, i, p, f, f, f, i]
    lib.fgo_gsf2d.argtypes = [p, p, p, i, i, i, i, f, f, f, i]
    lib.fgo_gsf3d.argtypes = [p, p, p, i, i, i, i, i, f, p, f, f, i]
    for name in ("fgo_geodesic2d", "fgo_geodesic3d", "fgo_signed_geodesic2d",
                 "fgo_signed_geodesic3d", "fgo_gsf2d", "fgo_gsf3d"):
        getattr(lib, name).restype = ctypes.c_int
    lib.fgo_geodesic2d_layout.argtypes = [p, p, p, i, i, i, i, f, f, f, i]
    lib.fgo_geodesic3d_layout.argtypes = [p, p, p, i, i, i, i, i, p, f, f, f, i]
    lib.fgo_geodesic2d_layout.restype = ctypes.c_int
    lib.fgo_geodesic3d_layout.restype = ctypes.c_int
    lib.fgo_toivanen2d.argtypes = [p, p, p, i, i, i, i, f, f, f, i]
    lib.fgo_toivanen3d.argtypes = [p, p, p, i, i, i, i, i, p, f, f, f, i]
    lib.fgo_toivanen2d.restype = ctypes.c_int
    lib.fgo_toivanen3d.restype = ctypes.c_int
    lib.fgo_num_threads.restype = ctypes.c_int
    lib.fgo_set_num_threads.argtypes = [i]
    lib.fgo_set_num_threads.restype = None
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_f32p)


def _prep(image, mask, nd):
    image = np.ascontiguousarray(image, dtype=np.float32)
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    if image.ndim != nd + 2 or mask.ndim != nd + 2:
        raise ValueError("oracle expects (1, C, spatial...) arrays")
    if image.shape[0] != 1 or mask.shape[0] != 1 or image.shape[2:] != mask.shape[2:]:
        raise ValueError("oracle expects batch 1 and matching spatial shapes")
    return image, mask, np.empty_like(mask)


def _spacing(spacing):
    s = np.ascontiguousarray(np.asarray(spacing, dtype=np.float32).reshape(3))
    return s


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle call failed with code {rc}")


def num_threads() -> int:
    return _load().fgo_num_threads()


def set_num_threads(n: int) -> None:
    """OpenMP thread count of later oracle calls."""
    _load().fgo_set_num_threads(int(n))


# ---- layout-faithful forms: the CPU baseline ------------------------------
# Same results as geodesic2d/3d; they copy the image and distance into each
# pass's layout like the reference's transposes (fastgeodis_cpu.cpp:180-195,
# 361-401), so their speed reflects the reference's CPU path.

def geodesic2d_layout(image, mask, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 2)
    _, C, H, W = image.shape
    _check(_load().fgo_geodesic2d_layout(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1],
                                         H, W, v, l_grad, l_eucl, int(iterations)))
    return out


def geodesic3d_layout(image, mask, spacing, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 3)
    _, C, D, H, W = image.shape
    s = _spacing(spacing)
    _check(_load().fgo_geodesic3d_layout(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1],
                                         D, H, W, _ptr(s), v, l_grad, l_eucl, int(iterations)))
    return out


# ---- FastGeodisCpp-level entry points (explicit l_grad / l_eucl) ---------

def geodesic2d(image, mask, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 2)
    _, C, H, W = image.shape
    _check(_load().fgo_geodesic2d(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1],
                                  H, W, v, l_grad, l_eucl, int(iterations)))
    return out


def geodesic3d(image, mask, spacing, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 3)
    _, C, D, H, W = image.shape
    s = _spacing(spacing)
    _check(_load().fgo_geodesic3d(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1],
                                  D, H, W, _ptr(s), v, l_grad, l_eucl, int(iterations)))
    return out


def signed_geodesic2d(image, mask, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 2)
    _, C, H, W = image.shape
    _check(_load().fgo_signed_geodesic2d(_ptr(image), _ptr(mask), _ptr(out), C,
                                         mask.shape[1], H, W, v, l_grad, l_eucl,
                                         int(iterations)))
    return out


def signed_geodesic3d(image, mask, spacing, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 3)
    _, C, D, H, W = image.shape
    s = _spacing(spacing)
    _check(_load().fgo_signed_geodesic3d(_ptr(image), _ptr(mask), _ptr(out), C,
                                         mask.shape[1], D, H, W, _ptr(s), v, l_grad,
                                         l_eucl, int(iterations)))
    return out


def gsf2d(image, mask, theta, v, lamb, iterations):
    image, mask, out = _prep(image, mask, 2)
    _, C, H, W = image.shape
    _check(_load().fgo_gsf2d(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1], H, W,
                             theta, v, lamb, int(iterations)))
    return out


def gsf3d(image, mask, theta, spacing, v, lamb, iterations):
    image, mask, out = _prep(image, mask, 3)
    _, C, D, H, W = image.shape
    s = _spacing(spacing)
    _check(_load().fgo_gsf3d(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1], D, H,
                             W, theta, _ptr(s), v, lamb, int(iterations)))
    return out


# ---- Toivanen's raster scan: accuracy cross-check ---------------------------
# (geodis_toivanen.cpp; called like FastGeodisCpp.generalised_geodesic*_toivanen)

def toivanen2d(image, mask, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 2)
    _, C, H, W = image.shape
    _check(_load().fgo_toivanen2d(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1], H, W,
                                  v, l_grad, l_eucl, int(iterations)))
    return out


def toivanen3d(image, mask, spacing, v, l_grad, l_eucl, iterations):
    image, mask, out = _prep(image, mask, 3)
    _, C, D, H, W = image.shape
    s = _spacing(spacing)
    _check(_load().fgo_toivanen3d(_ptr(image), _ptr(mask), _ptr(out), C, mask.shape[1], D, H,
                                  W, _ptr(s), v, l_grad, l_eucl, int(iterations)))
    return out


# ---- Python-wrapper-level entry points (FastGeodis/__init__.py) -----------

def wrapper_weights(lamb):
    """(l_grad, l_eucl) as the Python wrappers pass them: 1 - lamb in double,
    both narrowed to float32 at the binding (FastGeodis/__init__.py:63-65)."""
    return np.float32(lamb), np.float32(1.0 - float(lamb))


def py_geodesic2d(image, softmask, v, lamb, iter=2):
    lg, le = wrapper_weights(lamb)
    return geodesic2d(image, softmask, v, lg, le, iter)


def py_geodesic3d(image, softmask, spacing, v, lamb, iter=4):
    lg, le = wrapper_weights(lamb)
    return geodesic3d(image, softmask, spacing, v, lg, le, iter)


def py_signed_geodesic2d(image, softmask, v, lamb, iter=2):
    lg, le = wrapper_weights(lamb)
    return signed_geodesic2d(image, softmask, v, lg, le, iter)


def py_signed_geodesic3d(image, softmask, spacing, v, lamb, iter=4):
    lg, le = wrapper_weights(lamb)
    return signed_geodesic3d(image, softmask, spacing, v, lg, le, iter)


def py_gsf2d(image, softmask, theta, v, lamb, iter):
    return gsf2d(image, softmask, theta, v, lamb, iter)


def py_gsf3d(image, softmask, theta, spacing, v, lamb, iter):
    return gsf3d(image, softmask, theta, spacing, v, lamb, iter)


# ---- the reference's own CPU extension, when built ------------------------

_ref_mod = None


def reference_module():
    """Return the reference ``FastGeodisCpp`` module built by ``oracle/build.py``
    from the reference sources, or None if it has not been built here."""
    global _ref_mod
    if _ref_mod is not None:
        return _ref_mod
    if not os.path.isdir(_REF_DIR):
        return None
    libs = [f for f in os.listdir(_REF_DIR) if f.startswith("FastGeodisCpp") and f.endswith(".so")]
    if not libs:
        return None
    import importlib.util

    import torch  # noqa: F401  (loads libtorch for the extension)

    # loaded by path, never through sys.modules, so it cannot be confused
    # with this repository's own FastGeodisCpp module
    spec = importlib.util.spec_from_file_location("FastGeodisCpp", os.path.join(_REF_DIR, libs[0]))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    _ref_mod = mod
    return _ref_mod
