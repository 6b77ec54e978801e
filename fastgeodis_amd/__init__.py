"""fastgeodis_amd -- raster-scan generalised geodesic distance transforms on
AMD Instinct MI355X (gfx950).

Drop-in for the raster-scan entry points of FastGeodis 1.0.5
(``FastGeodis/__init__.py:37-162, 526-585``): same names, argument names,
order, defaults and exception types::

    import fastgeodis_amd as FastGeodis
    d = FastGeodis.generalised_geodesic3d(image, softmask, spacing, v, lamb, iter)

Computation happens in hand-written HIP kernels (``csrc/fg_kernels.hip``)
behind a C ABI (``include/fastgeodis_hip.h``), reached through the torch
extension ``_fastgeodis_hip`` (``csrc/fg_glue.cpp``), exposed as
``fastgeodis_amd.FastGeodisCpp`` like the reference's binding module.  There is no CPU
implementation: HIP tensors are processed on their device and the current
torch stream; CPU tensors are staged through the current HIP device.  If the
native modules are not built, importing this package raises ImportError.

Additions beyond the reference: :func:`geodesic_batch` (many uneven 2D/3D
instances in the same persistent launches), :func:`set_tuning`, and the
interactive-segmentation serving loop in :mod:`fastgeodis_amd.serving`.
"""
from __future__ import annotations

import importlib.util
import os
import sys
from typing import List, Optional, Sequence

import torch

__version__ = "0.1.0"

_LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")


def _load_native():
    import sysconfig

    path = os.path.join(_LIB_DIR, "_fastgeodis_hip" + sysconfig.get_config_var("EXT_SUFFIX"))
    if not os.path.isfile(path):
        raise ImportError(
            f"fastgeodis_amd native module not built ({path} missing); run "
            "`python fastgeodis_amd/_build.py` (hipcc for gfx950)")
    spec = importlib.util.spec_from_file_location("_fastgeodis_hip", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


# The binding module; the reference's wrappers call theirs FastGeodisCpp.
FastGeodisCpp = _load_native()


def _release_at_exit():
    # frees pinned host memory and events before the HIP runtime tears down
    try:
        FastGeodisCpp.release()
    except Exception:
        pass


import atexit  # noqa: E402

atexit.register(_release_at_exit)


def generalised_geodesic2d(
    image: torch.Tensor, softmask: torch.Tensor, v: float, lamb: float, iter: int = 2
):
    r"""Generalised geodesic distance, 2D (reference ``FastGeodis/__init__.py:37-65``).

    Args:
        image: (1, C, H, W) float32 image, any number of channels.
        softmask: (1, Cm, H, W) float32 soft mask in [0, 1]; 0 marks seeds.
        v: weight between the unary (mask) and spatial distances.
        lamb: 0.0 gives Euclidean distance, 1.0 geodesic distance.
        iter: number of iterations of the raster passes.

    Returns:
        (1, Cm, H, W) float32 distances (channel 0 transformed).
    """
    return FastGeodisCpp.generalised_geodesic2d(image, softmask, v, lamb, 1 - lamb, iter)


def generalised_geodesic3d(
    image: torch.Tensor,
    softmask: torch.Tensor,
    spacing: List,
    v: float,
    lamb: float,
    iter: int = 4,
):
    r"""Generalised geodesic distance, 3D (reference ``FastGeodis/__init__.py:68-98``).

    ``spacing`` is (s_D, s_H, s_W).  Other arguments as in
    :func:`generalised_geodesic2d`; tensors are (1, C, D, H, W).
    """
    return FastGeodisCpp.generalised_geodesic3d(image, softmask, spacing, v, lamb, 1 - lamb, iter)


def signed_generalised_geodesic2d(
    image: torch.Tensor, softmask: torch.Tensor, v: float, lamb: float, iter: int = 2
):
    r"""Signed generalised geodesic distance D(M) - D(1 - M), 2D
    (reference ``FastGeodis/__init__.py:101-129``)."""
    return FastGeodisCpp.signed_generalised_geodesic2d(image, softmask, v, lamb, 1 - lamb, iter)


def signed_generalised_geodesic3d(
    image: torch.Tensor,
    softmask: torch.Tensor,
    spacing: List,
    v: float,
    lamb: float,
    iter: int = 4,
):
    r"""Signed generalised geodesic distance, 3D (reference ``FastGeodis/__init__.py:132-162``)."""
    return FastGeodisCpp.signed_generalised_geodesic3d(image, softmask, spacing, v, lamb, 1 - lamb, iter)


def GSF2d(
    image: torch.Tensor,
    softmask: torch.Tensor,
    theta: float,
    v: float,
    lamb: float,
    iter: int,
):
    r"""Geodesic symmetric filtering, 2D (reference ``FastGeodis/__init__.py:526-553``)."""
    return FastGeodisCpp.GSF2d(image, softmask, theta, v, lamb, iter)


def GSF3d(
    image: torch.Tensor,
    softmask: torch.Tensor,
    theta: float,
    spacing: List,
    v: float,
    lamb: float,
    iter: int,
):
    r"""Geodesic symmetric filtering, 3D (reference ``FastGeodis/__init__.py:556-585``)."""
    return FastGeodisCpp.GSF3d(image, softmask, theta, spacing, v, lamb, iter)


# ---- additions -------------------------------------------------------------

OP_UNSIGNED, OP_SIGNED, OP_GSF = 0, 1, 2


class Instance:
    """One member of a batched call.

    op: "unsigned" | "signed" | "gsf".  For "unsigned"/"signed" the weights
    follow the Python wrappers (l_grad = lamb, l_eucl = 1 - lamb in double);
    for "gsf" lamb is passed raw and 1 - lamb is formed in float, as the
    reference's C++ GSF does.

    image_t: optional transposed copy of ``image`` (see :func:`transposed_image`)
    that the call uses instead of transposing the image itself (a caller that
    transforms one image many times keeps it).
    """

    __slots__ = ("image", "softmask", "spacing", "v", "lamb", "iter", "op", "theta", "image_t")

    def __init__(self, image, softmask, v, lamb, iter=None, spacing=None, op="unsigned", theta=0.0,
                 image_t=None):
        """See the class docstring."""
        self.image, self.softmask, self.v, self.lamb = image, softmask, v, lamb
        self.image_t = image_t
        nd = image.dim() - 2
        self.iter = iter if iter is not None else (2 if nd == 2 else 4)
        self.spacing = list(spacing) if spacing is not None else [1.0, 1.0, 1.0]
        if op not in ("unsigned", "signed", "gsf"):
            raise ValueError(f"unknown op {op!r}")
        self.op, self.theta = op, theta


def geodesic_batch(instances: Sequence[Instance], tuned: bool = False) -> List[torch.Tensor]:
    """Runs many 2D/3D instances of uneven sizes in the same persistent launches.

    All tensors must be on one HIP device.  Each returned tensor equals the
    single-instance call with the same arguments, bit for bit.  ``tuned``:
    use the tuning that :mod:`fastgeodis_amd.autotune` measured for the
    batch's size class (measuring it on first use) instead of the analytic
    defaults; results are identical either way.
    """
    if tuned and instances:
        from . import autotune

        best = autotune.tuning_for_batch(instances, device=instances[0].image.device)
        old = get_tuning()
        set_tuning(best["engine"], best["halo"], best["max_groups"], best["shape"])
        try:
            return geodesic_batch(instances)
        finally:
            set_tuning(**old)
    ops = {"unsigned": OP_UNSIGNED, "signed": OP_SIGNED, "gsf": OP_GSF}
    none = torch.empty(0)
    images_t = [i.image_t if i.image_t is not None else none for i in instances]
    return FastGeodisCpp.geodesic_batch(
        [i.image for i in instances],
        [i.softmask for i in instances],
        [[float(s) for s in i.spacing] if i.image.dim() == 5 else [] for i in instances],
        [ops[i.op] for i in instances],
        [float(i.v) for i in instances],
        [float(i.lamb) for i in instances],
        [(1 - float(i.lamb)) if i.op != "gsf" else 0.0 for i in instances],
        [float(i.theta) for i in instances],
        [int(i.iter) for i in instances],
        images_t if any(i.image_t is not None for i in instances) else [],
    )


def transposed_image(image: torch.Tensor) -> torch.Tensor:
    """The image in the layout of the W (3D) or column (2D) pass:
    (1, C, W, H, D) for a (1, C, D, H, W) image, (1, C, W, H) for (1, C, H, W)."""
    if image.dim() == 5:
        return image.permute(0, 1, 4, 3, 2).contiguous()
    if image.dim() == 4:
        return image.permute(0, 1, 3, 2).contiguous()
    raise ValueError("image must be (1, C, H, W) or (1, C, D, H, W)")


def set_tuning(engine: int = 1, halo: int = 0, max_groups: int = 0, shape: int = 0) -> None:
    """Engine (0 per-slice, 1 persistent tiled), halo depth K (slices between
    neighbour exchanges; 0 = automatic: the deepest halo whose tiles fit one
    per compute unit), workgroup cap (0 = occupancy limit) and 3D plane
    workgroup shape (0 auto = 5, 1 = 4 waves x 8 rows, 2 = 8 x 4, 3 = 16 x 2,
    4 = 8 x 4 with a deeper prefetch ring, 5 = 16 x 2 with the deeper ring,
    6 = 8 x 4 with the deeper ring and four more waves that perform the halo
    exchanges, 7 = 8 waves x 2 rows, a 16-row window, with the deeper ring)
    used by later calls on this host thread."""
    FastGeodisCpp.set_tuning(int(engine), int(halo), int(max_groups), int(shape))


def get_tuning():
    e, h, g, s = FastGeodisCpp.get_tuning()
    return {"engine": e, "halo": h, "max_groups": g, "shape": s}


def set_graphs(on: bool = True) -> None:
    """Graph replay on this host thread (default on): a call whose sequence of
    kernel launches repeats one seen before runs as one launch of a cached HIP
    graph updated with the call's arguments.  Results are identical either
    way; off issues every launch directly."""
    FastGeodisCpp.set_graphs(bool(on))


def check_status(device: Optional[int] = None) -> None:
    """Synchronises the device's current stream and raises RuntimeError if a
    launch on it reported an error (an expired bounded wait, whose call then
    returned invalid distances) since the last check.  Every call also checks,
    without synchronising, for errors left by earlier calls on its device."""
    if device is None:
        device = torch.cuda.current_device()
    FastGeodisCpp.check_status(int(device))


def native_library_paths() -> List[str]:
    """Paths of the native libraries this package loads."""
    import sysconfig

    return [os.path.join(_LIB_DIR, "libfastgeodis_hip.so"),
            os.path.join(_LIB_DIR, "_fastgeodis_hip" + sysconfig.get_config_var("EXT_SUFFIX"))]


__all__ = [
    "generalised_geodesic2d", "generalised_geodesic3d",
    "signed_generalised_geodesic2d", "signed_generalised_geodesic3d",
    "GSF2d", "GSF3d", "Instance", "geodesic_batch", "set_tuning", "get_tuning", "check_status", "set_graphs", "transposed_image",
]
