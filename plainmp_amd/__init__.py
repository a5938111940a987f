"""plainmp_amd: plainmp's FK + sphere-SDF collision path on MI355X (gfx950).

Module layout mirrors plainmp's Python API (``plainmp.kinematics``,
``plainmp.psdf``, ``plainmp.constraint``, ``plainmp.kdtree``) plus
``plainmp_amd.ompl`` (motion validators) and ``plainmp_amd.autograd``.
Every computation runs in the native library ``libplainmp_amd.so``; importing
this package without the built library raises.
"""

from __future__ import annotations

import os

# torch ships its own HIP runtime; loading it first makes the native library
# bind to that same runtime instead of a second copy.
try:  # pragma: no cover - torch is present in the supported environment
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(HERE, "libplainmp_amd.so")

try:
    from . import _native  # noqa: E402
except ImportError as e:  # pragma: no cover
    raise ImportError(
        "plainmp_amd native extension is not built; run `python plainmp_amd/build.py` "
        f"(or __graft_entry__.build()) first: {e}"
    ) from e

__all__ = ["LIBRARY_PATH", "_native"]
