"""Loader of the native module libpyCudaPolygon (the reference's loader contract,
pyCudaPolygon/pyCudaPolygonLink.py:8-12: the module's directory goes on sys.path and
`import libpyCudaPolygon` must succeed).  There is no fallback: a missing build raises."""
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.append(_HERE)

try:
    import libpyCudaPolygon  # noqa: E402
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError("libpyCudaPolygon is not built: run `python -m pycudapolygon_amd.build` "
                      f"({e})") from e
