"""MI355X (gfx950) implementation of pyCudaPolygon's overlap-force-minimization path.

    from pycudapolygon_amd import model        # the reference's Python class
    from pycudapolygon_amd import capi         # ctypes C ABI (include/pcp.h)
"""


def __getattr__(name):
    if name == "model":
        from .pyCudaPolygon import model
        return model
    raise AttributeError(name)
