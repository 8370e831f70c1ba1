"""Re-export of the module's *Enum types (reference pyCudaPolygon/enums.py:10-14)."""
from .pyCudaPolygonLink import libpyCudaPolygon as _lpcp

for _name in dir(_lpcp):
    if _name.endswith("Enum"):
        globals()[_name] = getattr(_lpcp, _name)
