"""MI355X-native finite-volume SIMPLE hot path with XCALibre's interface.

The package directory name is not a Python identifier; it is imported under
the module name ``xcal_amd`` (see ``load_package`` in __graft_entry__.py).

Submodules: ``mesh`` (host mesh builders), ``hip`` (ctypes bindings of the
gfx950 library libxcalhip.so), ``api`` (the reference-shaped user API),
``writers`` (VTK result output), ``build`` (compiles the native libraries
in-tree).
"""
from . import _abi  # noqa: F401
from . import _libs  # noqa: F401
from . import mesh  # noqa: F401
from . import hip  # noqa: F401
from . import writers  # noqa: F401
from . import api  # noqa: F401
