"""MI355X-native finite-volume SIMPLE hot path with XCALibre's interface.

The package is imported under the module name ``xcal_amd`` (the directory
name is not a Python identifier); see ``load_package`` in __graft_entry__.py.
"""
from . import _libs  # noqa: F401
from . import mesh  # noqa: F401
from . import _abi  # noqa: F401
