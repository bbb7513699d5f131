"""rwpt: MI355X-native random-walk particle tracking for MODPATH-RW's hot path.

Importable once ``modpath-rw_amd`` is on ``sys.path``.  The compute lives in
``librwpt.so`` (HIP, gfx950) behind the C ABI of ``include/rwpt.h``;
``rwpt.engine`` binds it, ``rwpt.problems`` synthesises structured-grid
inputs, ``rwpt.mf6`` reads MODFLOW 6 DIS output into them and
``rwpt.driver`` runs MPathRW's schedule of tracking intervals.
"""
from . import mf6, problems  # noqa: F401
from .engine import Engine, RwptError, library_path  # noqa: F401
