"""The Fortran binding module (include/rwpt.f90) compiled with amdflang and
linked against librwpt.so: the default options cross the boundary field
for field, and rwpt_kde_write called from Fortran writes the same bytes as
from Python.  No GPU is touched."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from rwpt import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/lib/llvm/bin/amdflang"

pytestmark = pytest.mark.skipif(not os.access(FC, os.X_OK), reason="amdflang not present")


def test_fortran_binding(tmp_path):
    lib = engine.library_path()
    exe = tmp_path / "binding_check"
    subprocess.run([FC, "-J", str(tmp_path), "-o", str(exe),
                    os.path.join(ROOT, "include", "rwpt.f90"),
                    os.path.join(ROOT, "tests", "fortran", "binding_check.f90"),
                    lib, f"-Wl,-rpath,{os.path.dirname(lib)}"],
                   check=True, cwd=tmp_path)
    out_f = tmp_path / "rows_f.out"
    res = subprocess.run([str(exe), str(out_f)], check=True, capture_output=True,
                         text=True, cwd=tmp_path).stdout.split("\n")
    o = engine.default_options()
    kinds = res[0].split()
    assert [int(kinds[1]), int(kinds[2])] == [o.advection_kind, o.time_step_kind]
    assert [float(kinds[3]), float(kinds[4])] == list(o.time_step_params)
    mask = [int(v) for v in res[1].split()[1:]]
    assert mask == [*o.dimension_mask, o.backward_tracking, o.stop_zone,
                    o.stop_at_weak_sinks, o.stop_at_weak_sources, o.steady_state,
                    o.precision]
    seed = [int(v, 16) for v in res[2].split()[1:]]
    assert seed == [o.base_seed[0], o.base_seed[1]]
    assert res[3].split() == ["rc", "0"]
    out_p = tmp_path / "rows_p.out"
    engine.kde_write(str(out_p), (0, 0, 0), (1.0, 1.0, 0.0), (3, 2, 1),
                     np.array([0.0, 1.5, -2.0, 3.25e-7, 4.0, 0.0]),
                     np.array([1.0, 2.0, 0.0, 4.0, 5.0, 6.0]), time_id=1, time=2.5,
                     group_id=1, column_format=1, histogram_scale=2.0)
    assert out_f.read_bytes() == out_p.read_bytes()
    assert len(out_f.read_text().splitlines()) == 3
