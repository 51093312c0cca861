"""The C++ facades (include/ge_facade.hpp, ge_facade3d.hpp) driven like the
reference's updateScene() and mode3D(): stage by stage with the reference's
call names, and fused through ge_step / ge3_step.  The programs compare both
states bit for bit."""
import subprocess

import pytest

from test_abi import _build_facade_program

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [5000, 60000])
def test_facade_staged_equals_fused(tmp_path, n):
    exe = _build_facade_program(tmp_path)
    r = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "staged==fused: yes" in r.stdout


@pytest.mark.parametrize("n", [8000])
def test_facade3d_staged_equals_fused(tmp_path, n):
    # mode3D with SPH3D: the 3D facade stage by stage against ge3_step
    exe = _build_facade_program(tmp_path, "facade_frame3d")
    r = subprocess.run([str(exe), str(n)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "staged==fused: yes" in r.stdout
