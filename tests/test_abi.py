"""CPU checks of the C-ABI boundary: the shared library loads, exports every
entry point include/ge_physics.h declares, its struct layouts match the
ctypes mirror, and the defaults equal the reference's parameters."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ge_physics.h")


SERVE_HEADER = os.path.join(ROOT, "include", "ge_serve.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ge_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import galaxy_step
    lib = galaxy_step.load_library()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(galaxy_step.lib.EXPORTS)


SCENE_HEADER = os.path.join(ROOT, "include", "ge_scene.h")


def test_library_exports_scene_header():
    from galaxy_step import scene
    lib = scene.load_scene_library()
    names = declared_functions(SCENE_HEADER)
    assert set(names) == set(scene.SCENE_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_library_exports_serve_header():
    from galaxy_step import serve
    lib = serve._lib()
    names = declared_functions(SERVE_HEADER)
    assert set(names) == set(serve.SERVE_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_serve_struct_layout_matches_header(tmp_path):
    from galaxy_step import serve as S
    src = tmp_path / "sz2.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ge_serve.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ge_tune_result), '
                   'sizeof(ge_server_config), sizeof(ge_request), sizeof(ge_request_result), '
                   'sizeof(ge_server_stats), offsetof(ge_request, out), offsetof(ge_server_stats, rounds));'
                   'return 0;}\n')
    exe = tmp_path / "sz2"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    want = [C.sizeof(S.TuneResult), C.sizeof(S.ServerConfig), C.sizeof(S.Request),
            C.sizeof(S.RequestResult), C.sizeof(S.ServerStats), S.Request.out.offset,
            S.ServerStats.rounds.offset]
    assert [int(v) for v in out] == want


def test_server_default_config():
    from galaxy_step import serve
    cfg = serve.default_server_config()
    assert cfg.hbm_budget_bytes == 200 << 30 and cfg.batching == 1 and cfg.max_systems == 512


def test_struct_layout_matches_header(tmp_path):
    import galaxy_step.lib as L
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ge_physics.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(ge_config), '
                   'sizeof(ge_particles), sizeof(ge_step_params), sizeof(ge_node), '
                   'offsetof(ge_step_params, compensated));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    want = [C.sizeof(L.GeConfig), C.sizeof(L.GeParticles), C.sizeof(L.StepParams),
            C.sizeof(L.GeNode), L.StepParams.compensated.offset]
    assert [int(v) for v in out] == want


def test_default_params_are_reference_defaults():
    import galaxy_step
    p = galaxy_step.default_params()
    # GE/include/parameters.h:114-133,311-318; SPH.h:17; neighborSearch.h:604
    assert p.theta == pytest.approx(0.8) and p.softening == pytest.approx(1.5)
    assert p.G == pytest.approx(6.674e-11)
    assert p.dt == pytest.approx(0.045) and p.direct_below == 5000
    assert p.sph_h == 3.0 and p.sph_cell == 3.0 and p.sph_unit_mass == pytest.approx(0.03)
    assert p.viscosity == pytest.approx(0.3) and p.delta == 19000.0
    assert p.sph_max_vel == 250.0 and p.periodic == 0 and p.temperature == 0


def test_create_reports_status_without_crashing():
    import galaxy_step
    from galaxy_step.lib import GeConfig
    lib = galaxy_step.load_library()
    cfg = GeConfig()
    lib.ge_default_config(C.byref(cfg))
    h = C.c_void_p()
    assert lib.ge_create(C.byref(h), None) == -1  # NULL config -> GE_E_INVALID
    cfg.capacity = 0
    assert lib.ge_create(C.byref(h), C.byref(cfg)) == -1
    cfg.capacity = 1000
    rc = lib.ge_create(C.byref(h), C.byref(cfg))
    assert rc in (0, -7)  # GE_OK on a GPU host, GE_E_NODEVICE otherwise
    if rc == 0:
        assert lib.ge_destroy(h) == 0
    assert lib.ge_version().decode().startswith("galaxy-step")


def test_config_generators_shapes():
    from galaxy_step import configs
    st, p = configs.config0()
    assert len(st["x"]) == 20000 and st["mass"][-1] == 5e4
    assert p["theta"] == 0.8 and p["dt"] == 0.01
    st, p = configs.config2(n_target=2000, n_impactor=500)
    assert len(st["x"]) == 2500 and p["sph"] == 1
    n, s = configs.config3_sizes()
    assert len(n) == 512 and n.min() >= 1000 and n.max() <= 131072
    reqs = configs.config4_stream(50, window_s=2.0)
    assert all(2000 <= r[1] <= 500000 and 10 <= r[2] <= 200 for r in reqs)


def _build_facade_program(tmp_path, name="facade_frame"):
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", name + ".cpp"),
                    "-L", os.path.join(ROOT, "galaxy-engine_amd"), "-lge_physics",
                    "-Wl,-rpath," + os.path.join(ROOT, "galaxy-engine_amd"), "-o", str(exe)], check=True)
    return exe


def test_cpp_facade_compiles_and_links(tmp_path):
    # the facade keeps the reference's call names over the C ABI; without a
    # device the program must fail cleanly with GE_E_NODEVICE, not crash
    exe = _build_facade_program(tmp_path)
    assert exe.exists()
    assert _build_facade_program(tmp_path, "facade_frame3d").exists()


HEADER3 = os.path.join(ROOT, "include", "ge_physics3d.h")


def test_library_exports_3d_header():
    from galaxy_step import lib3d
    lib = lib3d.load_library3()
    text = open(HEADER3).read()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ge3_[a-z_0-9]+)\s*\(", text, re.M)))
    assert len(names) >= 20
    assert set(names) == set(lib3d.EXPORTS3)
    for n in names:
        assert hasattr(lib, n), n


def test_3d_struct_layout_matches_header(tmp_path):
    from galaxy_step import lib3d as L
    src = tmp_path / "sz3.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ge_physics3d.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ge3_config), '
                   'sizeof(ge3_particles), sizeof(ge3_step_params), sizeof(ge3_node), '
                   'offsetof(ge3_step_params, compensated), offsetof(ge3_node, child), '
                   'offsetof(ge3_config, enable_sph), offsetof(ge3_particles, pfz), '
                   'offsetof(ge3_step_params, sph_h), offsetof(ge3_step_params, boundary_friction));return 0;}\n')
    exe = tmp_path / "sz3"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    want = [C.sizeof(L.Ge3Config), C.sizeof(L.Ge3Particles), C.sizeof(L.Step3Params), C.sizeof(L.Ge3Node),
            L.Step3Params.compensated.offset, L.Ge3Node.child.offset, L.Ge3Config.enable_sph.offset,
            L.Ge3Particles.pfz.offset, L.Step3Params.sph_h.offset, L.Step3Params.boundary_friction.offset]
    assert [int(v) for v in out] == want


def test_3d_default_params_are_reference_defaults():
    from galaxy_step import default_params3
    p = default_params3()
    # GE/include/parameters.h:99,114-117,133; globalLogic.cpp:1678
    assert p.G == pytest.approx(6.674e-11) and p.theta == pytest.approx(0.8)
    assert p.softening == pytest.approx(1.5) and p.dt == 1.0 and p.direct_below == 5000
    assert (p.domain_w, p.domain_h, p.domain_d) == (1300.0, 1300.0, 1300.0)
    assert p.periodic == 0 and p.sph == 0 and p.temperature == 0 and p.compensated == 1
    # SPH3D.h:13, neighborSearch.h:981, parameters.h:118, 311, 316
    assert p.sph_h == 3.0 and p.sph_cell == 3.0 and p.sph_unit_mass == pytest.approx(0.03)
    assert p.sph_max_vel == 250.0 and p.vertical_gravity == 3.0 and p.sph_ground == 0


def test_3d_create_reports_status_without_crashing():
    from galaxy_step import lib3d
    lib = lib3d.load_library3()
    cfg = lib3d.Ge3Config()
    lib.ge3_default_config(C.byref(cfg))
    h = C.c_void_p()
    assert lib.ge3_create(C.byref(h), None) == -1
    cfg.capacity = 0
    assert lib.ge3_create(C.byref(h), C.byref(cfg)) == -1
    cfg.capacity = 1000
    rc = lib.ge3_create(C.byref(h), C.byref(cfg))
    assert rc in (0, -7)
    if rc == 0:
        assert lib.ge3_destroy(h) == 0


RENDER_HEADER = os.path.join(ROOT, "include", "ge_render.h")


def test_library_exports_render_header():
    from galaxy_step import render
    lib = render.load_render_library()
    text = open(RENDER_HEADER).read()
    names = sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(ge3?_[a-z_0-9]+)\s*\(", text, re.M)))
    assert set(names) == set(render.RENDER_EXPORTS)
    for n in names:
        assert hasattr(lib, n), n


def test_render_struct_layout_matches_header(tmp_path):
    from galaxy_step import render as R
    from galaxy_step import lib as L
    from galaxy_step import lib3d as L3
    src = tmp_path / "szr.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ge_render.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", '
                   'sizeof(ge_render_fields), sizeof(ge_visual_params), sizeof(ge_ply_vertex), '
                   'sizeof(ge_galaxy_params), offsetof(ge_visual_params, p_color), '
                   'offsetof(ge_visual_params, max_color_acc), offsetof(ge_ply_vertex, radius), '
                   'offsetof(ge_galaxy_params, first_id), offsetof(ge_config, enable_render), '
                   'offsetof(ge3_config, enable_render), sizeof(ge_config) + sizeof(ge3_config));return 0;}\n')
    exe = tmp_path / "szr"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    want = [C.sizeof(R.RenderFields), C.sizeof(R.VisualParams), C.sizeof(R.PlyVertex), C.sizeof(R.GalaxyParams),
            R.VisualParams.p_color.offset, R.VisualParams.max_color_acc.offset, R.PlyVertex.radius.offset,
            R.GalaxyParams.first_id.offset, L.GeConfig.enable_render.offset, L3.Ge3Config.enable_render.offset,
            C.sizeof(L.GeConfig) + C.sizeof(L3.Ge3Config)]
    assert [int(v) for v in out] == want
    assert R.PLY_DTYPE.itemsize == C.sizeof(R.PlyVertex) == 20


def test_render_defaults_are_reference_defaults():
    from galaxy_step import render as R
    v = R.default_visual_params()
    # densitySize.h:7-12; particleColorVisuals.h:9, 24-25, 33-36; neighborSearch.h:441, 604
    assert (v.min_size, v.max_size, v.size_acc, v.size_max_neighbors) == (pytest.approx(0.17), pytest.approx(0.62),
                                                                         22.0, 200)
    assert v.density_color == 1 and v.density_size == 0 and v.force_size == 0 and v.search == 0
    assert v.p_color == R.rgba(0, 40, 68, 100) and v.s_color == R.rgba(155, 80, 40, 75)
    assert v.color_max_neighbors == 200 and v.max_color_acc == 40.0 and v.cell == 3.0 and v.radius == 3.5
    g = R.default_galaxy_params()
    # particlesSpawning.h:28-32, particlesSpawning.cpp:141, 165, 200, 222; parameters.h:297
    assert (g.visible, g.dark, g.outer_radius, g.scale_length) == (40000, 12000, 200.0, 90.0)
    assert (g.outer_radius_dm, g.radius_core_dm, g.mass_scatter) == (2000.0, 3.5, pytest.approx(0.75))
    assert g.mass_visible == np.float32(8.5e9) and g.mass_dark == np.float32(1.416e11)
