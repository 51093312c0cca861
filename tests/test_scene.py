"""Scene files (include/ge_scene.h) against an independent restatement of
SaveSystem's version-190 layout.

The reference ships no save file, so the byte layout is pinned by numpy
dtypes written here from the field order of SaveSystem's writes
(GE/src/UX/saveSystem.cpp:309-396) and reads (GE/include/UX/saveSystem.h:
228-342) and the field types of GE/include/Particles/particle.h: packed,
glm::vec2 / vec3 of floats, raylib Color as 4 bytes, bool as 1 byte."""
import numpy as np
import pytest

from galaxy_step.scene import Scene

SEP = b"\n---PARTICLE BINARY DATA---\n"


def record_dtype(dim):
    v = ("<f4", (dim,))
    phys = [("pos", v), ("predPos", v), ("vel", v), ("prevVel", v), ("predVel", v), ("acc", v),
            ("mass", "<f4"), ("press", "<f4"), ("pressTmp", "<f4"), ("pressF", v), ("dens", "<f4"),
            ("predDens", "<f4"), ("sphMass", "<f4"), ("restDens", "<f4"), ("stiff", "<f4"), ("visc", "<f4"),
            ("cohesion", "<f4"), ("temp", "<f4"), ("ke", "<f4"), ("prevKe", "<f4"), ("mortonKey", "<u8"),
            ("id", "<u4"), ("isHotPoint", "u1"), ("hasSolidified", "u1")]
    rend = [("color", "u1", (4,)), ("pColor", "u1", (4,)), ("sColor", "u1", (4,)), ("sphColor", "u1", (4,)),
            ("size", "<f4"), ("uniqueColor", "u1"), ("isSolid", "u1"), ("canBeSubdivided", "u1"),
            ("canBeResized", "u1"), ("isDarkMatter", "u1"), ("isSPH", "u1"), ("isSelected", "u1"),
            ("isGrabbed", "u1"), ("previousSize", "<f4"), ("neighbors", "<i4"), ("totalRadius", "<f4"),
            ("lifeSpan", "<f4"), ("sphLabel", "<u4"), ("isPinned", "u1"), ("isBeingDrawn", "u1"),
            ("spawnCorrectIter", "<i4"), ("turbulence", "<f4")]
    return np.dtype(phys + rend)  # numpy structured dtypes are packed unless align=True


def random_records(n, dim, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, record_dtype(dim))
    for name in r.dtype.names:
        f = r[name]
        if f.dtype.kind == "f":
            r[name] = rng.normal(0, 100, f.shape).astype(np.float32)
        elif name in ("isHotPoint", "hasSolidified", "uniqueColor", "isSolid", "canBeSubdivided", "canBeResized",
                      "isDarkMatter", "isSPH", "isSelected", "isGrabbed", "isPinned", "isBeingDrawn"):
            r[name] = rng.integers(0, 2, f.shape)
        else:
            r[name] = rng.integers(0, 200, f.shape)
    r["id"] = rng.permutation(np.arange(5, 5 + n)).astype(np.uint32)
    return r


YAML = ("GlobalTrails: false\n---\nTheta: 0.699999988\n---\nSoftening: 1.5\n---\nTimeMult: 2\n---\n"
        "GravityMultiplier: 1.49835181e+10\n---\nSPHEnabled: true\n---\nLoopingSpace: true\n---\n"
        "InfiniteDomain: false\n---\nSPHRadiusMult: 4\n---\nSPHMass: 0.0500000007\n---\nSPHDelta: 20000\n---\n"
        "DomainWidth: 3000\n---\nDomainHeight: 2000\n---\nDomainDepth3D: 900\n---\nis3DMode: {mode}\n")


def write_file(path, recs, is3d, tail=b"", version=190, ids=(1234, 7, 9)):
    head = np.array([version, *ids, len(recs)], np.uint32).tobytes()
    with open(path, "wb") as f:
        f.write(YAML.format(mode="true" if is3d else "false").encode())
        f.write(SEP)
        f.write(head)
        f.write(recs.tobytes())
        f.write(tail)


@pytest.mark.parametrize("is3d", [False, True])
def test_load_reads_the_v190_layout(tmp_path, is3d):
    dim = 3 if is3d else 2
    recs = random_records(257, dim, 1)
    tail = bytes(range(200)) * 3
    p = tmp_path / "Save_1.bin"
    write_file(p, recs, is3d, tail)
    with Scene.load(p) as s:
        info = s.info()
        assert info["is3d"] == int(is3d) and info["version"] == 190 and info["count"] == 257
        assert info["record_bytes"] == recs.dtype.itemsize == (208 if is3d else 180)
        assert (info["global_id"], info["global_shape_id"], info["global_wall_id"]) == (1234, 7, 9)
        assert info["tail_bytes"] == len(tail)
        assert s.records() == recs.tobytes()
        assert s.value("Theta") == "0.699999988" and s.value("Missing") is None
        # an unmodified scene saves to the same bytes
        q = tmp_path / "again.bin"
        s.save(q)
    assert q.read_bytes() == p.read_bytes()


def test_params_follow_the_reference_mapping(tmp_path):
    p = tmp_path / "s.bin"
    write_file(p, random_records(3, 2, 2), False)
    with Scene.load(p) as s:
        prm = s.params()
        # myVar.G = 6.674e-11 * gravityMultiplier (double product, stored as
        # float); timeFactor = 0.045f * timeStepMultiplier (globalLogic.cpp:1135, 1187)
        assert prm.G == np.float32(6.674e-11 * float(np.float32(1.49835181e10)))
        assert prm.dt == np.float32(np.float32(0.045) * np.float32(2.0)) and prm.dt_start == prm.dt
        assert prm.theta == np.float32(0.699999988) and prm.softening == 1.5
        assert prm.sph == 1 and prm.sph_h == 4.0 and prm.sph_unit_mass == np.float32(0.05)
        assert prm.delta == 20000.0 and prm.periodic == 1
        assert (prm.domain_w, prm.domain_h) == (3000.0, 2000.0)
        assert prm.viscosity == pytest.approx(0.3)  # absent key: the default
        p3 = s.params3()
        assert p3.domain_d == 900.0 and p3.periodic == 1
        # infiniteDomain switches the wrap off (isPeriodicBoundaryEnabled && !infiniteDomain)
        s.set_value("InfiniteDomain", "true")
        assert s.params().periodic == 0
        # set_params writes the inverse mapping
        prm.theta = 0.5
        prm.G = 1.0
        s.set_params(prm)
        back = s.params()
        assert back.theta == 0.5 and back.G == pytest.approx(1.0, rel=1e-6)


def test_yaml_forms_and_rejections(tmp_path):
    recs = random_records(4, 2, 3)
    head = np.array([190, 10, 0, 0, 4], np.uint32).tobytes()
    p = tmp_path / "forms.bin"
    p.write_bytes(b"# comment\nTheta: '0.5'\n...\n---\nis3DMode: no\nSoftening:   2.5  \n" + SEP + head +
                  recs.tobytes())
    with Scene.load(p) as s:
        assert s.params().theta == 0.5 and s.params().softening == 2.5 and s.info()["is3d"] == 0
        assert s.info()["tail_bytes"] == 0
        with pytest.raises(Exception):
            s.set_value("is3DMode", "true")  # the records are 2D
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"Theta: 1\n")
    with pytest.raises(Exception, match="separator"):
        Scene.load(bad)
    bad.write_bytes(b"is3DMode: false\n" + SEP + np.array([190, 1, 0, 0, 9], np.uint32).tobytes() + b"x" * 100)
    with pytest.raises(Exception, match="truncated"):
        Scene.load(bad)
    bad.write_bytes(b"is3DMode: false\n" + SEP + np.array([170, 1, 0, 0, 0], np.uint32).tobytes())
    with pytest.raises(Exception, match="version"):
        Scene.load(bad)


def test_new_scene_has_an_empty_tail(tmp_path):
    with Scene.new(is3d=True) as s:
        info = s.info()
        assert info["count"] == 0 and info["is3d"] == 1 and info["tail_bytes"] == 24
        s.save(tmp_path / "n.bin")
    raw = (tmp_path / "n.bin").read_bytes()
    assert raw.startswith(b"is3DMode: true\n")
    assert raw.endswith(SEP + np.array([190, 0, 0, 0, 0], np.uint32).tobytes() + bytes(24))
