"""GPU tests of the JSON scene path (LoadScene -> GenericScene -> C ABI)."""
from pathlib import Path

import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu
SCENES = Path(__file__).resolve().parent / "golden" / "scenes"


def test_scene_seeding_and_step_match_the_direct_api():
    sc = V.Scene(str(SCENES / "dam_break_walls.json"))
    assert sc.n == 16 ** 3
    x0, v0, rho0, pid = sc.get_state(ids=True)
    want = CF.lattice((16, 16, 16), 0.02, (0.02, 0.02, 0.08), 0.001, 0x5EED0000)
    np.testing.assert_array_equal(x0, want)
    assert not v0.any() and not rho0.any()
    sc.step()  # iterations = 2 substeps of dt = 0.0004 / 2
    grid = V.volume_map(V.VFS_SHAPE_BOX_WALLS, (0.8, 0.4, 0.4, 0.8, 0.4, 0.4), (32, 32, 32), 0.04)
    p = V.make_params(sc.n, 0.04, dt=0.0004, iterations=2, box_size=(1.6, 0.8, 0.8), flags=V.VFS_FLAG_TRACK_IDS)
    sim = V.Sim(p, [grid])
    sim.set_state(want)
    sim.step()
    for a, b in zip(sc.get_state(ids=True), sim.get_state(ids=True)):
        np.testing.assert_array_equal(a, b)
    orc = H.OracleSim(p, [grid])
    orc.set_state(want)
    orc.substeps(2)
    np.testing.assert_array_equal(sc.get_state()[0].astype(np.float64), orc.get_state()[0])
    sc.reset()
    np.testing.assert_array_equal(sc.get_state()[0], want)


def test_compact_lattice_seeding_like_the_reference():
    """SpawnCompactParticlesInBox: spacing cbrt(1/rho0), x outer, z inner."""
    sc = V.Scene(str(SCENES / "compact_block.json"))
    assert sc.n == 10 * 12 * 14
    x0 = sc.get_state()[0]
    # the lattice as GenericScene::Reset sizes it (generic_scene.cpp:42-57): float
    # spacing from the C library's cbrtf, one multiply and one add per coordinate
    want = CF.compact_block((10, 12, 14), 1000.0, pos=(0.0, 0.0, 1.5))
    np.testing.assert_array_equal(x0, want)
    sc.step()
    x1 = sc.get_state()[0]
    assert np.isfinite(x1).all() and not np.array_equal(x0, x1)


def test_cpp_facade_matches_the_c_abi(tmp_path):
    """tests/cpp/facade_scene (Simulation / LoadScene / GenericScene in C++)
    gives the same state as the ctypes scene path, bit for bit."""
    import subprocess
    exe = Path(__file__).resolve().parent / "cpp" / "facade_scene"
    out = tmp_path / "state.bin"
    r = subprocess.run([str(exe), str(SCENES / "dam_break_walls.json"), "3", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:4], np.uint32)[0])
    f = np.frombuffer(raw[4:], np.float32)
    x, v, rho = f[:3 * n].reshape(n, 3), f[3 * n:6 * n].reshape(n, 3), f[6 * n:7 * n]
    ids = np.frombuffer(raw[4 + 28 * n:], np.uint32)
    sc = V.Scene(str(SCENES / "dam_break_walls.json"))
    for _ in range(3):
        sc.step()
    ex, ev, er, eid = sc.get_state(ids=True)
    np.testing.assert_array_equal(x, ex)
    np.testing.assert_array_equal(v, ev)
    np.testing.assert_array_equal(rho, er)
    np.testing.assert_array_equal(ids, eid)


def test_two_fluid_blocks_are_seeded_one_after_the_other(tmp_path):
    """GenericScene::Reset (generic_scene.cpp:42-57): each block's compact lattice
    takes the next compMul(size) particles, in the order of the JSON."""
    import json

    blocks = [{"size": [6, 5, 4], "pos": [0.5, 0.0, 0.25]}, {"size": [3, 7, 2], "pos": [4.0, 1.0, 2.0]}]
    scene = {
        "name": "Two blocks", "fluidModel": "wcsph",
        "wcsphParameters": {"stiffness": 1000.0, "expoent": 7.0, "viscosityStrenght": 0.01,
                            "outerWallStiffness": 10000.0},
        "simulationParameters": {"gravity": [0.0, -9.81, 0.0], "smoothRadius": 0.2, "timeScale": 1.0,
                                 "iterations": 2, "dt": 0.004, "targetDensity": 800.0,
                                 "boundingBox": {"pos": [0, 0, 0], "size": [6, 4, 4]}},
        "fluidBlocks": blocks,
        "boundaryObjects": [],
    }
    path = tmp_path / "two_blocks.json"
    path.write_text(json.dumps(scene))
    sc = V.Scene(str(path))
    assert sc.n == 6 * 5 * 4 + 3 * 7 * 2
    want = np.concatenate([CF.compact_block(tuple(b["size"]), 800.0, pos=tuple(b["pos"])) for b in blocks])
    x0, v0, _, _ = sc.get_state()
    np.testing.assert_array_equal(x0, want)
    orc = H.OracleSim(sc.params, [])
    orc.set_state(x0, v0)
    sc.step()
    orc.substeps(2)
    np.testing.assert_array_equal(sc.get_state()[0].astype(np.float64), orc.get_state()[0])
    sc.reset()
    np.testing.assert_array_equal(sc.get_state()[0], want)
    sc.close()
