"""ParticleWriter frames and state files (CPU only): byte layout of a frame
(int32 N, then N x (x, y, z, c) float32, ParticleWriter.cpp:54-79),
round trips, the empty frame, and resume from a state file."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def pio(pkg):
    from importlib import import_module
    import __graft_entry__ as G
    return import_module(G.PKG_NAME + ".particle_io")


def test_frame_byte_layout(pio, tmp_path):
    p = np.array([[1.0, 2.0, 3.0], [4.5, -5.0, 6.25]], np.float32)
    c = np.array([0.5, 7.0], np.float32)
    f = tmp_path / "frame.txt"
    pio.write_frame(str(f), p, c)
    raw = f.read_bytes()
    assert len(raw) == 4 + 2 * 16
    assert np.frombuffer(raw[:4], "<i4")[0] == 2
    body = np.frombuffer(raw[4:], "<f4").reshape(2, 4)
    np.testing.assert_array_equal(body[:, :3], p)
    np.testing.assert_array_equal(body[:, 3], c)


def test_frame_round_trip_and_empty(pio, scene, tmp_path):
    x = scene.jittered_lattice((7, 5, 3), (0.1, 0.2, 0.3), 0.005, 0.1, seed=3)
    col = scene.uniform01(4, len(x)).astype(np.float32)
    f = tmp_path / "a.txt"
    pio.write_frame(str(f), x, col)
    p2, c2 = pio.read_frame(str(f))
    np.testing.assert_array_equal(p2, x)
    np.testing.assert_array_equal(c2, col)
    e = tmp_path / "e.txt"
    pio.write_frame(str(e), np.zeros((0, 3), np.float32))
    assert e.read_bytes() == b"\x00\x00\x00\x00"
    p3, c3 = pio.read_frame(str(e))
    assert p3.shape == (0, 3) and c3.shape == (0,)
    with pytest.raises(ValueError):
        pio.write_frame(str(f), x, col[:-1])
    (tmp_path / "bad.txt").write_bytes(np.array([3], "<i4").tobytes() + b"\x00" * 20)
    with pytest.raises(ValueError):
        pio.read_frame(str(tmp_path / "bad.txt"))


def test_writer_module_numbers_frames(pio, tmp_path):
    w = pio.ParticleWriter()
    w.setOutputPath(str(tmp_path) + "/")
    w.setNamePrefix("fluid_")
    x = np.arange(12, dtype=np.float32).reshape(4, 3)
    w.m_position.setValue(x)
    w.m_color_mapping.setValue(np.ones(4, np.float32))
    w.execute()
    w.execute()
    assert (tmp_path / "fluid_0.txt").exists() and (tmp_path / "fluid_1.txt").exists()
    np.testing.assert_array_equal(pio.read_frame(str(tmp_path / "fluid_1.txt"))[0], x)


def test_state_resume_matches_uninterrupted_run(pio, oracle, scene, tmp_path):
    """Steps 0..3 then save, load, steps 3..6 equal six steps in one go (the
    step is a function of (x, v) and the parameters)."""
    x = scene.jittered_lattice((6, 6, 6), (0.3, 0.5, 0.3), 0.005, 0.1, seed=8)
    v = np.zeros_like(x)
    P = oracle.make_params()
    x6, v6, _, _ = oracle.step(P, x, v, 6)
    x3, v3, _, _ = oracle.step(P, x, v, 3)
    f = str(tmp_path / "state.npz")
    params = {"smoothing_length": 0.0125, "iterations": 5}
    pio.save_state(f, x3, v3, params, step=3, scene_offsets=np.array([0, len(x)]))
    s = pio.load_state(f)
    assert s["step"] == 3 and s["params"] == params
    np.testing.assert_array_equal(s["scene_offsets"], [0, len(x)])
    xr, vr, _, _ = oracle.step(P, s["position"], s["velocity"], 3)
    np.testing.assert_array_equal(xr, x6)
    np.testing.assert_array_equal(vr, v6)
