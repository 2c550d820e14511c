"""Batched scenes (pbf_batch_step): independent scenes with shared parameters,
each in its own copy of the scene box, advanced together.

Bars: every scene's x, v and rho after a batched step are bit-identical to a
pbf_step of that scene alone (the batch only changes where a scene's cells
live, never a particle's neighbor set or summation order); scenes checked
against the oracle meet the single-scene parity bars.

Every call goes through the C ABI."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import assert_step_parity, dev, oracle_both, say  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def mixed_scenes(scene):
    """Scenes of ragged sizes: empty, single particle, wave-boundary counts,
    a dense blob, a block leaving the grid through its lower face, and two
    identical copies of one block."""
    out = []
    for k, n in enumerate([1, 0, 63, 64, 65, 1000, 4097]):
        side = max(1, int(np.ceil(n ** (1 / 3) - 1e-9)))
        while side ** 3 < n:
            side += 1
        out.append(scene.jittered_lattice((side, side, side), (0.3, 0.5, 0.3), 0.005, 0.1, seed=40 + k, count=n)
                   if n else np.zeros((0, 3), np.float32))
    out.append(scene.gaussian_blobs(1, 6000, 0.02, seed=47))
    oog = scene.jittered_lattice((12, 12, 12), (-0.035, 0.2, 0.4), 0.005, 0.1, seed=48)
    out.append(oog)  # first layers lie below the grid origin -0.025
    twin = scene.jittered_lattice((16, 10, 14), (0.6, 0.3, 0.2), 0.005, 0.1, seed=49)
    out += [twin, twin.copy()]
    return out


def concat(scenes):
    counts = [len(s) for s in scenes]
    off = np.zeros(len(scenes) + 1, np.int32)
    off[1:] = np.cumsum(counts)
    x = np.concatenate([np.asarray(s, np.float32).reshape(-1, 3) for s in scenes])
    return x, off


def velocities(scene, n, seed, scale=0.2):
    return (scale * (scene.uniform01(seed, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)


def serial(fluid, scenes, vels, nsteps):
    ctx = fluid.Context("cuda:0")
    res = []
    for x, v in zip(scenes, vels):
        xd, vd = dev(x), dev(v)
        rho = torch.zeros(len(x), dtype=torch.float32, device="cuda:0")
        if len(x):
            ctx.step(xd, vd, rho, nsteps)
        res.append((xd.cpu().numpy(), vd.cpu().numpy(), rho.cpu().numpy()))
    return res


def batched(fluid, scenes, vels, nsteps, ctx=None):
    ctx = ctx or fluid.Context("cuda:0")
    x, off = concat(scenes)
    v = np.concatenate([np.asarray(q, np.float32).reshape(-1, 3) for q in vels])
    xd, vd = dev(x), dev(v)
    rho = torch.zeros(len(x), dtype=torch.float32, device="cuda:0")
    offd = torch.from_numpy(off).to("cuda:0")
    ctx.batch_step(xd, vd, offd, rho, nsteps)
    xs, vs, rs = xd.cpu().numpy(), vd.cpu().numpy(), rho.cpu().numpy()
    return [(xs[a:b], vs[a:b], rs[a:b]) for a, b in zip(off[:-1], off[1:])], ctx


def assert_bit_identical(got, want):
    for s, (g, w) in enumerate(zip(got, want)):
        for name, a, b in zip("x v rho".split(), g, w):
            assert a.shape == b.shape, (s, name)
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"scene {s} {name}")


def test_batch_equals_serial_mixed(fluid, scene):
    say("batched vs serial, 12 ragged scenes incl. empty / out-of-grid / twins, 3 steps")
    scenes = mixed_scenes(scene)
    vels = [velocities(scene, len(s), 60 + k) for k, s in enumerate(scenes)]
    vels[8][:] = np.array([-2.0, 0.0, 0.0], np.float32)  # the out-of-grid block keeps leaving
    vels[10] = vels[9].copy()  # twins: same state
    got, ctx = batched(fluid, scenes, vels, 3)
    assert_bit_identical(got, serial(fluid, scenes, vels, 3))
    st = ctx.stats()
    assert st["out_of_grid"] > 0
    # twins: same input, same output
    for a, b in zip(got[9], got[10]):
        np.testing.assert_array_equal(a, b)


def test_batch_equals_serial_generator(fluid, scene):
    say("batched vs serial, 48 log-uniform lattice scenes (256..8192), 2 steps")
    b = scene.batched_lattice_scenes(nscenes=48, nmin=256, nmax=8192, seed=4)
    scenes = [b.pos[b.offsets[s]:b.offsets[s + 1]] for s in range(48)]
    vels = [np.zeros_like(s) for s in scenes]
    got, ctx = batched(fluid, scenes, vels, 2)
    assert_bit_identical(got, serial(fluid, scenes, vels, 2))
    st = ctx.stats()
    # the scenes' windows hold far fewer cells than 48 full grids
    assert 48 < st["cells_used"] < 48 * 2 * 85 ** 3 // 20  # sort cells: half a grid cell along x


def test_batch_staged_equals_serial(fluid, scene):
    """Staged lists in a batch (waves may straddle scenes, and the serial runs
    of small scenes keep unstaged lists) give the serial bits."""
    say("batched with staged lists vs serial, 12 ragged scenes, 3 steps")
    scenes = mixed_scenes(scene)
    vels = [velocities(scene, len(s), 80 + k) for k, s in enumerate(scenes)]
    ctx = fluid.Context("cuda:0", stage_min_particles=1, list_slots=256)
    got, ctx = batched(fluid, scenes, vels, 3, ctx)
    assert ctx.stats()["staged_waves"] > 0
    assert_bit_identical(got, serial(fluid, scenes, vels, 3))


def test_batch_more_than_1024_scenes(fluid, scene):
    say("batched vs serial, 1500 scenes of 8..120 particles, 1 step")
    u = scene.uniform01(71, 1500)
    counts = (8 + np.floor(u * 113)).astype(int)
    scenes = [scene.jittered_lattice((5, 5, 5), (0.4, 0.4, 0.4), 0.005, 0.1, seed=1000 + s, count=int(n))
              for s, n in enumerate(counts)]
    vels = [np.zeros_like(s) for s in scenes]
    got, _ = batched(fluid, scenes, vels, 1)
    assert_bit_identical(got, serial(fluid, scenes, vels, 1))


def test_batch_scenes_vs_oracle(fluid, oracle, scene):
    say("batched step vs oracle on three scenes of a batch, 1 step")
    scenes = mixed_scenes(scene)
    vels = [velocities(scene, len(s), 80 + k) for k, s in enumerate(scenes)]
    got, _ = batched(fluid, scenes, vels, 1)
    for s in (5, 7, 8):
        assert_step_parity(got[s], *oracle_both(oracle, scenes[s], vels[s], 1))


def test_batch_reuses_context(fluid, scene):
    """A context serves batches of changing scene counts and a plain step in
    between without stale state."""
    say("one context: 12-scene batch, single step, 3-scene batch")
    scenes = mixed_scenes(scene)
    vels = [np.zeros_like(s) for s in scenes]
    ctx = fluid.Context("cuda:0")
    got, _ = batched(fluid, scenes, vels, 1, ctx)
    x, v = dev(scenes[5]), dev(vels[5])
    ctx.step(x, v, None, 1)
    got3, _ = batched(fluid, scenes[4:7], vels[4:7], 1, ctx)
    want = serial(fluid, scenes, vels, 1)
    assert_bit_identical(got, want)
    assert_bit_identical(got3, want[4:7])
    np.testing.assert_array_equal(x.cpu().numpy(), want[5][0])


def test_batch_invalid_arguments(fluid, capi, scene):
    say("batched step argument checks")
    ctx = fluid.Context("cuda:0")
    x = dev(scene.jittered_lattice((4, 4, 4), (0.3, 0.3, 0.3), 0.005, 0.1, seed=3))
    v = torch.zeros_like(x)
    off = torch.tensor([0, 64], dtype=torch.int32, device="cuda:0")
    L = capi.lib()
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    assert L.pbf_batch_step(ctx.handle, p(x), p(v), None, p(off), 0, 64, 1, None) == capi.PBF_ERR_INVALID
    assert L.pbf_batch_step(ctx.handle, p(x), p(v), None, None, 1, 64, 1, None) == capi.PBF_ERR_INVALID
    assert L.pbf_batch_step(ctx.handle, p(x), p(v), None, p(off), 1, -1, 1, None) == capi.PBF_ERR_INVALID
    assert L.pbf_batch_step(ctx.handle, p(x), p(v), None, p(off), 1, 64, 1, None) == capi.PBF_OK
    with pytest.raises(ValueError):
        ctx.batch_step(x, v, off.to(torch.int64))
