"""Between the plain steps of one call the state stays in the cell order of
the step before (pbf_step and pbf_batch_step with more than one step): the
next sort reads it nearly in order and orders each cell by original index, so
the steps of one call must give the same bits as the same steps made one call
at a time, each of which reads and writes the caller's order.

A scene's out-of-grid bucket is in original-index order at any size (the
"oog_many" case holds thousands); in-grid cells of more than 1024 particles
keep arrival order, and no case here has one.

Every call goes through the C ABI."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import dev, say  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def cell_order_scene(scene, name):
    if name == "lattice":
        return scene.dam_break_32k()
    if name == "staged":  # above the staging threshold (262,144 particles)
        return scene.jittered_lattice((80, 64, 64), (0.1, 0.1, 0.1), 0.005, 0.1, seed=31)
    if name == "blob":
        return scene.gaussian_blobs(1, 20000, 0.02, seed=23)
    if name == "oog_many":  # about 9,000 of 11,520 particles below the grid's lower x face
        return scene.jittered_lattice((20, 24, 24), (-0.085, 0.45, 0.45), 0.005, 0.1, seed=13)
    # out-of-grid particles near the grid's lower x face, next to in-grid ones
    # (5 layers of 12 x 12 below the grid)
    return scene.jittered_lattice((20, 12, 12), (-0.047, 0.45, 0.45), 0.005, 0.1, seed=12)


def velocities(scene, n, seed, scale=0.5):
    return (scale * (scene.uniform01(seed, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)


def run(fluid, x, v, nsteps, calls, off=None):
    ctx = fluid.Context("cuda:0")
    xd, vd = dev(x), dev(v)
    rho = torch.zeros(len(x), dtype=torch.float32, device="cuda:0")
    offd = None if off is None else torch.from_numpy(off).to("cuda:0")
    for _ in range(calls):
        if offd is None:
            ctx.step(xd, vd, rho, nsteps)
        else:
            ctx.batch_step(xd, vd, offd, rho, nsteps)
    torch.cuda.synchronize()
    return (xd.cpu().numpy(), vd.cpu().numpy(), rho.cpu().numpy()), ctx.stats()


def assert_same_bits(a, b):
    for name, u, w in zip("x v rho".split(), a, b):
        np.testing.assert_array_equal(u.view(np.uint32), w.view(np.uint32), err_msg=name)


@pytest.mark.parametrize("scene_name", ["lattice", "staged", "blob", "oog", "oog_many"])
def test_steps_of_one_call_equal_single_steps(fluid, scene, scene_name):
    say(f"6 steps in one call vs 6 calls of one step, {scene_name}")
    x = cell_order_scene(scene, scene_name)
    v = velocities(scene, len(x), 17)
    one, st = run(fluid, x, v, 6, 1)
    many, _ = run(fluid, x, v, 1, 6)
    if scene_name == "oog":
        assert 0 < st["out_of_grid"] < 1024
    if scene_name == "oog_many":
        say(f"out of the grid: {st['out_of_grid']} of {len(x)}")
        assert st["out_of_grid"] > 4096
        again, _ = run(fluid, x, v, 6, 1)  # the same call again: the same bits
        assert_same_bits(one, again)
    if scene_name == "staged":
        assert st["staged_waves"] > 0
    assert_same_bits(one, many)
    # and both differ from the input (the steps ran)
    assert not np.array_equal(one[0], x)


def test_batch_steps_of_one_call_equal_single_steps(fluid, scene):
    say("batch: 5 steps in one call vs 5 calls of one step, 48 lattice scenes")
    b = scene.batched_lattice_scenes(nscenes=48, nmin=256, nmax=8192, seed=9)
    off = np.asarray(b.offsets, np.int32)
    x = np.asarray(b.pos, np.float32)
    v = velocities(scene, len(x), 21)
    one, _ = run(fluid, x, v, 5, 1, off)
    many, _ = run(fluid, x, v, 1, 5, off)
    assert_same_bits(one, many)


def test_cell_order_then_module_flags(fluid, scene):
    """A call with the box constraint keeps the caller's order between its
    steps; a plain call after it starts from the caller's arrays again."""
    say("plain, box, plain calls on one context vs single steps")
    x = scene.dam_break_32k()
    v = velocities(scene, len(x), 5)
    out = []
    for per_call in ((3, 3, 3), (1,) * 9):
        ctx = fluid.Context("cuda:0")
        xd, vd = dev(x), dev(v)
        rho = torch.zeros(len(x), dtype=torch.float32, device="cuda:0")
        done = 0
        for k in per_call:
            box = 3 <= done < 6
            ctx.set_params(flags=2 if box else 0)
            ctx.step(xd, vd, rho, k)
            done += k
        torch.cuda.synchronize()
        out.append((xd.cpu().numpy(), vd.cpu().numpy(), rho.cpu().numpy()))
    assert_same_bits(*out)
