"""hipGraph replay of the middle steps of a multi-step call: results
bit-identical to launching every kernel directly (PBF_FLAG_EAGER), across
replays with eager work in between, after the context's buffers grow and
after a parameter change (both force a new capture).

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


def bits(t):
    return t.cpu().numpy().view(np.uint32)


def run(fluid, capi, x0, v0, calls, eager, **params):
    """calls: list of step counts, one pbf_step call each, with a one-step
    call between them; returns the states after each call and the stats."""
    flags = capi.PBF_FLAG_EAGER if eager else 0
    ctx = fluid.Context("cuda:0", flags=flags, **params)
    x, v = dev(x0), dev(v0)
    rho = torch.zeros(len(x0), dtype=torch.float32, device="cuda:0")
    out = []
    for k in calls:
        ctx.step(x, v, rho, k)
        out.append((bits(x), bits(v), bits(rho)))
        ctx.step(x, v, None, 1)
    return out, ctx.stats()


@pytest.mark.parametrize("name", ["lattice", "blob", "staged"])
def test_graph_replay_equals_eager(fluid, capi, scene, name):
    say(f"graph replay vs eager launches: {name}")
    params = {}
    if name == "lattice":
        x0 = scene.dam_break_32k()
    elif name == "blob":
        x0 = scene.gaussian_blobs(2, 6000, 0.02, seed=23)
    else:  # staged lists (the 1M workload's list form) on a small input
        x0 = scene.jittered_lattice((40, 40, 40), (0.2, 0.4, 0.2), 0.005, 0.1, seed=24)
        params = dict(stage_min_particles=1, list_slots=256)
    v0 = (0.1 * (scene.uniform01(25, x0.size).reshape(-1, 3) - 0.5)).astype(np.float32)
    calls = [6, 3, 8]
    got, st = run(fluid, capi, x0, v0, calls, False, **params)
    want, st_e = run(fluid, capi, x0, v0, calls, True, **params)
    assert st["graph_replays"] == sum(k - 2 for k in calls)
    assert st["graph_captures"] >= 1 and st_e["graph_replays"] == 0
    for c, (g, w) in enumerate(zip(got, want)):
        for part, a, b in zip(("x", "v", "rho"), g, w):
            np.testing.assert_array_equal(a, b, err_msg=f"call {c} {part}")


def test_graph_batched_equals_eager(fluid, capi, scene):
    say("graph replay vs eager launches: 24-scene batch, 5 steps")
    b = scene.batched_lattice_scenes(nscenes=24, nmin=300, nmax=6000, seed=26)
    off = torch.from_numpy(b.offsets.astype(np.int32)).to("cuda:0")
    res = []
    for eager in (False, True):
        ctx = fluid.Context("cuda:0", flags=capi.PBF_FLAG_EAGER if eager else 0)
        x, v = dev(b.pos), dev(np.zeros_like(b.pos))
        ctx.batch_step(x, v, off, None, 5)
        ctx.batch_step(x, v, off, None, 5)
        res.append((bits(x), bits(v), ctx.stats()["graph_replays"]))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
    assert res[0][2] == 6 and res[1][2] == 0


def test_graph_recaptured_after_growth_and_params(fluid, capi, scene):
    """A larger input reallocates the context's buffers and a new smoothing
    length changes the kernels' constants: each forces a new capture, and
    the results still equal eager launches."""
    say("graph recapture after buffer growth and a parameter change")
    small = scene.jittered_lattice((10, 10, 10), (0.3, 0.5, 0.3), 0.005, 0.1, seed=27)
    big = scene.jittered_lattice((30, 30, 30), (0.3, 0.5, 0.3), 0.005, 0.1, seed=28)
    outs = []
    for eager in (False, True):
        ctx = fluid.Context("cuda:0", flags=capi.PBF_FLAG_EAGER if eager else 0)
        r = []
        for x0, h in ((small, 0.0125), (big, 0.0125), (big, 0.011), (small, 0.011)):
            ctx.set_params(smoothing_length=h)
            x, v = dev(x0), dev(np.zeros_like(x0))
            ctx.step(x, v, None, 4)
            r.append((bits(x), bits(v)))
        outs.append((r, ctx.stats()))
    (got, st), (want, _) = outs
    assert st["graph_captures"] == 4 and st["graph_replays"] == 8
    for k, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(g[0], w[0], err_msg=f"case {k} x")
        np.testing.assert_array_equal(g[1], w[1], err_msg=f"case {k} v")


def test_profiling_and_flags_disable_replay(fluid, capi, scene):
    say("no replay while profiling")
    x0 = scene.jittered_lattice((12, 12, 12), (0.3, 0.5, 0.3), 0.005, 0.1, seed=29)
    ctx = fluid.Context("cuda:0")
    ctx.set_profiling(True)
    x, v = dev(x0), dev(np.zeros_like(x0))
    ctx.step(x, v, None, 5)
    ctx.profile()
    ctx.set_profiling(False)
    assert ctx.stats()["graph_replays"] == 0
    ctx.step(x, v, None, 5)
    assert ctx.stats()["graph_replays"] == 3
