"""Launch settings (pbf_set_tuning) and the autotuner (pbf_autotune).

Bars: every setting (PBD-pass workgroups of 64, 128, 256 and 512 threads,
staged lists always or never, ordinary or persistent launches, one lane or a
lane pair per particle) gives bit-identical positions, velocities and density
to the default setting, on one scene, on a staged lattice and on a batch; the
autotuner leaves the caller's state unchanged, applies the setting it reports,
stores it in its JSON cache and reads it back from there. Every call goes
through the C ABI."""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import dev, say  # noqa: E402

SETTINGS = [(b, s, p) for p in (0, 1) for b in (64, 128, 256, 512) for s in (1, -1)]
# one lane and a lane pair per particle, forced (the default picks by size)
LANE_SETTINGS = [(b, s, 0, lanes) for lanes in (1, 2) for b in (64, 256) for s in (1, -1)]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def bits(a):
    return a.cpu().numpy().view(np.uint32)


def run(ctx, x0, v0, nsteps, off=None):
    x, v = dev(x0), dev(v0)
    rho = torch.zeros(len(x0), dtype=torch.float32, device="cuda:0")
    if off is None:
        ctx.step(x, v, rho, nsteps)
    else:
        ctx.batch_step(x, v, off, rho, nsteps)
    return bits(x), bits(v), bits(rho)


def assert_settings_identical(fluid, x0, v0, nsteps, off=None, **params):
    ctx = fluid.Context("cuda:0", **params)
    want = run(ctx, x0, v0, nsteps, off)
    for setting in [t + (0,) for t in SETTINGS] + LANE_SETTINGS:
        ctx.set_tuning(*setting)
        got = run(ctx, x0, v0, nsteps, off)
        if setting[3]:
            assert ctx.stats()["pass_lanes"] == setting[3]
        for name, g, w in zip(("x", "v", "rho"), got, want):
            np.testing.assert_array_equal(g, w, err_msg=f"pass_block, stage_min_particles, persistent, pass_lanes "
                                                        f"{setting}: {name}")
    return ctx


def test_settings_identical_dam_break(fluid, scene):
    say("launch settings, 32,768-particle dam break, 3 steps")
    x0 = scene.jittered_lattice((32, 32, 32), (0.10, 0.50, 0.10), 0.005, 0.1, seed=1)
    ctx = assert_settings_identical(fluid, x0, np.zeros_like(x0), 3)
    ctx.set_tuning(256, 0, 0)  # automatic lanes: a small search takes the lane pair
    run(ctx, x0, np.zeros_like(x0), 1)
    assert ctx.stats()["pass_lanes"] == 2


def test_settings_identical_staged_lattice_with_modules(fluid, scene):
    say("launch settings, 327,680-particle lattice (staged lists), 2 steps, viscosity and box")
    x0 = scene.jittered_lattice((80, 64, 64), (0.2, 0.3, 0.2), 0.005, 0.1, seed=2)
    ctx = assert_settings_identical(fluid, x0, np.zeros_like(x0), 2, flags=3)
    ctx.set_tuning(256, 1, 1)
    run(ctx, x0, np.zeros_like(x0), 1)
    assert ctx.stats()["staged_waves"] > 0 and ctx.stats()["pass_lanes"] == 1
    ctx.set_tuning(256, 1, 0)  # automatic lanes: a staged search above 196,608 particles keeps one lane
    run(ctx, x0, np.zeros_like(x0), 1)
    assert ctx.stats()["pass_lanes"] == 1
    ctx.set_tuning(256, -1, 0)  # ... and an unstaged one takes the lane pair
    run(ctx, x0, np.zeros_like(x0), 1)
    assert ctx.stats()["pass_lanes"] == 2


def test_settings_identical_batch(fluid, scene):
    b = scene.batched_lattice_scenes(nscenes=24, nmin=256, nmax=8192, seed=4)
    say(f"launch settings, batch of 24 scenes ({int(b.offsets[-1])} particles), 2 steps")
    off = torch.from_numpy(b.offsets.astype(np.int32)).to("cuda:0")
    assert_settings_identical(fluid, b.pos, np.zeros_like(b.pos), 2, off)


def test_lane_pair_identical_on_32bit_lists(fluid, scene):
    """A slab 8 lattice layers thick whose cell planes hold 41k particles each:
    975 of its 2,048 waves reach over more than 65,535 sorted indices and keep
    32-bit lists (a numpy restatement of the sort and the runs counts them),
    which the lane pair runs on one lane of the pair. One lane and the pair
    give the same bits, with and without staged lists."""
    x0 = scene.jittered_lattice((128, 128, 8), (0.18, 0.18, 0.4), 0.005, 0.1, seed=6)
    say(f"lane pair vs one lane, {len(x0)}-particle slab (32-bit lists), 2 steps")
    ctx = fluid.Context("cuda:0")
    want = None
    for stage in (-1, 1):
        for lanes in (1, 2):
            ctx.set_tuning(256, stage, 0, lanes)
            got = run(ctx, x0, np.zeros_like(x0), 2)
            assert ctx.stats()["pass_lanes"] == lanes
            if want is None:
                want = got
                continue
            for name, g, w in zip(("x", "v", "rho"), got, want):
                np.testing.assert_array_equal(g, w, err_msg=f"stage {stage}, lanes {lanes}: {name}")


def test_autotune_and_cache(fluid, scene, tmp_path):
    x0 = scene.jittered_lattice((32, 32, 32), (0.10, 0.50, 0.10), 0.005, 0.1, seed=1)
    x, v = dev(x0), dev(np.zeros_like(x0))
    cache = tmp_path / "tune.json"
    ctx = fluid.Context("cuda:0")
    r = ctx.autotune(x, v, steps=3, cache=cache)
    say(f"autotune, 32,768 particles: {r}")
    assert r["candidates"] == 1 + len(SETTINGS) and not r["from_cache"]
    assert (r["pass_block"], r["stage_min_particles"], r["persistent"]) in SETTINGS + [(256, 0, 0)]
    assert 0 < r["best_ms"] <= r["default_ms"]
    assert ctx.tuning() == {"pass_block": r["pass_block"], "stage_min_particles": r["stage_min_particles"],
                            "persistent": r["persistent"], "pass_lanes": 0}
    np.testing.assert_array_equal(bits(x), x0.view(np.uint32))  # the caller's state is not touched
    np.testing.assert_array_equal(bits(v), np.zeros_like(x0).view(np.uint32))
    stored = json.loads(cache.read_text())
    assert stored[r["key"]]["pass_block"] == r["pass_block"]
    # a new context with the same workload key reads the setting back
    ctx2 = fluid.Context("cuda:0")
    r2 = ctx2.autotune(x, v, steps=3, cache=cache)
    assert r2["from_cache"] and r2["candidates"] == 0 and r2["key"] == r["key"]
    assert ctx2.tuning() == ctx.tuning()
    # another workload key adds an entry and keeps the first
    xs = dev(x0[:4096])
    r3 = fluid.Context("cuda:0").autotune(xs, torch.zeros_like(xs), steps=2, cache=cache)
    assert r3["key"] != r["key"] and not r3["from_cache"]
    assert set(json.loads(cache.read_text())) == {r["key"], r3["key"]}
    # a dense blob of the same size has its own key: the mean neighbor count is part of it
    xb = dev(scene.gaussian_blobs(1, 32768, 0.02, seed=3))
    r4 = fluid.Context("cuda:0").autotune(xb, torch.zeros_like(xb), steps=2, cache=cache)
    say(f"keys: lattice {r['key']}, blob {r4['key']}")
    assert " k2^6 " in r["key"] and r4["key"] != r["key"] and not r4["from_cache"]
    assert r4["key"].split(" k2^")[0] == r["key"].split(" k2^")[0]


def test_invalid_tuning_rejected(fluid, capi):
    ctx = fluid.Context("cuda:0")
    with pytest.raises(capi.PBFError):
        ctx.set_tuning(96, 0)
    with pytest.raises(capi.PBFError):
        ctx.set_tuning(256, 0, 2)
    with pytest.raises(capi.PBFError):
        ctx.set_tuning(256, 0, 0, 3)
    with pytest.raises(capi.PBFError):  # the lane pair has no persistent form
        ctx.set_tuning(256, 0, 1, 2)
    assert ctx.tuning() == {"pass_block": 256, "stage_min_particles": 0, "persistent": 0, "pass_lanes": 0}
