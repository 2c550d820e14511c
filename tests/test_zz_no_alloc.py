"""pbf_pbd_constrain_noalloc (DensityPBD::constrain with the library's fixed
per-device scratch) against pbf_pbd_constrain, the entry point it varies, on
the same caller NeighborList: two calls in a row, the second with fewer
particles, so the second runs on scratch the first left behind. Bars as the
step parity: x within 1e-6 m, v within 1e-3 m/s, density within 1e-5
relative (the float atomics of both sum in an undefined order, as the
reference's do). Also: no device memory is taken by the calls, and a count
beyond the scratch is refused. Every call goes through the C ABI."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import RHO_REL, V_ABS, X_ABS, dev, rel_err, say  # noqa: E402

MAX_PARTICLES = 1 << 20  # PBF_NOALLOC_MAX_PARTICLES


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def compressed_case(scene, dims, seed):
    """A lattice at spacing 0.0042 (density above rest: lambda is active) with
    small seeded velocities, and its neighbor list from pbf_neighbor_query."""
    x = scene.jittered_lattice(dims, (0.3, 0.5, 0.3), 0.0042, 0.1, seed=seed)
    n = len(x)
    v = (0.05 * (scene.uniform01(seed + 100, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)
    return x, v


def on_device(x, v):
    return dev(x), dev(v), torch.zeros(len(x), dtype=torch.float32, device="cuda:0")


def host(state):
    return tuple(t.cpu().numpy() for t in state)


def test_noalloc_constrain_equals_constrain(fluid, scene):
    ctx = fluid.Context("cuda:0")
    ref_ctx = fluid.Context("cuda:0")
    cases = [compressed_case(scene, (48, 40, 36), 21), compressed_case(scene, (17, 23, 11), 22)]
    lists, refs, states = [], [], []
    for x, v in cases:
        nl = ctx.neighbor_query(dev(x))
        lists.append(nl)
        ref = on_device(x, v)
        ref_ctx.pbd_constrain(ref[0], ref[1], nl, ref[2])
        refs.append(host(ref))
        states.append(on_device(x, v))
    warm = on_device(*cases[1])
    ctx.pbd_constrain_noalloc(warm[0], warm[1], lists[1], warm[2])  # code objects loaded before the reading
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    for st, nl in zip(states, lists):  # every buffer exists already: torch allocates nothing here
        ctx.pbd_constrain_noalloc(st[0], st[1], nl, st[2])
        torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(0)[0]
    for k, ((x, v), nl, ref, st) in enumerate(zip(cases, lists, refs, states)):
        got = host(st)
        ex = np.abs(got[0].astype(np.float64) - ref[0]).max()
        ev = np.abs(got[1].astype(np.float64) - ref[1]).max()
        er = rel_err(got[2], ref[2]).max()
        moved = np.abs(ref[0].astype(np.float64) - x).max()
        say(f"no-alloc DensityPBD::constrain, call {k + 1}: {len(x)} particles, {nl.elements.numel()} pairs; "
            f"vs pbf_pbd_constrain: x {ex:.3e} m, v {ev:.3e} m/s, rho {er:.3e} (largest correction {moved:.3e} m)")
        assert moved > 1e-5  # the constraint did move particles
        assert ex <= X_ABS and ev <= V_ABS and er <= RHO_REL
    say(f"device memory free before and after the calls: {free0} / {free1} bytes")
    assert free1 >= free0


def test_noalloc_constrain_refuses_more_than_its_scratch(fluid, capi):
    ctx = fluid.Context("cuda:0")
    n = MAX_PARTICLES + 1
    x = torch.zeros((n, 3), dtype=torch.float32, device="cuda:0")
    v = torch.zeros_like(x)
    index = torch.arange(n, dtype=torch.int32, device="cuda:0")
    nl = fluid.NeighborList(index, index.clone())
    with pytest.raises(capi.PBFError) as e:
        ctx.pbd_constrain_noalloc(x, v, nl, None)
    assert e.value.status == -5 and "exceed" in str(e.value)


def test_noalloc_constrain_empty_input(fluid):
    ctx = fluid.Context("cuda:0")
    x = torch.zeros((0, 3), dtype=torch.float32, device="cuda:0")
    nl = fluid.NeighborList(torch.zeros(0, dtype=torch.int32, device="cuda:0"),
                            torch.zeros(0, dtype=torch.int32, device="cuda:0"))
    ctx.pbd_constrain_noalloc(x, torch.zeros_like(x), nl, None)
