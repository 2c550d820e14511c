"""Parity on the benchmark workloads that the other GPU tests cover only in
part: the Gaussian-blob density of the 1M blob workload (8 x 131,072
particles, sigma 0.05) and the 256-scene batch (n_s log-uniform in
[2048, 65536]).

Bars: neighbor sets bit-exact; after one step positions within 1e-6 m and
velocities within 1e-3 m/s of the FP32 restatement, density within 1e-5
relative of FP64 (or twice the FP32 restatement's own distance from FP64);
a batch bit-identical to each scene stepped alone.

Every call goes through the C ABI."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import (assert_same_sets, assert_step_parity, dev, oracle_both,  # noqa: E402
                         row_checksums, say)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_blob_density_one_step(fluid, oracle, scene):
    """One blob of 16,384 particles at sigma 0.025 has the number density of
    the 1M blob workload (131,072 per blob at sigma 0.05: N / sigma^3 is the
    same), so its centre holds the ~550 neighbors of that workload. Sets,
    positions, velocities and density after one step against the oracle."""
    say("dense Gaussian blob: 16,384 particles, sigma 0.025, one step")
    x = scene.gaussian_blobs(1, 16384, 0.025, seed=3)
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    gpu = model.step_arrays(x, v, 1)
    nl = model.ctx.step_neighbors(len(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    cnt = np.diff(off)
    say(f"neighbors: mean {cnt.mean():.1f}, max {cnt.max()}")
    assert cnt.max() > 450  # the centre of the 1M blob workload
    assert_same_sets(nl, off, el)
    f32, f64 = oracle_both(oracle, x, v, 1)
    assert_step_parity(gpu, f32, f64)


def test_blob_density_three_steps(fluid, oracle, scene):
    """Three steps of the dense blob. It explodes (lambda is large at the
    centre, speeds reach 22 m/s), and after three steps the FP32 restatement
    is already 2e-4 m and 0.1 m/s from FP64 pointwise: a trajectory this
    sensitive is compared through its aggregates, mean density and kinetic
    energy within 0.5% (the bar for long runs), with the pointwise distances
    reported."""
    say("dense Gaussian blob: 16,384 particles, sigma 0.025, three steps (aggregates)")
    x = scene.gaussian_blobs(1, 16384, 0.025, seed=3)
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    xg, vg, rg = model.step_arrays(x, v, 3)
    f32, f64 = oracle_both(oracle, x, v, 3)
    for name, got in (("gpu", (xg, vg, rg)), ("fp32 oracle", f32[:3])):
        say(f"{name} vs fp64 after 3 steps: max |dx| {np.abs(got[0] - f64[0]).max():.3e} m, "
            f"max |dv| {np.abs(got[1] - f64[1]).max():.3e} m/s")
    for what, get in (("mean density", lambda s: float(np.mean(np.asarray(s[2], np.float64)))),
                      ("kinetic energy", lambda s: float(np.sum(np.asarray(s[1], np.float64) ** 2)))):
        a, b = get((xg, vg, rg)), get(f64)
        say(f"{what}: gpu {a:.6g}, fp64 {b:.6g}, rel {abs(a - b) / abs(b):.3e}")
        assert abs(a - b) <= 5e-3 * abs(b), what


def test_blobs_1m_neighbor_checksums(fluid, oracle, scene):
    """The full 1M blob workload (8 x 131,072, sigma 0.05, seed 3): per-row
    neighbor checksums against the oracle."""
    say("1M Gaussian blobs: per-row neighbor checksums vs oracle")
    x = scene.blobs_1m()
    ctx = fluid.Context("cuda:0")
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    go = nl.offsets().cpu().numpy()
    ge = nl.elements.cpu().numpy()
    say(f"pairs {off[-1]}, max row {np.diff(off).max()}")
    np.testing.assert_array_equal(go, off)
    for a, b in zip(row_checksums(go, ge), row_checksums(off, el)):
        np.testing.assert_array_equal(a, b)


def test_batched_256_scenes_equal_serial(fluid, oracle, scene):
    """The 256-scene batch (n_s log-uniform in [2048, 65536], seed 4, about
    4.6M particles): two batched steps bit-identical to each scene stepped
    alone, and three scenes (smallest, median, largest) against the oracle."""
    b = scene.batched_lattice_scenes(nscenes=256, nmin=2048, nmax=65536, seed=4)
    n = int(b.offsets[-1])
    say(f"256-scene batch: {n} particles, two steps, batch vs serial")
    dv = torch.device("cuda:0")
    off = torch.from_numpy(b.offsets.astype(np.int32)).to(dv)
    xb = dev(b.pos)
    vb = torch.zeros_like(xb)
    rb = torch.zeros(n, dtype=torch.float32, device=dv)
    ctx = fluid.Context("cuda:0")
    ctx.reserve(n, 256)
    ctx.batch_step(xb, vb, off, rb, 2)
    xs_b, vs_b, rs_b = xb.cpu().numpy(), vb.cpu().numpy(), rb.cpu().numpy()
    sc = fluid.Context("cuda:0")
    sc.reserve(int(b.counts.max()))
    for s in range(256):
        a, e = int(b.offsets[s]), int(b.offsets[s + 1])
        x1 = dev(b.pos[a:e])
        v1 = torch.zeros_like(x1)
        r1 = torch.zeros(e - a, dtype=torch.float32, device=dv)
        sc.step(x1, v1, r1, 2)
        for name, got, want in (("x", xs_b[a:e], x1), ("v", vs_b[a:e], v1), ("rho", rs_b[a:e], r1)):
            np.testing.assert_array_equal(got.view(np.uint32), want.cpu().numpy().view(np.uint32),
                                          err_msg=f"scene {s} {name}")
    order = np.argsort(b.counts)
    for s in (int(order[0]), int(order[128]), int(order[-1])):
        a, e = int(b.offsets[s]), int(b.offsets[s + 1])
        say(f"scene {s}: {e - a} particles vs oracle, two steps")
        x0 = b.pos[a:e]
        assert_step_parity((xs_b[a:e], vs_b[a:e], rs_b[a:e]), *oracle_both(oracle, x0, np.zeros_like(x0), 2))
