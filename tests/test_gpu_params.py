"""The fused step at parameters other than the benchmark's, against the oracle.

Cases:
- the reference's own defaults: h = 0.006 with the 0.005 sampling distance
  (PositionBasedFluidModel.cpp:25) and 3 iterations (DensityPBD.cu:242),
  about 7 neighbors per particle;
- a wide kernel, h/d = 4: about 270 neighbors per particle, more mask words
  per lane than the neighbor pass keeps in LDS;
- one iteration with a four times longer time step and gravity along z;
- h/d = 1.8 with 4 iterations.
Bars: neighbor sets bit-exact; one step of the fused step and of the module
sequence (NeighborQuery::compute, integrate, DensityPBD::constrain) within the
parity bars (x 1e-6 m, v 1e-3 m/s, rho 1e-5 relative, each against FP64 with
twice the FP32 oracle's own distance from FP64 as the floor). Every call goes
through the C ABI."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import assert_same_sets, assert_step_parity, dev, oracle_both, say  # noqa: E402

# name: (h, d, iterations, dt, gravity, lattice dims)
CASES = {
    "reference_defaults": (0.006, 0.005, 3, 0.001, (0.0, -9.8, 0.0), (24, 24, 24)),
    "wide_kernel": (0.02, 0.005, 2, 0.001, (0.0, -9.8, 0.0), (20, 20, 20)),
    "one_iteration_long_dt": (0.0125, 0.005, 1, 0.004, (0.0, 0.0, -9.8), (24, 24, 24)),
    "h_over_d_1_8": (0.009, 0.005, 4, 0.001, (0.0, -9.8, 0.0), (24, 24, 24)),
}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("case", list(CASES))
def test_parameters_vs_oracle(fluid, oracle, scene, case):
    h, d, iters, dt, g, dims = CASES[case]
    x = scene.jittered_lattice(dims, (0.4, 0.4, 0.4), d, 0.1, seed=7)
    v = np.zeros_like(x)
    say(f"{case}: h {h}, d {d}, {iters} iterations, dt {dt}, gravity {g}, {len(x)} particles")
    okw = dict(h=h, d=d, iters=iters, dt=dt, g=g)
    gkw = dict(sampling_distance=d, iterations=iters, dt=dt, gravity=list(g))

    ctx = fluid.Context("cuda:0", smoothing_length=h, **gkw)
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(**okw), x)
    st = ctx.stats()
    say(f"{case}: {off[-1]} pairs ({off[-1] / len(x):.1f} per particle), mask words needed "
        f"{st['mask_words_needed']}, fallback waves {st['mask_fallback_waves']}")
    assert_same_sets(nl, off, el)

    model = fluid.PositionBasedFluidModel(h, "cuda:0", viscosity=False, **gkw)
    gpu = model.step_arrays(x, v, 1)
    ref = oracle_both(oracle, x, v, 1, **okw)
    assert_step_parity(gpu, *ref)
    # the module sequence (NeighborQuery::compute, integrate, DensityPBD::constrain) at the same parameters
    say(f"{case}: module path")
    assert_step_parity(model.step_arrays(x, v, 1, modules=True), *ref)
