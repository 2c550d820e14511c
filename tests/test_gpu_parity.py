"""Parity of libpbf.so (HIP, gfx950) against the CPU oracle.

Bars: neighbor sets bit-exact (integer work); positions within 1e-6 m (1e-6 of
the unit scene box) and velocities within 1e-3 m/s (the position bar divided
by dt) of the FP32 restatement; over 100 steps mean density and kinetic
energy within 0.5%; free fall bit-exact.

Density after a step is held to 1e-5 relative against the FP64 instantiation
of the oracle, or to twice the FP32 restatement's own deviation from FP64 on
the same input when that is larger: fp32 positions near 0.5 m carry 6e-8 m
of rounding per update, which moves the density of sparse (surface) particles
by ~1e-5 relative after five iterations, so two fp32 evaluations that differ
only in summation order (the reference's own order is that of float atomics)
legitimately differ at that level. A single density pass on identical inputs
(test_density_on_caller_list) meets 1e-5 directly.

Every call goes through the C ABI."""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from parity_util import (RHO_REL, V_ABS, X_ABS, assert_same_sets, assert_step_parity, dev,  # noqa: E402,F401
                         oracle_both, rel_err, row_checksums, row_keys, say)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


# ----------------------------------------------------------- neighbor search
def test_neighbor_sets_bit_exact_dam_break(fluid, oracle, scene):
    say("neighbor sets, 32768-particle dam break")
    x = scene.dam_break_32k()
    ctx = fluid.Context("cuda:0")
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    assert_same_sets(nl, off, el)
    st = ctx.stats()
    assert st["neighbor_pairs"] == off[-1]
    assert st["grid_dims"] == [85, 85, 85]
    assert st["out_of_grid"] == 0


def test_neighbor_sets_gaussian_blob(fluid, oracle, scene):
    say("neighbor sets, uneven occupancy (Gaussian blobs)")
    x = scene.gaussian_blobs(2, 12000, 0.03, seed=21)
    ctx = fluid.Context("cuda:0")
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    assert_same_sets(nl, off, el)
    assert np.diff(off).max() > 150  # dense centre


def test_neighbor_sets_crowded_cell(fluid, oracle, scene):
    """More than 1024 particles in one cell (the in-cell rank falls back to
    arrival order there); the sets must not change."""
    say("neighbor sets, 1200 particles in one cell")
    u = scene.uniform01(33, 3 * 1200).reshape(1200, 3)
    x = (0.5 + 0.011 * u).astype(np.float32)
    ctx = fluid.Context("cuda:0")
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    assert_same_sets(nl, off, el)


def half_cell_pair_cloud():
    """36 copies (0.05 apart in y and z) of a pair closer than h whose half-cell
    indices along x differ by 3: x_i = 0.106249996 (q = 10.499999, half cell 20,
    reference cell 10) and x_j = 0.11874999 (q = 11.5, half cell 23, reference
    cell 11). Found by a search over float32 positions near half-cell faces."""
    xi, xj = np.float32(0.106249996), np.float32(0.11874999)
    pts = []
    for k in range(6):
        for l in range(6):
            y, z = np.float32(0.3 + 0.05 * k), np.float32(0.3 + 0.05 * l)
            pts += [(xi, y, z), (xj, y, z)]
    return np.array(pts, dtype=np.float32)


def test_neighbor_sets_half_cell_rounding_pair(fluid, oracle):
    """The sort grid halves the cells along x and visits 5 half cells; a pair
    that rounding puts 3 half cells apart is found by the stencil's rounding
    extension, as the reference's 3-cell stencil finds it."""
    say("neighbor sets, pairs 3 half cells apart")
    x = half_cell_pair_cloud()
    ctx = fluid.Context("cuda:0")
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    assert np.all(np.diff(off) == 2)  # self and the partner
    assert_same_sets(nl, off, el)


def test_neighbor_query_capacity_error(fluid, capi, scene):
    import ctypes as C
    say("neighbor query capacity error")
    x = dev(scene.jittered_lattice((8, 8, 8), (0.3, 0.3, 0.3), 0.005, 0.1, seed=2))
    ctx = fluid.Context("cuda:0")
    off = torch.empty(513, dtype=torch.int32, device="cuda:0")
    el = torch.empty(10, dtype=torch.int32, device="cuda:0")
    total = C.c_longlong()
    st = capi.lib().pbf_neighbor_query(ctx.handle, C.c_void_p(x.data_ptr()), 512, C.c_void_p(off.data_ptr()),
                                       C.c_void_p(el.data_ptr()), 10, C.byref(total), None)
    assert st == capi.PBF_ERR_CAPACITY
    assert total.value > 10
    assert b"need" in capi.lib().pbf_last_error(ctx.handle)


# ------------------------------------------------------------- fused step
def test_one_step_parity_dam_break(fluid, oracle, scene):
    say("fused step, 32768-particle dam break, 1 step")
    x = scene.dam_break_32k()
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    gpu = model.step_arrays(x, v, 1)
    assert_step_parity(gpu, *oracle_both(oracle, x, v, 1))


def test_module_path_parity_dam_break(fluid, oracle, scene):
    say("module path NeighborQuery -> integrate -> DensityPBD, 1 step")
    x = scene.dam_break_32k()
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    gpu = model.step_arrays(x, v, 1, modules=True)
    assert_step_parity(gpu, *oracle_both(oracle, x, v, 1))


def test_step_neighbor_export(fluid, oracle, scene):
    say("CSR export after pbf_step")
    x = scene.dam_break_32k()
    ctx = fluid.Context("cuda:0")
    xd, vd = dev(x), dev(np.zeros_like(x))
    ctx.step(xd, vd, None, 1)
    nl = ctx.step_neighbors(len(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    assert_same_sets(nl, off, el)


def test_hundred_step_aggregates(fluid, oracle, scene):
    say("fused step, 32768 particles, 100 steps, mean density and kinetic energy")
    x = scene.dam_break_32k()
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    xs, vs, rho = model.step_arrays(x, v, 100)
    xo, vo, ro, _ = oracle.step(oracle.make_params(), x, v, 100)
    assert abs(rho.mean() / ro.mean() - 1) <= 5e-3
    ke, keo = 0.5 * (vs.astype(np.float64) ** 2).sum(1).mean(), 0.5 * (vo.astype(np.float64) ** 2).sum(1).mean()
    assert abs(ke / keo - 1) <= 5e-3


def test_step_is_bitwise_reproducible(fluid, scene):
    say("fused step reproducibility, 2 runs x 3 steps")
    x = scene.dam_break_32k()
    outs = []
    for _ in range(2):
        model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
        outs.append(model.step_arrays(x, np.zeros_like(x), 3))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


def test_free_fall_bit_exact(fluid):
    say("free fall, 20 steps")
    x0 = np.array([[0.2, 0.5, 0.2], [0.5, 0.6, 0.5], [0.8, 0.7, 0.3]], np.float32)
    v0 = np.array([[0.1, 0.0, -0.2], [0.0, 0.3, 0.0], [-0.05, 0.0, 0.07]], np.float32)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    xs, vs, rho = model.step_arrays(x0, v0, 20)
    ex, ev = x0.copy(), v0.copy()
    g = np.array([0.0, -9.8, 0.0], np.float32)
    for _ in range(20):
        ev = ev + (np.float32(0) + g) * np.float32(0.001)
        ex = ex + ev * np.float32(0.001)
    np.testing.assert_array_equal(vs, ev)
    np.testing.assert_array_equal(xs, ex)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 4097])
def test_ragged_sizes(fluid, oracle, scene, n):
    say(f"fused step, ragged n={n}")
    x = scene.jittered_lattice((17, 17, 17), (0.3, 0.5, 0.3), 0.005, 0.1, seed=n, count=n)
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    gpu = model.step_arrays(x, v, 2)
    assert_step_parity(gpu, *oracle_both(oracle, x, v, 2))


def test_empty_input(fluid):
    say("n = 0")
    ctx = fluid.Context("cuda:0")
    x = torch.empty((0, 3), dtype=torch.float32, device="cuda:0")
    ctx.step(x, x.clone(), None, 3)
    nl = ctx.neighbor_query(x)
    assert nl.index.numel() == 0 and nl.elements.numel() == 0 and nl.size() == 0


def test_particle_count_limit(fluid, capi, oracle, scene):
    """2^27 particles or more in one call (the passes' 32-bit byte offsets of
    16-byte records) are refused with PBF_ERR_UNSUPPORTED before any access
    to the caller's buffers, and the context keeps working afterwards."""
    import ctypes as C
    say("n = 2^27: refused, then a 512-particle step on the same context")
    x_h = scene.jittered_lattice((8, 8, 8), (0.3, 0.3, 0.3), 0.005, 0.1, seed=4)
    ctx = fluid.Context("cuda:0")
    x, v = dev(x_h), torch.zeros((len(x_h), 3), dtype=torch.float32, device="cuda:0")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    L = capi.lib()
    assert L.pbf_step(ctx.handle, p(x), p(v), None, 1 << 27, 1, None) == capi.PBF_ERR_UNSUPPORTED
    assert b"2^27" in L.pbf_last_error(ctx.handle)
    off = torch.empty(16, dtype=torch.int32, device="cuda:0")
    total = C.c_longlong()
    st = L.pbf_neighbor_query(ctx.handle, p(x), 1 << 27, p(off), p(off), 16, C.byref(total), None)
    assert st == capi.PBF_ERR_UNSUPPORTED
    nl = ctx.neighbor_query(x)
    assert_same_sets(nl, *oracle.neighbor_query(oracle.make_params(), x_h))


@pytest.mark.parametrize("scene_name,reserve", [
    ("lattice", {"list_slots": 1}),                     # list full: words re-tested (kept in LDS only)
    ("blob", {"list_slots": 1, "mask_words": 512}),     # list full: dense waves read global bit masks
    ("blob", {"list_slots": 1, "mask_words": 1}),       # both full: every wave re-tests
])
def test_reservation_fallbacks_are_identical(fluid, scene, scene_name, reserve):
    """Reservations too small for the scene make waves iterate global bit masks
    or re-test their candidates in every pass; the results must be
    bit-identical to the neighbor-list path. Automatic reservations (the
    default) grow for the next call; explicit ones stay as given."""
    say(f"neighbor storage fallback {scene_name} {reserve}")
    if scene_name == "lattice":
        x = scene.dam_break_32k()
    else:
        x = scene.gaussian_blobs(1, 20000, 0.02, seed=23)  # > 24 mask words per lane at the centre
    v = np.zeros_like(x)
    a = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    a.step_arrays(x, v, 1)
    ra = a.step_arrays(x, v, 1)  # second call: reservations sized for the scene, list path
    assert a.ctx.stats()["list_slots_needed"] <= a.ctx.stats()["list_slots_reserved"]
    b = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False, **reserve)
    rb = b.step_arrays(x, v, 1)
    st = b.ctx.stats()
    assert st["list_slots_needed"] > st["list_slots_reserved"]
    if reserve.get("mask_words") == 1:
        assert st["mask_fallback_waves"] > 0
    elif scene_name == "blob":
        assert st["mask_words_needed"] > 0 and st["mask_fallback_waves"] == 0
    for u, w in zip(ra, rb):
        np.testing.assert_array_equal(u, w)
    rb2 = b.step_arrays(x, v, 1)  # explicit reservations: the same fallbacks again
    st = b.ctx.stats()
    assert st["list_slots_needed"] > st["list_slots_reserved"]
    for u, w in zip(ra, rb2):
        np.testing.assert_array_equal(u, w)


def staged_scene(scene, name):
    if name == "lattice":
        return scene.dam_break_32k()
    if name == "blob":  # dense centre: its waves keep unstaged lists or masks
        return scene.gaussian_blobs(1, 20000, 0.02, seed=23)
    # out-of-grid particles near the grid's lower x face, next to in-grid ones
    return scene.jittered_lattice((20, 16, 16), (-0.047, 0.45, 0.45), 0.005, 0.1, seed=12)


@pytest.mark.parametrize("scene_name", ["lattice", "blob", "oog"])
def test_staged_lists_are_identical(fluid, scene, scene_name):
    """Staged lists (each wavefront's PBD passes read the neighbors in its own
    z-plane rows from a copy in LDS) give the same bits as unstaged lists, over
    several steps, and the same CSR export."""
    say(f"staged vs unstaged lists, {scene_name}, 3 steps")
    x = staged_scene(scene, scene_name)
    v = np.zeros_like(x)
    outs, stats, csr = [], [], []
    for smin in (1, -1):
        m = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False, list_slots=320, mask_words=512,
                                          stage_min_particles=smin)
        outs.append(m.step_arrays(x, v, 3))
        stats.append(m.ctx.stats())
        nl = m.ctx.step_neighbors(len(x))
        csr.append((nl.offsets().cpu().numpy(), nl.elements.cpu().numpy()))
    nw = (len(x) + 63) // 64
    say(f"staged waves {stats[0]['staged_waves']} of {nw}")
    if scene_name == "lattice":
        assert stats[0]["staged_waves"] >= 0.9 * nw
    else:
        assert stats[0]["staged_waves"] > 0
    assert stats[1]["staged_waves"] == 0
    for u, w in zip(*outs):
        np.testing.assert_array_equal(u, w)
    np.testing.assert_array_equal(csr[0][0], csr[1][0])
    np.testing.assert_array_equal(csr[0][1], csr[1][1])


def test_out_of_grid_particles(fluid, oracle, scene):
    """Particles outside the padded grid query but are not hashed
    (GridHash.cu:85,194); the one-sided pairs they create are reproduced."""
    say("out-of-grid particles near the grid's lower x face")
    x = scene.jittered_lattice((8, 8, 8), (-0.047, 0.5, 0.5), 0.005, 0.1, seed=12)
    v = np.zeros_like(x)
    p = oracle.make_params(g=(0.0, 0.0, 0.0))
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False, gravity=(0.0, 0.0, 0.0))
    gpu = model.step_arrays(x, v, 1)
    c3, cid = oracle.cell3(p, x)
    n_oog = int((cid < 0).sum())
    assert 0 < n_oog < len(x)
    assert model.ctx.stats()["out_of_grid"] == n_oog
    assert_step_parity(gpu, *oracle_both(oracle, x, v, 1, g=(0.0, 0.0, 0.0)))


def test_nonfinite_positions_reported(fluid):
    say("NaN / Inf positions reported")
    x = np.array([[0.5, 0.5, 0.5], [np.nan, 0.5, 0.5], [0.5, np.inf, 0.5]], np.float32)
    ctx = fluid.Context("cuda:0")
    ctx.neighbor_query(dev(x))
    st = ctx.stats()
    assert st["nonfinite"] == 2 and st["out_of_grid"] == 2


# ----------------------------------------------------- module-granular calls
def test_integrate_bit_exact(fluid, oracle, scene):
    say("pbf_integrate with force")
    x = scene.jittered_lattice((10, 10, 10), (0.3, 0.5, 0.3), 0.005, 0.1, seed=3)
    v = (scene.uniform01(5, x.size).reshape(x.shape) - 0.5).astype(np.float32)
    f = (scene.uniform01(6, x.size).reshape(x.shape) * 20 - 10).astype(np.float32)
    ctx = fluid.Context("cuda:0")
    xd, vd, fd = dev(x), dev(v), dev(f)
    ctx.integrate(xd, vd, fd)
    xo, vo = oracle.integrate(oracle.make_params(), x, v, f)
    np.testing.assert_array_equal(vd.cpu().numpy(), vo)
    np.testing.assert_array_equal(xd.cpu().numpy(), xo)


def test_density_on_caller_list(fluid, oracle, scene):
    say("pbf_density on the oracle's CSR")
    x = scene.dam_break_32k()
    p = oracle.make_params()
    off, el = oracle.neighbor_query(p, x)
    ctx = fluid.Context("cuda:0")
    nl = fluid.NeighborList(torch.from_numpy(off[:-1].copy()).to("cuda:0"), torch.from_numpy(el).to("cuda:0"))
    rho = torch.empty(len(x), dtype=torch.float32, device="cuda:0")
    ctx.density(dev(x), nl, rho)
    ro = oracle.density(p, x, off, el)
    assert (np.abs(rho.cpu().numpy() - ro) / ro).max() <= RHO_REL


def test_pbd_constrain_on_asymmetric_list(fluid, oracle, scene):
    """DensityPBD over a caller list that is not symmetric (one-sided pairs):
    the pairwise scatter of the reference, not the symmetric gather."""
    say("pbf_pbd_constrain, asymmetric caller list")
    x = scene.jittered_lattice((10, 10, 10), (0.3, 0.5, 0.3), 0.0045, 0.1, seed=9)
    v = np.zeros_like(x)
    p = oracle.make_params()
    off, el = oracle.neighbor_query(p, x)
    # drop every third row's last neighbor
    keep = np.ones(len(el), bool)
    rows = np.arange(len(x))
    drop = off[1:][rows % 3 == 0] - 1
    keep[drop] = False
    cnt = np.diff(off) - (rows % 3 == 0)
    off2 = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    el2 = el[keep].astype(np.int32)
    ctx = fluid.Context("cuda:0")
    nl = fluid.NeighborList(torch.from_numpy(off2[:-1].copy()).to("cuda:0"), torch.from_numpy(el2).to("cuda:0"))
    xd, vd = dev(x), dev(v)
    rho = torch.empty(len(x), dtype=torch.float32, device="cuda:0")
    ctx.pbd_constrain(xd, vd, nl, rho)
    f32 = oracle.pbd_constrain(p, x, v, off2, el2)
    f64 = oracle.pbd_constrain(oracle.make_params(double=True), x.astype(np.float64), v.astype(np.float64), off2, el2)
    assert_step_parity((xd.cpu().numpy(), vd.cpu().numpy(), rho.cpu().numpy()), f32, f64)


def test_invalid_params_rejected(fluid, capi):
    say("parameter validation")
    with pytest.raises(capi.PBFError) as e:
        fluid.Context("cuda:0", smoothing_length=-1.0)
    assert e.value.status == capi.PBF_ERR_INVALID
    ctx = fluid.Context("cuda:0")
    with pytest.raises(capi.PBFError):
        ctx.set_params(iterations=-2)
    assert ctx.params()["iterations"] == 5


# ------------------------------------------------------- full-size properties
def test_allocation_failure_recovers(fluid, capi, scene):
    """A reservation the device cannot hold (list_slots = 2^30 per particle:
    about 2 TB) fails with PBF_ERR_ALLOC before any kernel runs; after the
    parameter is set back, the same context steps again and gives the bits of
    a fresh context."""
    say("allocation failure: list reservation of 2^30 slots per particle, then recovery")
    x = scene.jittered_lattice((10, 10, 10), (0.3, 0.5, 0.3), 0.005, 0.1, seed=3)
    ctx = fluid.Context("cuda:0", list_slots=1 << 30)
    xd, vd = dev(x), dev(np.zeros_like(x))
    with pytest.raises(capi.PBFError) as e:
        ctx.step(xd, vd, None, 1)
    assert e.value.status == capi.PBF_ERR_ALLOC and "hipMalloc" in str(e.value)
    np.testing.assert_array_equal(xd.cpu().numpy(), x)  # nothing ran
    ctx.set_params(list_slots=0)
    got = []
    for c in (ctx, fluid.Context("cuda:0")):
        xd, vd = dev(x), dev(np.zeros_like(x))
        c.step(xd, vd, None, 3)
        got.append((xd.cpu().numpy(), vd.cpu().numpy()))
    np.testing.assert_array_equal(got[0][0].view(np.uint32), got[1][0].view(np.uint32))
    np.testing.assert_array_equal(got[0][1].view(np.uint32), got[1][1].view(np.uint32))


def test_million_particle_neighbor_checksums(fluid, oracle, scene):
    say("1,048,576-particle lattice: per-row neighbor checksums vs oracle")
    x = scene.lattice_1m()
    ctx = fluid.Context("cuda:0")
    nl = ctx.neighbor_query(dev(x))
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    go = nl.offsets().cpu().numpy()
    ge = nl.elements.cpu().numpy()
    np.testing.assert_array_equal(go, off)
    for a, b in zip(row_checksums(go, ge), row_checksums(off, el)):
        np.testing.assert_array_equal(a, b)


def test_million_particle_one_step(fluid, oracle, scene):
    say("1,048,576-particle lattice: one fused step vs oracle")
    x = scene.lattice_1m()
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    gpu = model.step_arrays(x, v, 1)
    assert_step_parity(gpu, *oracle_both(oracle, x, v, 1))


def test_million_particle_twenty_steps(fluid, oracle, scene):
    """The benchmark workload (1,048,576-particle lattice) over the benchmark's
    20 timed steps: mean density and kinetic energy within 0.5% of the oracle's
    FP32 run (the bar for long runs); the pointwise distance is reported."""
    say("1,048,576-particle lattice: 20 fused steps vs oracle (aggregates)")
    x = scene.lattice_1m()
    v = np.zeros_like(x)
    model = fluid.PositionBasedFluidModel(0.0125, "cuda:0", viscosity=False)
    xs, vs, rho = model.step_arrays(x, v, 20)
    xo, vo, ro, _ = oracle.step(oracle.make_params(), x, v, 20)
    ke, keo = (0.5 * (a.astype(np.float64) ** 2).sum(1).mean() for a in (vs, vo))
    rm, rmo = float(np.mean(rho, dtype=np.float64)), float(np.mean(ro, dtype=np.float64))
    say(f"mean rho {rm:.6g} vs {rmo:.6g}, KE {ke:.6g} vs {keo:.6g}; max |dx| {np.abs(xs - xo).max():.3e} m, "
        f"max |dv| {np.abs(vs - vo).max():.3e} m/s")
    assert abs(rm / rmo - 1) <= 5e-3 and abs(ke / keo - 1) <= 5e-3
