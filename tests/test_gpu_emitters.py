"""ParticleFluid with emitters on the GPU (ParticleFluid.cpp:33-124): the
particle count grows frame by frame, and every frame equals a direct
PositionBasedFluidModel step of the concatenated input — the emitters'
segments in the order added, each one's carried particles then its new ones.
Every tenth frame that direct step is also held against the oracle's step
(with ImplicitViscosity, as the model runs it) on the same concatenated
input, to the one-step parity bars."""
from importlib import import_module

import numpy as np
import pytest
import torch


@pytest.fixture(scope="module")
def em(pkg):
    import __graft_entry__ as G
    return import_module(G.PKG_NAME + ".emitters")


@pytest.mark.gpu
def test_fluid_with_emitters_equals_direct_steps(fluid, em, oracle):
    from parity_util import assert_step_parity, oracle_both, say
    dev = torch.device("cuda:0")
    pf = fluid.ParticleFluid()
    direct = fluid.PositionBasedFluidModel()
    e1 = em.ParticleEmitterSquare(device=dev, seed=5)
    e1.varLocation((0.3, 0.8, 0.3))
    e2 = em.ParticleEmitterRound(device=dev, seed=6)
    e2.varLocation((0.7, 0.8, 0.6))
    e2.varRotation((30, 0, 20))
    e2.sampling_distance = np.float32(0.004)
    pf.addParticleEmitter(e1)
    pf.addParticleEmitter(e2)
    pf.resetStatus()
    n1 = n2 = 0
    counts = []
    for frame in range(60):
        xb = pf.currentPosition.getValue().clone()
        vb = pf.currentVelocity.getValue().clone()
        print(f"frame {frame} n={xb.shape[0]}", flush=True)
        stepped = pf.advance(0.001)
        g = [torch.from_numpy(e.gen_pos).to(dev) for e in (e1, e2)]
        gv = [torch.from_numpy(e.gen_vel).to(dev) for e in (e1, e2)]
        x_in = torch.cat([xb[:n1], g[0], xb[n1:n1 + n2], g[1]])
        v_in = torch.cat([vb[:n1], gv[0], vb[n1:n1 + n2], gv[1]])
        n1 += len(e1.gen_pos)
        n2 += len(e2.gen_pos)
        assert stepped == n1 + n2 == pf.currentPosition.getElementCount()
        counts.append(stepped)
        if stepped == 0:
            continue
        xi, vi = x_in.cpu().numpy(), v_in.cpu().numpy()  # the model steps x_in and v_in in place
        direct.m_position.setValue(x_in.contiguous())
        direct.m_velocity.setValue(v_in.contiguous())
        direct.m_forceDensity.setValue(torch.zeros_like(x_in))
        direct.step(0.001)
        torch.cuda.synchronize()
        assert torch.equal(pf.currentPosition.getValue(), direct.m_position.getValue()), frame
        assert torch.equal(pf.currentVelocity.getValue(), direct.m_velocity.getValue()), frame
        assert torch.equal(pf.model.m_density.getValue(), direct.m_density.getValue()), frame
        if frame % 10 == 9:
            say(f"emitter frame {frame}: {stepped} particles vs oracle")
            got = tuple(a.cpu().numpy() for a in (direct.m_position.getValue(), direct.m_velocity.getValue(),
                                                  direct.m_density.getValue()))
            assert_step_parity(got, *oracle_both(oracle, xi, vi, 1, flags=1))
        assert e1.currentPosition.getElementCount() == n1 and e2.currentPosition.getElementCount() == n2
    assert counts[-1] > counts[len(counts) // 2] > 0
    assert torch.isfinite(pf.currentPosition.getValue()).all()
    assert pf.model.ctx.stats()["out_of_grid"] == 0
