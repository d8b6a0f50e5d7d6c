"""The 64-instance batch workload: the frozen (b0, psi0) table is the stated
draw, and instance i gets the stated top nside, frequency list and seed."""
import numpy as np

from hammurabix_amd import Param, configs


def test_frozen_table_is_the_stated_draw():
    rng = np.random.default_rng(20251202)
    b0 = rng.uniform(4, 8, 64)
    psi0 = rng.uniform(20, 35, 64)
    table = np.array(configs.BATCH_B0_PSI0)
    assert table.shape == (64, 2)
    np.testing.assert_allclose(table[:, 0], b0, atol=5e-5)
    np.testing.assert_allclose(table[:, 1], psi0, atol=5e-5)


def test_instance_assignment():
    for i in (0, 1, 2, 3, 5, 63):
        par = Param(text=configs.batch_instance(i))
        top = (64, 128, 256)[i % 3]
        g = par.grid_obs
        assert g.nside_shell == [top // 4, top // 2, top]
        assert g.nside_sync == [top] * (1 + i % 4)
        assert g.sim_sync_freq == [f * 1e9 for f in (0.408, 1.4, 23.0, 30.0)[:1 + i % 4]]
        assert par.brnd_seed == i + 1
        d = par.los_desc()
        b0, psi0 = configs.BATCH_B0_PSI0[i]
        assert abs(d.breg[0] - b0 * 1e-6) < 1e-15
        assert abs(d.breg[1] - np.deg2rad(psi0)) < 1e-12
