"""Device diagnostics against the read-back state: node sampling equals the
moments at those nodes, per-moment means and the running mean flow equal
numpy's on read-back snapshots, turbulent kinetic energy equals its
definition on the read-back fields."""
import numpy as np
import pytest

import mrlbm

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("storage", [mrlbm.STORE_FP32, mrlbm.STORE_FP16])
def test_sample_and_means(storage):
    shape = (32, 24, 16)
    say("storage", storage)
    with mrlbm.Lattice(shape, q=19, tau=0.56, tile=(8, 8, 8), storage=storage) as lat:
        lat.init_taylor_green(0.05)
        lat.step(25)
        m = lat.moments()  # (10, N), x-fastest
        rng = np.random.default_rng(3)
        pts = np.stack([rng.integers(0, s, 50) for s in shape], axis=1)
        pts = np.vstack([pts, [[0, 0, 0], [31, 23, 15]]])
        got = lat.sample(pts)
        j = pts[:, 0] + shape[0] * (pts[:, 1] + shape[1] * pts[:, 2])
        assert np.array_equal(got, m[:, j].T)
        means = lat.moment_means()
        assert np.allclose(means, m.mean(axis=1), rtol=1e-12, atol=1e-15)
        with pytest.raises(mrlbm.MrlbmError):
            lat.sample([[32, 0, 0]])


def test_mean_flow_and_turbulent_ke():
    shape = (24, 16, 16)
    with mrlbm.Lattice(shape, q=19, tau=0.55, tile=(8, 8, 8)) as lat:
        lat.init_taylor_green(0.05)
        with pytest.raises(mrlbm.MrlbmError):
            lat.turbulent_ke()  # nothing accumulated yet
        snaps = []
        for _ in range(4):
            lat.step(7)
            lat.mean_flow_accumulate()
            snaps.append(lat.macroscopic())
        (rho_m, ux_m, uy_m, uz_m), count = lat.mean_flow()
        assert count == 4
        ref = [np.mean([s[k] for s in snaps], axis=0) for k in range(4)]
        for a, b in zip((rho_m, ux_m, uy_m, uz_m), ref):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-15)
        rho, ux, uy, uz = snaps[-1]
        tke_ref = np.mean(0.5 * rho * ((ux - ref[1]) ** 2 + (uy - ref[2]) ** 2 + (uz - ref[3]) ** 2))
        tke = lat.turbulent_ke()
        say("tke", tke, tke_ref)
        assert tke == pytest.approx(tke_ref, rel=1e-10, abs=1e-20)
        lat.mean_flow_reset()
        with pytest.raises(mrlbm.MrlbmError):
            lat.mean_flow()
