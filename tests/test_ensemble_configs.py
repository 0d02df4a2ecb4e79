"""The ensemble .ini generator writes configurations the oracle's parameter
reader accepts, with the advertised sizes, seeds and parameters."""
import numpy as np

from flusi_amd.ensemble_configs import write_ensemble
from oracle import flusi_oracle as fo


def test_default_ensemble_files(tmp_path):
    paths = write_ensemble(str(tmp_path))
    assert len(paths) == 48
    radii, sizes = [], []
    for p in paths:
        prm, _ = fo.get_params(p)
        sizes.append(prm.nx)
        assert prm.nx == prm.ny == prm.nz
        assert np.isclose(prm.xl, 2 * np.pi) and prm.eps == 1e-3 and prm.nu == 1e-2
        assert prm.iTimeMethodFluid == "RK2" and prm.dt_fixed > 0
        if "sphere" in p:
            radii.append(0.5 * prm.length)
    assert sorted(set(sizes)) == [64, 96, 128] and all(sizes.count(n) == 16 for n in (64, 96, 128))
    assert len(radii) == 24 and min(radii) >= 0.3 and max(radii) <= 0.8
    # the radii are the first draws of default_rng(5)
    assert np.allclose(radii, np.random.default_rng(5).uniform(0.3, 0.8, 24))
