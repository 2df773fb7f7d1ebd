"""Shared test set-up: markers, paths, seeded datasets."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def oracle():
    from oracle import lko
    lko.lib()
    return lko


@pytest.fixture(scope="session")
def datasets(tmp_path_factory):
    """Small seeded problems covering the path's cases (case-only, case-control with
    covariates, row/column masks, several 1920-bit blocks, B < 32, gamma_type 1,
    kappa_squared > 0, a ragged tiny case)."""
    from lakcluster_c_amd import gen
    root = str(tmp_path_factory.mktemp("lkc_data"))
    specs = {
        "caseonly": dict(nA=200, ncols=700, nZ=0, T=1, gamma=0.05, B=32, seed=1),
        "casectrl": dict(nA=150, ncols=900, nZ=170, T=3, gamma=0.05, B=32, seed=2),
        "masked": dict(nA=130, ncols=1100, nZ=90, T=3, gamma=0.08, B=32, seed=3,
                       row_mask_frac=0.2, col_mask_frac=0.1),
        "blocks": dict(nA=96, ncols=4000, nZ=64, T=2, gamma=0.1, B=32, seed=4),
        "b16_g1": dict(nA=120, ncols=600, nZ=100, T=3, gamma=0.05, B=16, seed=5, gamma_type=1),
        "kappa": dict(nA=110, ncols=500, nZ=80, T=3, gamma=0.06, B=24, seed=6, kappa_squared=0.5),
        "tiny": dict(nA=5, ncols=40, nZ=3, T=2, gamma=0.2, B=32, seed=7, planted=()),
        "wide_z": dict(nA=60, ncols=800, nZ=200, T=3, gamma=0.04, B=32, seed=8),
    }
    out = {}
    for name, kw in specs.items():
        ds = gen.make_dataset(root, name, **kw)
        out[name] = ds
    return out
