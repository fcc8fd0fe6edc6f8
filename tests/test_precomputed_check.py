"""Host-side check of precomputed operator tables (gnufft_check_precomputed):
valid tables in the reference's std::sort order and in the stable order
pass; each kind of malformed table is rejected with INVALID_ARGUMENT before
anything reaches the device (CPU only: the check makes no HIP call)."""

import numpy as np
import pytest

from gpunufft_amd import workloads as W

INVALID = 1


@pytest.fixture(scope="module")
def capi():
    from gpunufft_amd import capi as C
    C.load()
    return C


GEOMS = [
    ("radial2d", lambda: W.radial_2d(32, 64), (32, 32), 2.0, 3, 8),
    ("koosh3d", lambda: W.koosh_3d(64, 24), (16, 16, 16), 1.5, 3, 8),
    ("rand2d_sw16", lambda: np.random.default_rng(3).uniform(-0.5, 0.5, (2, 900)).astype(np.float32),
     (48, 40), 2.0, 3, 16),
]


@pytest.mark.parametrize("geom", GEOMS, ids=[g[0] for g in GEOMS])
@pytest.mark.parametrize("order", ["std", "stable"])
def test_valid_tables_pass(oracle, capi, geom, order):
    _, tf, img, osf, kw, sw = geom
    traj = tf()
    pre = oracle.precompute(traj, img, osf, kw, sw, order=order)
    rc = capi.check_precomputed(img, sw, osf, pre["traj_sorted"], pre["data_indices"],
                                pre["sector_data_count"], pre["sector_centers"])
    assert rc == 0, capi.load().gnufft_last_error()


def _tables(oracle):
    traj = W.radial_2d(32, 64)
    img, osf, sw = (32, 32), 2.0, 8
    pre = oracle.precompute(traj, img, osf, 3, sw, order="std")
    return img, osf, sw, pre


def _run(capi, img, osf, sw, t):
    return capi.check_precomputed(img, sw, osf, t["traj_sorted"], t["data_indices"],
                                  t["sector_data_count"], t["sector_centers"])


def _copy(pre):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in pre.items()}


@pytest.mark.parametrize("defect", ["index_out_of_range", "index_duplicate", "count_first_nonzero",
                                    "count_decreasing", "count_total", "center_layout",
                                    "sample_in_wrong_sector"])
def test_malformed_tables_rejected(oracle, capi, defect):
    img, osf, sw, pre = _tables(oracle)
    t = _copy(pre)
    M = t["data_indices"].size
    sdc = t["sector_data_count"]
    if defect == "index_out_of_range":
        t["data_indices"][5] = M
    elif defect == "index_duplicate":
        t["data_indices"][7] = t["data_indices"][8]
    elif defect == "count_first_nonzero":
        t["sector_data_count"][0] = 1
    elif defect == "count_decreasing":
        s = int(np.nonzero(np.diff(sdc))[0][3])  # a non-empty sector
        t["sector_data_count"][s + 1] = sdc[s] - 1 if sdc[s] > 0 else 0
        if t["sector_data_count"][s + 1] >= t["sector_data_count"][s]:
            t["sector_data_count"][s] = t["sector_data_count"][s + 1] + 1
    elif defect == "count_total":
        t["sector_data_count"][-1] = M - 1
    elif defect == "center_layout":
        t["sector_centers"][3] += 1
    elif defect == "sample_in_wrong_sector":
        # swap the coordinates of the first samples of two different sectors
        nz = np.nonzero(np.diff(sdc))[0]
        a, b = int(sdc[nz[0]]), int(sdc[nz[-1]])
        tr = t["traj_sorted"]
        tr[:, [a, b]] = tr[:, [b, a]]
    rc = _run(capi, img, osf, sw, t)
    assert rc == INVALID, defect
    assert capi.load().gnufft_last_error().startswith(b"check_precomputed")
