"""Known answers of the host arithmetic boundary, re-checked on the GPU host.

The sites and orientation rows come from the host's glibc rand() / random_r
and libstdc++ normal_distribution (CellularAutomataManager.cpp:153-213,
CellularAutomataManager_3D.cpp:838-850).  The GPU parity tests compare the
library with the oracle on the same host, so a different libc there would move
both sides together; these checks pin the absolute values measured on glibc
2.35 / GCC 11.4 on whatever host the GPU tests run.
"""
import numpy as np
import pytest

from test_oracle_kat import (test_glibc_rand_rank1_seed, test_glibc_rand_seed0,  # noqa: F401
                             test_glibc_rand_seed1_equals_seed0, test_glibc_type3_recurrence,
                             test_minstd_rand0)

pytestmark = pytest.mark.gpu


def test_host_layer_draws_of_the_nucleation_example(amca):
    """The reference's nucleation example (60^3 cells): 270 bulk draws, 269
    distinct sites; the first site cells, undercoolings and orientation angles
    as the host layer draws them."""
    from amca3d_amd import decks
    s = amca.Simulation(decks.nucleation_example_deck()).prepare()
    cell, dt, eu = s.initial_sites()
    assert s.info()["n_site_draws"] == 270
    assert len(cell) == 269
    assert cell[:5].tolist() == [4000, 5319, 5462, 6307, 6391]
    want_dt = [float.fromhex(x) for x in ("0x1.06b36d577e0e2p+2", "0x1.85751552ca6eep+1",
                                          "0x1.6c213c46795ddp+0")]
    assert np.array_equal(dt[:3], np.array(want_dt))
    want_eu = [float.fromhex(x) for x in ("0x1.65f90267f03a8p-1", "0x1.d9ebf7a4ab41dp+0",
                                          "0x1.cfcd51f92648dp-1", "0x1.dd16c6122a514p+1",
                                          "0x1.3b2f75723e792p+0", "0x1.d07f6b83c6456p-2")]
    assert np.array_equal(eu[:2].ravel(), np.array(want_eu))
