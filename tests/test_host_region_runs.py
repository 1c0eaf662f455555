"""lh_region_runs: the region layout of CreateRegionIndexSets
(lulesh-init.cc:401-510) as runs, the input of the device-side set-up.
Expanded, the runs must give the oracle's regNumList (the oracle is pinned to
the reference's own region lists) and the host Domain's."""
import numpy as np
import pytest

import lulesh_amd as la
from oracle_binding import OracleDomain


@pytest.mark.parametrize("s,nr,balance", [(1, 1, 1), (4, 3, 1), (10, 11, 1), (30, 11, 1),
                                          (24, 21, 2), (17, 7, 0), (45, 21, 1)])
def test_runs_expand_to_the_region_numbers(s, nr, balance):
    reg, ln = la.region_runs(s, nr, balance)
    assert ln.min() >= 1 and int(ln.sum()) == s ** 3
    assert reg.min() >= 1 and reg.max() <= nr
    # the draw never repeats a region in consecutive runs
    if nr > 1:
        assert not np.any(reg[1:] == reg[:-1])
    o = OracleDomain(s, nr, balance, 1)
    assert np.array_equal(np.repeat(reg, ln), o.ints("regNumList"))
    h = la.HostDomain(s, nr, balance, 1)
    assert np.array_equal(np.repeat(reg, ln), h.ints("regNumList"))


def test_capacity_and_bad_arguments():
    lib = la.host_lib()
    n = lib.lh_region_runs(8, 11, 1, None, None, 0)
    assert n == len(la.region_runs(8, 11, 1)[0])
    assert lib.lh_region_runs(0, 11, 1, None, None, 0) == -1
    assert lib.lh_region_runs(8, 0, 1, None, None, 0) == -1
    assert lib.lh_region_runs(8, 11, 1, None, None, 5) == -1  # capacity without arrays
