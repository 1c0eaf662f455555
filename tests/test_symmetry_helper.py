"""The numpy restatement of the end-of-run symmetry figures
(helpers.symmetry, lulesh-util.cc:197-212), used on device-built runs, equals
the oracle's own (orc_symmetry) on runs where e still holds zeros (x/0 and
0/0 in the relative difference) and later."""
import pytest

from helpers import symmetry
from oracle_binding import OracleDomain


@pytest.mark.parametrize("s,its", [(8, 3), (8, 200), (12, 40)])
def test_symmetry_helper_equals_oracle(s, its):
    o = OracleDomain(s)
    assert o.run(its) == 0
    assert symmetry(o.real("e"), s) == o.symmetry()
