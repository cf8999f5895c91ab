"""On-device pack / unpack / checksum / seeded fill against the host restatements."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
RULES = ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic", "fluid", "critters", "cyclic"]


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("rule", RULES)
def test_pack_unpack_checksum(rule, layout):
    import torch
    from cellato_amd import DeviceGrid, grid as hg, rule as get_rule, iid_values
    r = get_rule(rule)
    for w, h in [(192, 40), (4160, 24), (64, 8)]:
        v = iid_values(rule, w, h, 3)
        d = DeviceGrid(r, w, h, layout)
        d.pack_values(torch.from_numpy(v).cuda())
        np.testing.assert_array_equal(d.download(), hg.pack(r.index_of(v), r.planes, layout))
        np.testing.assert_array_equal(d.unpack_values().cpu().numpy(), v)
        assert d.checksum_device() == O.checksum(v)


@pytest.mark.parametrize("rule", RULES)
def test_fill_random_matches_host_generator(rule):
    from cellato_amd import DeviceGrid, rule as get_rule
    from cellato_amd.rules import device_random_values
    for layout in (0, 1):
        d = DeviceGrid(get_rule(rule), 512, 48, layout)
        d.fill_random(12345)
        np.testing.assert_array_equal(d.to_plane_grid().to_standard(), device_random_values(rule, 512, 48, 12345))


def test_pack_rejects_unknown_state():
    import torch
    from cellato_amd import DeviceGrid, rule as get_rule
    d = DeviceGrid(get_rule("brian"), 64, 8, 0)
    with pytest.raises(ValueError):
        d.pack_values(torch.full((8, 64), 3, dtype=torch.int32, device="cuda"))
