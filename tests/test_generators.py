"""Host side of the seeded generators (no GPU): distributions per dictionary index and the
restatement of the device generator."""
import numpy as np

from cellato_amd.rules import iid_index_probs, device_random_values, iid_values, all_rules


def test_index_probabilities_sum_to_one():
    for r in all_rules():
        p = iid_index_probs(r.name)
        assert len(p) == len(r.states) and abs(sum(p) - 1) < 1e-9


def test_device_generator_frequencies():
    v = device_random_values("wire", 1024, 256, 7)
    freq = {s: (v == s).mean() for s in range(4)}
    assert abs(freq[0] - 0.80) < 0.01 and abs(freq[3] - 0.15) < 0.01
    assert abs(freq[1] - 0.03) < 0.005 and abs(freq[2] - 0.02) < 0.005
    bits = device_random_values("fluid", 512, 256, 1)
    for b in range(4):
        assert abs(((bits >> b) & 1).mean() - 0.25) < 0.01


def test_mix64_known_values():
    from cellato_amd.rules import mix64
    # splitmix64 of 0 and 1 (Steele, Lea, Flood 2014 reference outputs)
    assert int(mix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(mix64(np.uint64(1))) == 0x910A2DEC89025CC1


def test_iid_values_frequencies():
    v = iid_values("excitable", 512, 512, 3)
    assert abs((v == 0).mean() - 0.90) < 0.01 and abs((v == 1).mean() - 0.05) < 0.005
