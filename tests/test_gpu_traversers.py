"""The Python traverser mirrors (init / run / fetch_result, per-step callback) on the GPU."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_traverser_lifecycle():
    import cellato_amd as C
    v = C.iid_values("wire", 1024, 96, 3)
    for cls, T in [(C.SimpleTraverser, 0), (C.TemporalTraverser, 3)]:
        for layout in (0, 1):
            tr = cls()
            tr.init(C.PlaneGrid.from_standard("wire", v, layout), C.RunParams(automaton="wire", temporal_steps=T))
            tr.run(5)
            tr.run(6)   # continues at time step 5
            out = tr.fetch_result()
            np.testing.assert_array_equal(out.to_standard(), O.run_standard("wire", v, 11))
            assert out.get_checksum() == "-".join(map(str, O.checksum(O.run_standard("wire", v, 11))))


def test_callback_mode_sees_every_generation():
    import cellato_amd as C
    v = C.iid_values("traffic", 256, 32, 1)
    seen = []
    tr = C.TemporalTraverser()
    tr.init(C.PlaneGrid.from_standard("traffic", v), C.RunParams(temporal_steps=2))
    tr.run(4, callback=lambda i, g: seen.append((i, g.to_standard())))
    assert [i for i, _ in seen] == [0, 1, 2, 3, 4]
    for i, g in seen:
        np.testing.assert_array_equal(g, O.run_standard("traffic", v, i))


def test_bit_array_traverser_lifecycle():
    import cellato_amd as C
    for rule, wb, w, h in [("cyclic", 32, 600, 20), ("critters", 64, 640, 24), ("forest-fire", 64, 512, 9)]:
        v = C.iid_values(rule, w, h, 4)
        tr = C.BitArrayTraverser()
        tr.init(C.BitArrayGrid.from_standard(rule, v, wb), C.RunParams(automaton=rule, word_size=wb))
        tr.run(4)
        tr.run(7)   # continues at time step 4
        out = tr.fetch_result()
        np.testing.assert_array_equal(out.to_standard(), O.run_standard(rule, v, 11))
        assert out.x_size_physical() == w // (wb // C.rule(rule).planes)
