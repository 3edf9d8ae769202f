"""ps_timing_sample: events around one launch in `period` of each enabled
kernel; the count read back is the bracketed launches."""
import pytest

from particlesolver_amd import System, capi, scenes

pytestmark = pytest.mark.gpu


def test_sampled_timing_counts(gpu_available):
    sc = scenes.dam_break(8, 8, 8)
    with System(sc.params) as g:
        sc.load_into(g)
        g.timing(("lambda", "solve", "finalize"))
        g.timing_sample(7)
        g.timing_reset()
        g.step(sc.dt, 10)                  # 40 lambda and solve launches, 10 finalize
        t = g.timing_read()
        assert t["lambda"][1] == t["solve"][1] == 6      # launches 0, 7, ..., 35
        assert t["finalize"][1] == 2                     # launches 0 and 7
        assert t["lambda"][0] > 0.0
        g.timing_sample(1)
        g.timing_reset()
        g.step(sc.dt, 2)
        assert g.timing_read()["solve"][1] == 8
        with pytest.raises(capi.PsError):
            g.timing_sample(0)
