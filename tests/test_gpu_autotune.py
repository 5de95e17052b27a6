"""Launch shapes (fa_ctx_set_launch_shape) change only the speed.  A wave
cap leaves every result bit for bit (one wave scores a pose whatever the
wave).  The two neighbour-scan variants visit the same candidates in the
same emission order, so statuses, contact records and every term of a
scored pose are bit for bit the same; the clash pre-pass adds its wall terms
in lane order, which differs between the variants, so the wall of a clash
pose (and its cf) can differ in the last bits.  The tuner leaves its choice
on the context and bad shapes are rejected."""
import numpy as np
import pytest

from flexaidds_amd import abi, autotune, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.mark.parametrize("config", ["config0", "flex1", "config2"])
def test_every_shape_gives_the_same_results(hip, config):
    if config == "config0":
        s = synth.make_config0(n_poses=600)
    elif config == "flex1":
        s = synth.make_config1(pop=600)
    else:
        s = synth.make_config2(n_ligands=6, pop=100)
    with hip.Context() as c:
        c.upload_system(s)
        ref = [c.eval(li, s.genes[li]) for li in range(len(s.ligands))]
        ops = c.opcount()
        shapes = autotune.candidates(c)
        assert len(shapes) == 6 and {f for _, f in shapes} == {0, abi.FA_SHAPE_CELL_LISTS}
        for waves, flags in shapes:
            c.set_launch_shape(waves, flags)
            assert c.launch_shape() == (waves, flags)
            for li in range(len(s.ligands)):
                out = c.eval(li, s.genes[li])
                r = ref[li]
                if flags == 0:
                    assert out.tobytes() == r.tobytes(), (waves, flags, li)
                    continue
                assert np.array_equal(out["status"], r["status"])
                ok = r["status"] != 2
                assert out[ok].tobytes() == r[ok].tobytes(), (waves, flags, li)
                cl = ~ok
                for f in ("wal", "cf", "app"):
                    assert np.all(np.abs(out[f][cl] - r[f][cl]) <= 1e-12 * np.abs(r[f][cl])), f
            if len(s.ligands) == 1:
                assert c.opcount() == ops
        assert sum(int((r["status"] == 0).sum()) for r in ref) > 50


def test_tune_sets_the_fastest_shape(hip):
    s = synth.make_config1(pop=2000)
    g = np.ascontiguousarray(s.genes[0])
    with hip.Context() as c:
        c.upload_system(s)
        res = autotune.tune(c, lambda: c.eval(0, g), reps=2)
        assert len(res["table"]) == 6
        # the first shape (most waves first) within 3% of the fastest
        fastest = min(r["ms"] for r in res["table"])
        first = next(r for r in res["table"] if r["ms"] <= fastest * 1.03)
        assert (res["waves"], res["flags"], res["ms"]) == (first["waves"], first["flags"], first["ms"])
        assert c.launch_shape() == (res["waves"], res["flags"])
        full = c.launch_waves()
        assert 0 < res["waves"] <= full
        with pytest.raises(hip.FAError):
            c.set_launch_shape(-1, 0)
        with pytest.raises(hip.FAError):
            c.set_launch_shape(0, 8)
        c.set_launch_shape(10 * full, 0)                  # capped at the resident slots
        assert c.launch_shape() == (full, 0)
