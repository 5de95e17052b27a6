"""Poses the device cannot score never pass as energies: they carry
cf = +inf, rank after every scored pose in GA selection, weigh 0 in the
ensemble, and are counted; invalid ligand indices and gene strides are
rejected; a resident system cannot be replaced under a live GA or pool."""
import ctypes as C

import numpy as np
import pytest

from flexaidds_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


def _thermo(oracle, e, T=300.0):
    e = np.ascontiguousarray(e, np.float64)
    out = abi.Thermo()
    oracle.fa_oracle_thermo(e.ctypes.data, None, len(e), T, C.byref(out))
    return out


def test_unscored_poses_never_survive_and_weigh_nothing(hip, oracle):
    # half of the cleft grid sits inside a 150-atom shell (over the device's
    # contact capacity: FA_CAPACITY), half far from every atom (scored)
    s = synth.make_shell_system(n_shell=150, n_poses=8)
    lim = synth.ga_limits(s.ligands[0])
    pop, new = 48, 96
    with hip.Context() as c:
        c.upload_system(s)
        c.reset_totals()
        ga = hip.GA(c, lim, population=pop, offspring=new, seed=99, temperature=300.0)
        _, _, r0 = ga.population()
        energies = [r0["cf"][:pop]]
        n_cap = int((r0["status"][:pop] == abi.FA_CAPACITY).sum())
        for gen in range(4):
            ga.run(1, graph=gen % 2 == 1)
            _, _, r = ga.population()
            energies.append(r["cf"][pop:pop + new])
            n_cap += int((r["status"][pop:pop + new] == abi.FA_CAPACITY).sum())
            assert np.all(r["status"][:pop] == abi.FA_OK), gen      # survivors are scored poses
            assert np.all(np.isfinite(r["cf"][:pop])), gen
        th, n = ga.thermo()
        tot = c.totals()
        ga.close()
    e = np.concatenate(energies)
    assert n_cap > 20 and np.isposinf(e).sum() == n_cap
    assert tot["capacity"] == n_cap
    assert n == len(e)
    ref = _thermo(oracle, e[np.isfinite(e)])
    assert np.isfinite(th["log_Z"])
    assert abs(th["log_Z"] - ref.log_Z) <= 1e-9 * abs(ref.log_Z) + 1e-12
    assert abs(th["mean_energy"] - ref.mean_energy) <= 1e-9 * abs(ref.mean_energy) + 1e-12


def test_bad_ligand_index_and_gene_stride(hip):
    import torch
    s = synth.make_config0(n_poses=16)
    G = s.ligands[0].n_genes
    with hip.Context() as c:
        c.upload_system(s)
        c.reset_totals()
        g = torch.from_numpy(np.ascontiguousarray(s.genes[0])).cuda()
        lig = torch.tensor([0, 7, -1, 0] * 4, dtype=torch.int32, device="cuda")
        out = torch.zeros(16 * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        c.eval_device(16, g.data_ptr(), G, lig.data_ptr(), out.data_ptr(), st)
        res = np.frombuffer(out.cpu().numpy().tobytes(), abi.CF_DTYPE)
        bad = np.array([0, 1, 1, 0] * 4, bool)
        assert np.all(res["status"][bad] == abi.FA_BAD_LIGAND)
        assert np.all(np.isposinf(res["cf"][bad]))
        assert np.all(res["status"][~bad] != abi.FA_BAD_LIGAND)
        assert c.totals()["bad_ligand"] == int(bad.sum())
        with pytest.raises(hip.FAError):  # rows narrower than the ligand's genes
            c.eval_device(16, g.data_ptr(), G - 1, 0, out.data_ptr(), st)


def test_upload_refused_while_a_ga_or_pool_lives(hip):
    s = synth.make_config0(n_poses=8)
    lim = synth.ga_limits(s.ligands[0])
    with hip.Context() as c:
        c.upload_system(s)
        ga = hip.GA(c, lim, population=16, offspring=16, seed=1)
        with pytest.raises(hip.FAError):
            c.upload_system(s)
        with pytest.raises(hip.FAError):
            c.set_phase_profiling(True)
        ga.run(1, graph=True)                                    # still usable
        ga.close()
        pool = hip.Pool(c, population=16, max_slots=2, hbm_budget=1 << 26)
        with pytest.raises(hip.FAError):
            c.upload_system(s)
        pool.close()
        c.upload_system(s)                                       # allowed again
        out = c.eval(0, s.genes[0])
        assert out.shape == (8,)


def test_undefined_faces_are_their_own_status(hip, oracle):
    """Config-4 GA poses (tests/golden/config4_undefined_face.npz, children
    of the blind dock's first generations) on which a face has no second
    point on its first edge's line: the reference's order_faces reads past
    the face's point list there (LIB/Vcontacts.cpp:794-805).  Device and
    oracle both report FA_UNDEFINED_FACE (cf = +inf), counted apart from
    device capacity overflows, of which there are none."""
    import os
    import oracle_lib as ol
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "config4_undefined_face.npz"))
    s, _ = synth.make_config4(n_ligands=64, pop=8)
    d = abi.Descriptors(s.receptor, s.ligands, s.params)
    with hip.Context() as c:
        c.upload_system(s)
        c.reset_totals()
        for ic, k in zip(fx["ic"], fx["lig"]):
            g = ic[None, :s.ligands[int(k)].n_genes]
            out = c.eval(int(k), g)
            with ol.gpu_mirror(oracle):
                ref = ol.oracle_eval(oracle, d, int(k), g)
            assert out["status"][0] == ref["status"][0] == abi.FA_UNDEFINED_FACE
            assert np.isinf(out["cf"][0]) and out["cf"][0] > 0
        t = c.totals()
    assert t["undefined_face"] == len(fx["lig"]) and t["capacity"] == 0
