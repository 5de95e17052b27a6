"""The allocation-free evaluation entry point against the ordinary device
entry point: identical results bit for bit, on two calls in a row (the
second with fewer poses, so the shared static scratch carries the first
call's contents), through the C ABI."""
import numpy as np
import pytest

from flexaidds_amd import abi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from flexaidds_amd import hip as h
    return h


@pytest.mark.parametrize("config", ["flex1", "config4"])
def test_noalloc_equals_device_entry_point(hip, config):
    import torch
    if config == "flex1":
        s = synth.make_config1(pop=1500)
        lig_of = np.zeros(1500, np.int32)
        genes = s.genes[0]
    else:  # uneven ligands on the 20,000-atom globule, widened grids included
        s, _ = synth.make_config4(n_ligands=3, pop=300)
        G = max(l.n_genes for l in s.ligands)
        genes = np.zeros((900, G))
        lig_of = np.repeat(np.arange(3, dtype=np.int32), 300)
        for k, g in enumerate(s.genes):
            genes[300 * k:300 * (k + 1), :g.shape[1]] = g
    n = len(genes)
    G = genes.shape[1]
    assert hip._lib.fa_noalloc_scratch_bytes() == 128 << 20
    with hip.Context() as c:
        c.upload_system(s)
        dg = torch.from_numpy(np.ascontiguousarray(genes)).cuda()
        dl = torch.from_numpy(lig_of).cuda()
        st = torch.cuda.current_stream().cuda_stream
        for call, m in enumerate((n, n // 3)):   # second call smaller
            a = torch.zeros(m * abi.CF_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            b = torch.full((m * abi.CF_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device="cuda")
            c.eval_device(m, dg.data_ptr(), G, dl.data_ptr(), a.data_ptr(), st)
            c.eval_device_noalloc(m, dg.data_ptr(), G, dl.data_ptr(), b.data_ptr(), st)
            torch.cuda.synchronize()
            ra = np.frombuffer(a.cpu().numpy().tobytes(), abi.CF_DTYPE)
            rb = np.frombuffer(b.cpu().numpy().tobytes(), abi.CF_DTYPE)
            assert (ra["status"] == 0).sum() > 0, call
            assert ra.tobytes() == rb.tobytes(), call


def test_noalloc_rejects_bad_calls(hip):
    s = synth.make_config0(n_poses=8)
    with hip.Context() as c:
        with pytest.raises(hip.FAError):                 # nothing uploaded
            c.eval_device_noalloc(1, 1, 10, 0, 1)
        c.upload_system(s)
        with pytest.raises(hip.FAError):                 # rows narrower than the genes
            c.eval_device_noalloc(1, 1, s.ligands[0].n_genes - 1, 0, 1)
