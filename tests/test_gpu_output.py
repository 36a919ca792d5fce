"""Volume output (awp_amd/output.py): velbuffer_H samples every NTISKP steps,
asynchronous copies into the pinned host ring, one file per component every
WRITE_STEP samples.  Every record of every file must equal the wavefield of
its step sampled on the host by velbuffer's rule (u, v at level k, w one
level below), bit for bit; the copies of later samples must not disturb
earlier ones."""
import numpy as np
import pytest

from awp_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_volume_output_files_hold_the_sampled_wavefield(torch_gpu, tmp_path):
    from awp_amd.output import VolumeOutput
    from awp_amd.solver import Solver
    w = configs.anelastic(nx=40, ny=36, nz=32, nt=40, DH=200.0, DT=0.01)
    g = Solver(w.medium(), w.sources, w.receivers, nt_record=w.nt)
    d = g.dims
    nbg, ned, nskp = (1, 2, 0), (37, 33, 20), (3, 2, 4)
    out = VolumeOutput(d, nbg, ned, nskp, ntiskp=3, write_step=4,
                       basenames=("SX", "SY", "SZ"), directory=str(tmp_path))
    print(f"output dims={d} rec={out.rec}", flush=True)
    g.start()
    want = {}
    for _ in range(30):
        g.step()
        f = g.f
        out.sample(g.cur_step, f["u1"], f["v1"], f["w1"])
        if g.cur_step % 3 == 0:
            u, v, ww = (f[n].cpu().numpy() for n in ("u1", "v1", "w1"))
            i = 10 + np.arange(nbg[0], ned[0] + 1, nskp[0])
            j = 10 + np.arange(nbg[1], ned[1] + 1, nskp[1])
            k = d.nzt + 31 - np.arange(nbg[2], ned[2] + 1, nskp[2])
            I, J, K = np.meshgrid(i, j, k, indexing="ij")
            want[g.cur_step] = [a[I, J, K].transpose(2, 1, 0) for a in (u, v)] + \
                [ww[I, J, K - 1].transpose(2, 1, 0)]
    torch_gpu.cuda.synchronize()
    out.close()
    # samples at steps 3, 6, ..., 30: files after the 4th and 8th sample
    assert [p.rsplit("/", 1)[1] for p in out.files] == \
        ["SX_0_0000012", "SY_0_0000012", "SZ_0_0000012",
         "SX_0_0000024", "SY_0_0000024", "SZ_0_0000024"]
    for name in out.files:
        c = "XYZ".index(name.rsplit("/", 1)[1][1])
        step_end = int(name[-7:])
        data = out.read(name)
        for r in range(4):
            step = step_end - 3 * (3 - r)
            assert np.array_equal(data[r], want[step][c]), (name, r)
