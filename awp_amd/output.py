"""Volume output of pmcl3d on one rank (reference src/awp/pmcl3d.c:2252-2360):
every NTISKP steps velbuffer_H samples u, v, w (and the plastic strain with
NVE = 3) on a strided sub-volume into device buffers; the buffers go to a
host ring of WRITE_STEP records, and every WRITE_STEP records the ring is
written to one file per component, named "<base>_<grid>_<step:07d>" and
holding WRITE_STEP x rec_nzt x rec_nyt x rec_nxt native floats (the NOBGIO
write path with the single rank's file view).

The device-to-host copies run on a side stream and overlap the following
steps.  Two sets of device buffers alternate, so the next sample never waits
for the previous copy unless both sets are still in flight.  Two pinned host
rings alternate too: a full ring is written to its files by a background
thread while the next one fills, so file writes do not stall the launches.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import _lib


class VolumeOutput:
    def __init__(self, dims, nbg, ned, nskp, ntiskp: int, write_step: int, basenames,
                 grid: int = 0, nve: int = 1, follow_bathy: int = 0, bathy=None,
                 directory: str = "."):
        """nbg, ned, nskp: (x, y, z) first, last and stride of the sampled
        cells, in the reference's 0-based interior coordinates (z counted down
        from the free surface).  basenames: 3 names (4 with NVE = 3)."""
        self.dims = dims
        self.nbg, self.ned, self.nskp = tuple(nbg), tuple(ned), tuple(nskp)
        if min(self.nskp) < 1 or ntiskp < 1 or write_step < 1:
            raise ValueError("strides, NTISKP and WRITE_STEP must be positive")
        self.rec = tuple((e - b) // s + 1 for b, e, s in zip(self.nbg, self.ned, self.nskp))
        if min(self.rec) < 1:
            raise ValueError(f"empty output window {self.nbg}..{self.ned}")
        self.ntiskp, self.write_step = int(ntiskp), int(write_step)
        self.grid, self.nve, self.follow = int(grid), int(nve), int(follow_bathy)
        ncomp = 4 if nve == 3 else 3
        if len(basenames) != ncomp:
            raise ValueError(f"NVE={nve} writes {ncomp} components")
        self.basenames = list(basenames)
        self.directory = directory
        self.bathy = bathy
        n = self.rec[0] * self.rec[1] * self.rec[2]
        self.nrec = n
        self.dev = [[torch.zeros(n, device="cuda") for _ in range(4)] for _ in range(2)]
        self.hosts = [torch.empty((ncomp, self.write_step, n), dtype=torch.float32,
                                  pin_memory=True) for _ in range(2)]
        self.ring = 0
        self.pending = [None, None]  # file-write future of each host ring
        self.writer = ThreadPoolExecutor(max_workers=1)
        self.side = torch.cuda.Stream()
        self.copied = [None, None]   # event: the copy out of buffer set s has finished
        self.turn = 0
        self.files: list[str] = []

    def _slot(self, cur_step: int) -> int:
        # pmcl3d.c:2281: ((cur_step/NTISKP + WRITE_STEP - 1) % WRITE_STEP)
        return (cur_step // self.ntiskp + self.write_step - 1) % self.write_step

    def sample(self, cur_step: int, u1, v1, w1, neta=None) -> None:
        """Call after step cur_step (the reference's loop position)."""
        if cur_step % self.ntiskp:
            return
        lib = _lib.load()
        compute = torch.cuda.current_stream()
        s = self.turn
        self.turn ^= 1
        if self.copied[s] is not None:
            compute.wait_event(self.copied[s])
        bx, by, bz, beta = self.dev[s]
        if self.nve == 3 and neta is None:
            raise ValueError("NVE = 3 writes the plastic strain: pass neta")
        eta = neta if neta is not None else u1   # read only when NVE = 3
        nb, ne, ns = self.nbg, self.ned, self.nskp
        lib.velbuffer_H(u1.data_ptr(), v1.data_ptr(), w1.data_ptr(), eta.data_ptr(),
                        bx.data_ptr(), by.data_ptr(), bz.data_ptr(), beta.data_ptr(), self.nve,
                        nb[0], ne[0], ns[0], nb[1], ne[1], ns[1], nb[2], ne[2], ns[2],
                        self.rec[0], self.rec[1], self.rec[2], compute.cuda_stream, self.follow,
                        None if self.bathy is None else self.bathy.data_ptr(), self.grid)
        _lib.check()
        ready = torch.cuda.Event()
        ready.record(compute)
        slot = self._slot(cur_step)
        host = self.hosts[self.ring]
        if self.pending[self.ring] is not None:
            self.pending[self.ring].result()   # its previous contents are on disk
            self.pending[self.ring] = None
        with torch.cuda.stream(self.side):
            self.side.wait_event(ready)
            for c in range(host.shape[0]):
                host[c, slot].copy_(self.dev[s][c], non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.side)
        self.copied[s] = done
        if (cur_step // self.ntiskp) % self.write_step == 0:
            self.side.synchronize()
            names = [os.path.join(self.directory, f"{base}_{self.grid:1d}_{cur_step:07d}")
                     for base in self.basenames]
            self.files.extend(names)
            self.pending[self.ring] = self.writer.submit(self._write, host, names)
            self.ring ^= 1

    @staticmethod
    def _write(host, names) -> None:
        for c, name in enumerate(names):
            host[c].numpy().tofile(name)

    def close(self) -> None:
        """Wait for the files still being written."""
        for k, f in enumerate(self.pending):
            if f is not None:
                f.result()
                self.pending[k] = None
        self.writer.shutdown(wait=True)

    def read(self, name: str) -> np.ndarray:
        """A written file as (WRITE_STEP, rec_nzt, rec_nyt, rec_nxt)."""
        return np.fromfile(name, dtype=np.float32).reshape(
            self.write_step, self.rec[2], self.rec[1], self.rec[0])
