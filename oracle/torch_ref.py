"""Vectorised float64 restatement of one grid's NVE = 1 time step in torch
(CPU), differentiable in the medium arrays the kernels read: lam (1/lambda),
mu (1/mu), d1 (rho), with the free-surface coefficient lam_mu rebuilt from
lam and mu as pmcl3d builds it (src/awp/pmcl3d.c:941-952).

TEST INFRASTRUCTURE ONLY: reverse-mode differentiation of this forward is the
CPU reference gradient that the hand-written GPU adjoint is checked against.
It follows the oracle's arithmetic (oracle/awp_oracle.cpp: dvelcx, fs_ghosts,
dstrqc, addsrc, fstr) term by term, without its float32 roundings; a test
checks its traces against the FP64 oracle.
"""
from __future__ import annotations

import numpy as np
import torch

NGSL, ALIGN = 8, 32
C1, C2 = 9.0 / 8.0, -1.0 / 24.0


class TorchForward:
    def __init__(self, medium, sources, receivers, nt: int):
        d = medium.dims
        self.d = d
        self.nt = nt
        self.DH, self.DT = float(medium.DH), float(medium.DT)
        self.dth = self.DT / self.DH
        self.dh1 = 1.0 / self.DH
        t64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
        self.qp, self.qs = t64(medium.qp), t64(medium.qs)
        self.vx1, self.vx2, self.wwo = t64(medium.vx1), t64(medium.vx2), t64(medium.wwo)
        self.ww = torch.from_numpy(np.ascontiguousarray(medium.ww, dtype=np.int64))
        self.coeff = t64(medium.coeff)
        self.dcx, self.dcy, self.dcz = t64(medium.dcrjx), t64(medium.dcrjy), t64(medium.dcrjz)
        self.src = None
        if sources is not None and sources.npsrc > 0:
            ps = sources.psrc(d)
            self.src = [(int(x + 1 + NGSL), int(y + 1 + NGSL), int(z + ALIGN - 1))
                        for x, y, z in ps]
            self.rates = sources.rates  # (npsrc, nst, 6): xx yy zz xz yz xy
        idx = receivers.array_index(d)
        self.rec = [tuple(int(v) for v in r) for r in idx]
        self.I = slice(2 + NGSL, d.nxt + NGSL + 2)
        self.J = slice(2 + NGSL, d.nyt + NGSL + 2)
        self.top = d.nzt + ALIGN - 1

    # shifted views over a box
    @staticmethod
    def _s(A, box, a=0, b=0, c=0):
        (i0, i1), (j0, j1), (k0, k1) = box
        return A[i0 + a:i1 + a, j0 + b:j1 + b, k0 + c:k1 + c]

    def _qweight(self, q):
        """Production Q weight (dstrqc): quadratic fit below the branch,
        2 wwo q above, divided by wwo, per cell of the stress box."""
        ww = self._s(self.ww, self.sbox)
        wwo = self._s(self.wwo, self.sbox)
        ca = self.coeff[ww * 2 - 2]
        cb = self.coeff[ww * 2 - 1]
        quad = ca * (2 * q) * (2 * q) + cb * (2 * q)
        lin = 2.0 * wwo * q
        w = torch.where(1.0 / (q * 2.0) <= 200.0, quad, lin)
        return w / wwo

    def run(self, lam, mu, d1):
        """Traces (nt, nrec, 3) of a run from rest."""
        d = self.d
        sh = d.shape
        z = lambda: torch.zeros(sh, dtype=torch.float64)
        f = {n: z() for n in ("u1", "v1", "w1", "xx", "yy", "zz", "xy", "xz", "yz",
                               "r1", "r2", "r3", "r4", "r5", "r6")}
        I, J, top = self.I, self.J, self.top
        vbox = ((I.start, I.stop), (J.start, J.stop), (ALIGN, d.nzt + ALIGN))
        self.sbox = sbox = ((I.start, I.stop), (J.start, J.stop), (ALIGN + 3, top + 1))
        s = self._s
        dcrj_v = (self.dcx[I][:, None, None] * self.dcy[J][None, :, None] *
                  self.dcz[ALIGN:d.nzt + ALIGN][None, None, :])
        dcrj_s = (self.dcx[I][:, None, None] * self.dcy[J][None, :, None] *
                  self.dcz[ALIGN + 3:top + 1][None, None, :])
        # velocity coefficients
        fd1 = self.dth / (0.25 * (s(d1, vbox) + s(d1, vbox, 0, -1) + s(d1, vbox, 0, 0, -1) +
                                  s(d1, vbox, 0, -1, -1)))
        fd2 = self.dth / (0.25 * (s(d1, vbox) + s(d1, vbox, 1) + s(d1, vbox, 0, 0, -1) +
                                  s(d1, vbox, 1, 0, -1)))
        fd3 = self.dth / (0.25 * (s(d1, vbox) + s(d1, vbox, 1) + s(d1, vbox, 0, -1) +
                                  s(d1, vbox, 1, -1)))
        # stress coefficients
        sl8 = sum(s(lam, sbox, a, -b, -c) for a in (0, 1) for b in (0, 1) for c in (0, 1))
        sm8 = sum(s(mu, sbox, a, -b, -c) for a in (0, 1) for b in (0, 1) for c in (0, 1))
        xm = 16.0 / sm8
        xl = 8.0 / sl8 + xm
        xmu1 = 2.0 / (s(mu, sbox) + s(mu, sbox, 0, 0, -1))
        xmu2 = 2.0 / (s(mu, sbox) + s(mu, sbox, 0, -1))
        xmu3 = 2.0 / (s(mu, sbox) + s(mu, sbox, 1))
        sp8 = sum(s(self.qp, sbox, a, -b, -c) for a in (0, 1) for b in (0, 1) for c in (0, 1))
        ss8 = sum(s(self.qs, sbox, a, -b, -c) for a in (0, 1) for b in (0, 1) for c in (0, 1))
        qpaw = self._qweight(0.0625 * sp8)
        hw = self._qweight(0.0625 * ss8)
        h1w = self._qweight(0.25 * (s(self.qs, sbox) + s(self.qs, sbox, 0, 0, -1)))
        h2w = self._qweight(0.25 * (s(self.qs, sbox) + s(self.qs, sbox, 0, -1)))
        h3w = self._qweight(0.25 * (s(self.qs, sbox) + s(self.qs, sbox, 1)))
        vx1 = s(self.vx1, sbox)
        vx2 = s(self.vx2, sbox)
        wwo = s(self.wwo, sbox)
        DT, dth, dh1 = self.DT, self.dth, self.dh1
        h = -xm * hw * dh1 * vx1
        h1 = -xmu1 * h1w * dh1 * vx1
        h2 = -xmu2 * h2w * dh1 * vx1
        h3 = -xmu3 * h3w * dh1 * vx1
        qpa = -qpaw * xl * dh1 * vx1
        XM = xm * dth + DT * h
        XMU1 = xmu1 * dth + DT * h1
        XMU2 = xmu2 * dth + DT * h2
        XMU3 = xmu3 * dth + DT * h3
        XL = xl * dth
        vx1p = DT * (1 + vx2 * vx1)
        # free-surface coefficient from lam, mu at the surface
        t_xl = 1.0 / lam[I, J, top]
        lam_mu = t_xl / (2.0 / mu[I, J, top] + t_xl)
        nx, ny = d.nxt, d.nyt
        gi = torch.arange(I.start, I.stop) - NGSL - 1
        gj = torch.arange(J.start, J.stop) - NGSL - 1
        c_x = (gi < nx).to(torch.float64)[:, None]
        c_y = (gj > 1).to(torch.float64)[None, :]
        traces = []
        for n in range(1, self.nt + 1):
            if n == 1 and self.src is not None:
                f = self._addsrc(f, 1)
            u, v, w = f["u1"], f["v1"], f["w1"]
            xx, yy, zz, xy, xz, yz = (f[k] for k in ("xx", "yy", "zz", "xy", "xz", "yz"))
            b = vbox
            Xu = (C1 * (s(xx, b) - s(xx, b, -1)) + C2 * (s(xx, b, 1) - s(xx, b, -2)) +
                  C1 * (s(xy, b) - s(xy, b, 0, -1)) + C2 * (s(xy, b, 0, 1) - s(xy, b, 0, -2)) +
                  C1 * (s(xz, b) - s(xz, b, 0, 0, -1)) + C2 * (s(xz, b, 0, 0, 1) - s(xz, b, 0, 0, -2)))
            Xv = (C1 * (s(xy, b, 1) - s(xy, b)) + C2 * (s(xy, b, 2) - s(xy, b, -1)) +
                  C1 * (s(yy, b, 0, 1) - s(yy, b)) + C2 * (s(yy, b, 0, 2) - s(yy, b, 0, -1)) +
                  C1 * (s(yz, b) - s(yz, b, 0, 0, -1)) + C2 * (s(yz, b, 0, 0, 1) - s(yz, b, 0, 0, -2)))
            Xw = (C1 * (s(xz, b, 1) - s(xz, b)) + C2 * (s(xz, b, 2) - s(xz, b, -1)) +
                  C1 * (s(yz, b) - s(yz, b, 0, -1)) + C2 * (s(yz, b, 0, 1) - s(yz, b, 0, -2)) +
                  C1 * (s(zz, b, 0, 0, 1) - s(zz, b)) + C2 * (s(zz, b, 0, 0, 2) - s(zz, b, 0, 0, -1)))
            un, vn, wn = u.clone(), v.clone(), w.clone()
            un[I, J, ALIGN:d.nzt + ALIGN] = (s(u, b) + fd1 * Xu) * dcrj_v
            vn[I, J, ALIGN:d.nzt + ALIGN] = (s(v, b) + fd2 * Xv) * dcrj_v
            wn[I, J, ALIGN:d.nzt + ALIGN] = (s(w, b) + fd3 * Xw) * dcrj_v
            u, v, w = un, vn, wn
            # free-surface ghosts at top+1
            t = top
            ip = slice(I.start + 1, I.stop + 1)
            im = slice(I.start - 1, I.stop - 1)
            jp = slice(J.start + 1, J.stop + 1)
            jm = slice(J.start - 1, J.stop - 1)
            ug = u[I, J, t] - (w[I, J, t] - w[im, J, t])
            vg = v[I, J, t] - (w[I, jp, t] - w[I, J, t])
            vs1 = c_x * (u[ip, J, t] - (w[ip, J, t] - w[I, J, t]))
            vs2 = c_y * (v[I, jm, t] - (w[I, J, t] - w[I, jm, t]))
            wg = w[I, J, t - 1] - lam_mu * ((vs1 - ug) + (u[ip, J, t] - u[I, J, t]) +
                                           (vg - vs2) + (v[I, J, t] - v[I, jm, t]))
            u, v, w = u.clone(), v.clone(), w.clone()
            u[I, J, t + 1] = ug
            v[I, J, t + 1] = vg
            w[I, J, t + 1] = wg
            # stress + memory variables
            b = sbox
            N1 = C1 * (s(u, b, 1) - s(u, b)) + C2 * (s(u, b, 2) - s(u, b, -1))
            N2 = C1 * (s(v, b) - s(v, b, 0, -1)) + C2 * (s(v, b, 0, 1) - s(v, b, 0, -2))
            N3 = C1 * (s(w, b) - s(w, b, 0, 0, -1)) + C2 * (s(w, b, 0, 0, 1) - s(w, b, 0, 0, -2))
            E4 = (C1 * (s(u, b, 0, 1) - s(u, b)) + C2 * (s(u, b, 0, 2) - s(u, b, 0, -1)) +
                  C1 * (s(v, b) - s(v, b, -1)) + C2 * (s(v, b, 1) - s(v, b, -2)))
            E5 = (C1 * (s(u, b, 0, 0, 1) - s(u, b)) + C2 * (s(u, b, 0, 0, 2) - s(u, b, 0, 0, -1)) +
                  C1 * (s(w, b) - s(w, b, -1)) + C2 * (s(w, b, 1) - s(w, b, -2)))
            E6 = (C1 * (s(v, b, 0, 0, 1) - s(v, b)) + C2 * (s(v, b, 0, 0, 2) - s(v, b, 0, 0, -1)) +
                  C1 * (s(w, b, 0, 1) - s(w, b)) + C2 * (s(w, b, 0, 2) - s(w, b, 0, -1)))
            S = N1 + N2 + N3
            tmp = XL * S + DT * (qpa * S)
            a1 = qpa * S
            new = {}

            def normal(sig, r, O, twice):
                frt = -h * O + a1
                tt = s(sig, b) + tmp - XM * O + vx1p * s(r, b)
                if twice:
                    tt = tt * dcrj_s
                rn = vx2 * s(r, b) + wwo * frt
                frt2 = frt * (wwo - 1) + vx2 * s(r, b) * (1 - vx1)
                return (tt + DT * frt2) * dcrj_s, rn

            def shear(sig, r, E, H, XMU):
                frt = H * E
                tt = s(sig, b) + XMU * E + vx1p * s(r, b)
                rn = vx2 * s(r, b) + wwo * frt
                frt2 = frt * (wwo - 1) + vx2 * s(r, b) * (1 - vx1)
                return (tt + DT * frt2) * dcrj_s, rn

            new["xx"], new["r1"] = normal(xx, f["r1"], N2 + N3, False)
            new["yy"], new["r2"] = normal(yy, f["r2"], N1 + N3, True)
            new["zz"], new["r3"] = normal(zz, f["r3"], N1 + N2, True)
            new["xy"], new["r4"] = shear(xy, f["r4"], E4, h1, XMU1)
            new["xz"], new["r5"] = shear(xz, f["r5"], E5, h2, XMU2)
            new["yz"], new["r6"] = shear(yz, f["r6"], E6, h3, XMU3)
            g = {"u1": u, "v1": v, "w1": w}
            for k_, val in new.items():
                a_ = f[k_].clone()
                a_[I, J, ALIGN + 3:top + 1] = val
                g[k_] = a_
            f = g
            if self.src is not None and n < self.rates.shape[1]:
                f = self._addsrc(f, n + 1)
            # stress imaging above the surface
            zz, xz, yz = f["zz"].clone(), f["xz"].clone(), f["yz"].clone()
            zz[I, J, t + 1] = -zz[I, J, t]
            zz[I, J, t + 2] = -zz[I, J, t - 1]
            xz[I, J, t + 1] = -xz[I, J, t - 1]
            xz[I, J, t + 2] = -xz[I, J, t - 2]
            yz[I, J, t + 1] = -yz[I, J, t - 1]
            yz[I, J, t + 2] = -yz[I, J, t - 2]
            xz[I, J, t] = 0.0
            yz[I, J, t] = 0.0
            f["zz"], f["xz"], f["yz"] = zz, xz, yz
            traces.append(torch.stack([torch.stack([f["u1"][i, j, k], f["v1"][i, j, k],
                                                    f["w1"][i, j, k - 1]])
                                       for (i, j, k) in self.rec]))
        return torch.stack(traces)

    def _addsrc(self, f, sample):
        vtst = self.DT / (self.DH ** 3)
        out = dict(f)
        comps = (("xx", 0), ("yy", 1), ("zz", 2), ("xz", 3), ("yz", 4), ("xy", 5))
        for name, c in comps:
            a = out[name].clone()
            for n, (i, j, k) in enumerate(self.src):
                a[i, j, k] = a[i, j, k] - vtst * float(self.rates[n, sample - 1, c])
            out[name] = a
        return out
