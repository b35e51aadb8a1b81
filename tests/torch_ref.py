"""Naive autograd reference of the ARD time step in torch (test infrastructure).

A differentiable restatement of the reference's step, used only to check the
hand-written adjoint: torch records every operation of every step and
backpropagates through the whole time loop (memory grows with the number of
steps, hence "naive"). It follows the same algorithm as the CPU oracle:
  sources      F[cell] = sample            (ARD/partition.cpp:187-197)
  DCT update   M+ = .999 (2 M cos - M- + 2 F^/w^2 (1 - cos)), transforms
               REDFT10 x 1/(2 sqrt(2N)), REDFT01 x 1/sqrt(2N)
                                            (ARD/dct_partition.cpp:48-73, ARD/dct_volume.cpp:38-68)
  PML update   Grote-Sim leapfrog           (ARD/pml_partition.cpp:113-200)
  interface    6th-order forcing, last writer wins (ARD/boundary.cpp:21-173)
  recorder     pressure at the recorder cell (ARD/recorder.cpp:40-57)
The topology (PML slabs, boundaries, recorder owners) comes from a host-only
context of the library; the mode tables and PML profiles from the oracle.
Recorders use partition-local coordinates (the scenes' ARD_REC_LOCAL mode).
"""
import math

import numpy as np
import torch
import torch.nn.functional as Fn

import ard_oracle as O
import ard_scenes as S
import ard_sim as A

DAMP = float(np.float32(0.999))  # R(0.999f) of ARD/dct_partition.cpp:68
C6 = [2.0, -27.0, 270.0, -490.0, 270.0, -27.0, 2.0]
C4 = [1.0, -8.0, 0.0, 8.0, -1.0]
CB = [[0, 0, -2, 2, 0, 0], [0, -2, 27, -27, 2, 0], [-2, 27, -270, 270, -27, 2],
      [2, -27, 270, -270, 27, -2], [0, 2, -27, 27, -2, 0], [0, 0, 2, -2, 0, 0]]


def _cos_mats(n, dtype, device):
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    c2 = 2.0 * np.cos(np.pi * k * (j + 0.5) / n)             # REDFT10: Y_k = sum_j c2[k][j] x_j
    c3 = np.where(k.T == 0, 1.0, 2.0 * np.cos(np.pi * k.T * (j.T + 0.5) / n))  # REDFT01: Y_j = sum_k c3[j][k] X_k
    return (torch.tensor(c2, dtype=dtype, device=device), torch.tensor(c3, dtype=dtype, device=device))


def _shift(x, axis, s):
    """y[i] = x[i + s] along axis (zero outside), for a [D][H][W] volume (axis 0 = x)."""
    dim = 2 - axis
    n = x.shape[dim]
    if abs(s) >= n:
        return torch.zeros_like(x)
    if s > 0:
        y = torch.narrow(x, dim, s, n - s)
        pad = [0, 0] * 3
        pad[2 * (2 - dim) + 1] = s
        return Fn.pad(y, pad)
    y = torch.narrow(x, dim, 0, n + s)
    pad = [0, 0] * 3
    pad[2 * (2 - dim)] = -s
    return Fn.pad(y, pad)


def _dct_fft(v, dim, inverse):
    """REDFT10 (inverse: REDFT01) along dim through a complex FFT of length 2n:
    REDFT10  Y_k = Re(e^{-i pi k / 2n} FFT_2n([x, reversed x])_k)
    REDFT01  Y_j = Re(2n IFFT_2n(Z)_j),  Z_k = c_k X_k e^{i pi k / 2n} (k < n), c_0 = 1, c_k = 2"""
    n = v.shape[dim]
    k = torch.arange(n, device=v.device, dtype=v.dtype)
    shape = [1] * v.dim()
    shape[dim] = n
    ph = torch.polar(torch.ones_like(k), (math.pi / (2 * n)) * k).reshape(shape)
    if not inverse:
        ext = torch.cat([v, torch.flip(v, [dim])], dim=dim)
        f = torch.fft.fft(ext, dim=dim)
        return (torch.narrow(f, dim, 0, n) * ph.conj()).real
    c = torch.where(k == 0, torch.ones_like(k), 2 * torch.ones_like(k)).reshape(shape)
    z = torch.cat([v * c * ph, torch.zeros_like(v, dtype=ph.dtype)], dim=dim)
    return (2 * n) * torch.narrow(torch.fft.ifft(z, dim=dim), dim, 0, n).real


class TorchARD:
    def __init__(self, scene, device="cpu", dtype=torch.float64, dct="matrix"):
        assert scene.recorder_mode == S.REC_LOCAL
        assert dct in ("matrix", "fft")
        self.sc, self.device, self.dtype, self.dct = scene, device, dtype, dct
        host = A.Simulation(scene, device=-1, finalize=False)
        self.parts = host.partitions()
        self.bnds = host.boundaries()
        host.close()
        # DCT partitions keep the scene's box order (ids 0 .. nboxes-1)
        self.rec_part = [S.recorder_owner(scene.boxes, r) for r in scene.recorders]
        h, dt, c0 = scene.h, scene.dt, scene.c0
        self.mats = {}
        self.tab = {}
        self.prof = {}
        for i, p in enumerate(self.parts):
            is_pml, ptype, _, xs, ys, zs, w, hh, d = p
            if is_pml:
                z = O.pml_profile(h, c0, scene.pml_layers, ptype, w, hh, d, double=True)
                self.prof[i] = torch.tensor(z, dtype=dtype, device=device)  # [3][D][H][W], scale 1
            else:
                w2, cwt = O.mode_tables(h, dt, c0, w, hh, d, double=True)
                n = w * hh * d
                self.tab[i] = (torch.tensor(w2, dtype=dtype, device=device),
                               torch.tensor(cwt, dtype=dtype, device=device),
                               1.0 / (2.0 * math.sqrt(2.0 * n)), 1.0 / math.sqrt(2.0 * n))
                for L in (w, hh, d):
                    if L not in self.mats:
                        self.mats[L] = _cos_mats(L, dtype, device)
        self.src = []
        for s in scene.sources:  # owner: the DCT partition containing the cell
            for i, p in enumerate(self.parts):
                if p[0]:
                    continue
                xs, ys, zs, w, hh, d = p[3:]
                if xs <= s.x < xs + w and ys <= s.y < ys + hh and zs <= s.z < zs + d:
                    self.src.append((i, s.z - zs, s.y - ys, s.x - xs))
                    break
            else:
                self.src.append(None)

    def _dct3(self, v, inverse):
        if self.dct == "fft":
            for dim in (2, 1, 0):
                v = _dct_fft(v, dim, inverse)
            return v
        d, hh, w = v.shape
        c2w, c3w = self.mats[w]
        c2h, c3h = self.mats[hh]
        c2d, c3d = self.mats[d]
        Aw, Ah, Ad = (c3w, c3h, c3d) if inverse else (c2w, c2h, c2d)
        v = torch.matmul(v, Aw.t().contiguous())                          # along x
        v = torch.matmul(Ah, v)                                           # along y (per z plane)
        v = torch.matmul(Ad, v.reshape(d, hh * w)).reshape(d, hh, w)      # along z
        return v

    def run(self, waves, theta, steps):
        """Recorder traces [steps][nrec] for source waveforms `waves` (list of 1-D
        tensors, None for Gaussian sources) and PML damping scales theta [6]."""
        dev, dt_, sc = self.device, self.dtype, self.sc
        h, dt, c0 = sc.h, sc.dt, sc.c0
        st = {}
        for i, p in enumerate(self.parts):
            w, hh, d = p[6:9]
            z = lambda: torch.zeros((d, hh, w), dtype=dt_, device=dev)
            if p[0]:
                st[i] = {"p": z(), "q": z(), "phi": [z(), z(), z()], "f": z()}
            else:
                st[i] = {"M": z(), "Mp": z(), "F": z(), "p": z()}
        kb = sc.absorption * c0 * c0 / (180.0 * h * h)
        courant = np.float32(c0) * np.float32(dt) / np.float32(h)
        out = []
        for t in range(steps):
            # sources
            for si, loc in enumerate(self.src):
                if loc is None:
                    continue
                i, z0, y0, x0 = loc
                if waves[si] is None:
                    arg = (np.float32(np.pi) * ((np.float32(2) * courant * np.float32(t)) / np.float32(6) - np.float32(2))) ** 2
                    val = torch.tensor(float(np.float32(1e9) * np.exp(-np.float32(arg))), dtype=dt_, device=dev)
                else:
                    val = waves[si][t] if t < waves[si].shape[0] else torch.zeros((), dtype=dt_, device=dev)
                F = st[i]["F"].clone()
                F[z0, y0, x0] = val
                st[i]["F"] = F
            # partition updates
            for i, p in enumerate(self.parts):
                s = st[i]
                if p[0]:
                    ptype = p[1]
                    zeta = [theta[ptype] * self.prof[i][a] for a in range(3)]
                    P, Q = s["p"], s["q"]
                    lap = sum(C6[m] * _shift(P, a, m - 3) for a in range(3) for m in range(7)) / (180.0 * h * h)
                    dphi = sum(C4[m] * _shift(s["phi"][a], a, m - 2) for a in range(3) for m in range(5) if m != 2) / (12.0 * h)
                    du = [sum(C4[m] * _shift(P, a, m - 2) for m in range(5) if m != 2) / (12.0 * h) for a in range(3)]
                    zs_ = zeta[0] + zeta[1] + zeta[2]
                    zz2 = zeta[0] * zeta[1] + zeta[1] * zeta[2] + zeta[0] * zeta[2]
                    pn = 2 * P - Q + dt * dt * (c0 * c0 * lap - zs_ * (P - Q) / dt - zz2 * P + dphi + s["f"])
                    phin = [s["phi"][a] - dt * zeta[a] * s["phi"][a] + dt * (zs_ - 2 * zeta[a]) * du[a] for a in range(3)]
                    st[i] = {"p": pn, "q": P, "phi": phin, "f": torch.zeros_like(P)}
                else:
                    w2, cwt, sf, si_ = self.tab[i]
                    Fh = self._dct3(s["F"], False) * sf
                    Mn = DAMP * (2 * s["M"] * cwt - s["Mp"] + (2 * Fh / w2) * (1 - cwt))
                    pv = self._dct3(Mn, True) * si_
                    st[i] = {"M": Mn, "Mp": s["M"], "F": s["F"], "p": pv}
            # interface forcing, boundaries in list order (last writer wins)
            for (typ, a, b, xs, xe, ys, ye, zs, ze) in self.bnds:
                pa, pb = self.parts[a], self.parts[b]
                ax = typ  # 0 x, 1 y, 2 z
                a_end = pa[3 + ax] + pa[6 + ax]
                lo_hi = (xs, xe, ys, ye, zs, ze)
                a_first = lo_hi[2 * ax] <= a_end and lo_hi[2 * ax + 1] >= a_end
                L_, R_ = (a, b) if a_first else (b, a)
                pl, pr = self.parts[L_], self.parts[R_]
                # tangential ranges in each side's local coordinates, as [D][H][W] slices
                def sl(part, nrm):
                    o = part[3:6]
                    rng = [slice(xs - o[0], xe - o[0]), slice(ys - o[1], ye - o[1]), slice(zs - o[2], ze - o[2])]
                    rng[ax] = nrm
                    return (rng[2], rng[1], rng[0])
                nl = pl[6 + ax]
                PL = [st[L_]["p"][sl(pl, nl - 3 + n)] for n in range(3)]
                PR = [st[R_]["p"][sl(pr, n)] for n in range(3)]
                self_l, self_r = not pl[0], not pr[0]
                fkey_l = "f" if pl[0] else "F"
                fkey_r = "f" if pr[0] else "F"
                for r in range(6):
                    m = r - 3
                    s1 = sum(CB[r][n] * PL[n] for n in range(3))
                    s2 = sum(CB[r][n + 3] * PR[n] for n in range(3))
                    if m < 0:
                        val = (s1 + s2 if self_l else s2) * kb
                        G = st[L_][fkey_l].clone()
                        G[sl(pl, nl + m)] = val
                        st[L_][fkey_l] = G
                    else:
                        val = (s1 + s2 if self_r else s1) * kb
                        G = st[R_][fkey_r].clone()
                        G[sl(pr, m)] = val
                        st[R_][fkey_r] = G
            # recorders (partition-local coordinates)
            row = []
            for r, (x, y, z) in enumerate(sc.recorders):
                i = self.rec_part[r]
                p = self.parts[i]
                row.append(st[i]["p"][z - p[5], y - p[4], x - p[3]])
            out.append(torch.stack(row))
        return torch.stack(out)
