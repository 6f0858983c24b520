"""The reference's per-site random number streams (site_random.f90) and its
Gaussian noise (gaussian.F90), vectorised over the lattice with numpy.

Every (x, y, t) site owns an independent generator: a 48-bit linear
congruential generator held as two 24-bit halves in doubles (exact integer
arithmetic), single-precision outputs, and the Numerical Recipes shuffle over
a 97-entry table (site_random.f90:28-150).  Seeding adds the site's global
index to the seed (site_random.f90:64-72), so the streams, and hence the
noise fields, are identical to the reference's for any lattice split.

gauss0 / gaussp follow gaussian.F90:18-106 in single precision: one uniform
per site for the modulus, a second for the angle.
"""
import numpy as np

_DM1, _DM2 = 15184245.0, 2651554.0
_DX24, _DX48 = 16777216.0, 281474976710656.0
_TPI = np.float32(2.0) * np.float32(np.arccos(np.float32(-1.0)))


class SiteRandom:
    """Generator state for an lx x ly x lt lattice, arrays indexed [x, y, t]."""

    def __init__(self, lx, ly, lt, seed):
        self.shape = (lx, ly, lt)
        x = np.arange(lx)[:, None, None]
        y = np.arange(ly)[None, :, None]
        t = np.arange(lt)[None, None, :]
        # seed_offset (site_random.f90:64-72) on one process: ksize = lx
        off = (x + lx * y + lx * lx * t).astype(np.float64)
        dseed = off + float(seed)
        self.ds2 = np.trunc(dseed / _DX24)
        self.ds1 = dseed - self.ds2 * _DX24
        # init_random: the first rano call fills the shuffle table
        for _ in range(97):
            self._rranf()
        self.v = np.stack([self._rranf() for _ in range(97)])  # [97, x, y, t]
        self.y = self._rranf()
        self.rano()

    def _dranf(self, idx=Ellipsis):
        """One step of the 48-bit generator of every site (idx = Ellipsis) or
        of the one site idx = (x, y, t)."""
        ds1, ds2 = self.ds1[idx], self.ds2[idx]
        dl = ds1 * _DM1
        dc = np.trunc(dl / _DX24)
        dl = dl - dc * _DX24
        du = ds1 * _DM2 + ds2 * _DM1 + dc
        ds2 = du - np.trunc(du / _DX24) * _DX24
        self.ds1[idx] = dl
        self.ds2[idx] = ds2
        return (ds2 * _DX24 + dl) / _DX48

    def _rranf(self, idx=Ellipsis):
        return np.asarray(self._dranf(idx)).astype(np.float32)

    def rano(self):
        """One shuffled single-precision uniform per site (site_random.f90:33-60)."""
        j = (np.float32(97.0) * self.y).astype(np.int64)  # int() truncates toward 0
        j = np.clip(j, 0, 96)
        ix = np.indices(self.shape)
        self.y = self.v[(j,) + tuple(ix)]
        out = self.y.copy()
        self.v[(j,) + tuple(ix)] = self._rranf()
        return out

    def rano_at(self, ix, iy, it):
        """rano(y, i, ix, iy, it) for one site, 0-based (site_random.f90:33-60):
        advances only that site's stream.  rano_at(0, 0, 0) is the
        reference's rano(yran, idum, 1, 1, 1), which ends MD trajectories
        (evolution.F90:159-161) and draws the Metropolis test
        (dwf3d_lib.F90:309)."""
        s = (ix, iy, it)
        j = min(max(int(np.float32(97.0) * self.y[s]), 0), 96)
        out = np.float32(self.v[(j,) + s])
        self.y[s] = out
        self.v[(j,) + s] = self._rranf(s)
        return out

    def _gauss(self, two):
        r = self.rano()
        with np.errstate(divide="ignore"):
            mod = np.sqrt((np.float32(-2.0) if two else np.float32(-1.0)) * np.log(r))
        theta = _TPI * self.rano()
        ps = np.empty(self.shape + (2,), dtype=np.float32)
        ps[..., 0] = mod * np.sin(theta)
        ps[..., 1] = mod * np.cos(theta)
        return ps

    def gauss0(self):
        """Unit-variance complex noise per site, [x, y, t, re/im] (gaussian.F90:63-106)."""
        return self._gauss(False)

    def gaussp(self):
        """Momentum refresh noise, [x, y, t, 2] (gaussian.F90:18-60)."""
        return self._gauss(True)
