"""CPU restatement of the vision front end for one fly (test infrastructure only).

Restates MotionEye.see (flyzero/vision.py:137-191: motion, medulla transient / sustained and
colour channels) in float32 numpy with the device's order: per pixel ((r + g) + b) / 3 / 255,
summed over the cell row by row, column by column, divided by cell^2; the sustained channel's
surround is the plain sum over the clipped window (row-major) divided by its size, where the
reference differences a float32 integral image and finishes in float64 (the two agree to a few
float32 ulps of the grid sums).  Nothing in the product imports this module.

Keys: 0-7 motion (polarity * 4 + direction: right, left, down, up); 8 / 9 ON / OFF transient;
10 / 11 ON / OFF sustained contrast; 12 blue; 13 green.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


class OracleEye:
    def __init__(self, key, gy, gx, py, frame_shape, cell=4, shift=1, tau_hp=4.0, tau_delay=2.0, gain=1000.0,
                 max_rate=150.0, y_min=0.0, transient_gain=0.0, sustained_gain=0.0, surround=3, chromatic_gain=0.0):
        self.key, self.gy, self.gx, self.py = (np.asarray(a) for a in (key, gy, gx, py))
        self.h, self.w = frame_shape
        self.c, self.k = cell, shift
        self.gh, self.gw = self.h // cell, self.w // cell
        self.a_hp, self.a_d = f32(1.0 / tau_hp), f32(1.0 / tau_delay)
        self.gain, self.max_rate, self.y_cut = f32(gain), f32(max_rate), f32(y_min * self.h)
        self.tr, self.su, self.ch, self.r = f32(transient_gain), f32(sustained_gain), f32(chromatic_gain), surround
        self.lp = None
        self.d = None

    def grid(self, frame, channel=None):
        c, gh, gw = self.c, self.gh, self.gw
        img = frame[: gh * c, : gw * c].astype(f32)
        if channel is None:
            pix = (img[..., 0] + img[..., 1] + img[..., 2]) / f32(3.0) / f32(255.0)
        else:
            pix = img[..., channel] / f32(255.0)
        s = np.zeros((gh, gw), f32)
        for dy in range(c):
            for dx in range(c):
                s = s + pix[dy::c, dx::c]
        return s / f32(c * c)

    def see(self, frame):
        lum = self.grid(frame)
        if self.lp is None:
            self.lp = lum.copy()
            self.d = [np.zeros_like(lum), np.zeros_like(lum)]
        self.lp = self.lp + self.a_hp * (lum - self.lp)
        hp = lum - self.lp
        sig = [np.maximum(hp, f32(0.0)), np.maximum(-hp, f32(0.0))]
        k = self.k
        maps = {}
        for pol in (0, 1):
            s, d = sig[pol], self.d[pol]
            right = np.zeros_like(s)
            down = np.zeros_like(s)
            right[:, :-k] = d[:, :-k] * s[:, k:] - s[:, :-k] * d[:, k:]
            down[:-k] = d[:-k] * s[k:] - s[:-k] * d[k:]
            maps[pol * 4 + 0] = np.maximum(right, f32(0.0))
            maps[pol * 4 + 1] = np.maximum(-right, f32(0.0))
            maps[pol * 4 + 2] = np.maximum(down, f32(0.0))
            maps[pol * 4 + 3] = np.maximum(-down, f32(0.0))
        feats = {}
        if self.tr != 0:
            feats[8], feats[9] = self.tr * sig[0], self.tr * sig[1]
        if self.su != 0:
            con = lum - self.surround_mean(lum)
            feats[10] = self.su * np.maximum(con, f32(0.0))
            feats[11] = self.su * np.maximum(-con, f32(0.0))
        if self.ch != 0:
            feats[12], feats[13] = self.ch * self.grid(frame, 2), self.ch * self.grid(frame, 1)
        for key, mp in feats.items():
            maps[key] = mp / self.gain
        m = np.zeros(len(self.key), f32)
        for key, mp in maps.items():
            sel = self.key == key
            m[sel] = mp[self.gy[sel], self.gx[sel]]
        rates = self.max_rate * np.minimum(np.maximum(self.gain * m, f32(0.0)), f32(1.0))
        rates[self.py.astype(f32) < self.y_cut] = 0.0
        for pol in (0, 1):
            self.d[pol] = self.d[pol] + self.a_d * (sig[pol] - self.d[pol])
        return rates.astype(f32)

    def surround_mean(self, lum):
        """Window sums in row-major window order; a clipped entry adds an exact 0 (lum >= 0)."""
        gh, gw, r = self.gh, self.gw, self.r
        ys, xs = np.arange(gh)[:, None], np.arange(gw)[None, :]
        s = np.zeros((gh, gw), f32)
        pad = np.zeros((gh + 2 * r, gw + 2 * r), f32)
        pad[r:r + gh, r:r + gw] = lum
        for oy in range(-r, r + 1):
            for ox in range(-r, r + 1):
                s = s + pad[r + oy:r + oy + gh, r + ox:r + ox + gw]
        cnt = ((np.minimum(ys + r + 1, gh) - np.maximum(ys - r, 0)) *
               (np.minimum(xs + r + 1, gw) - np.maximum(xs - r, 0)))
        return s / cnt.astype(f32)
