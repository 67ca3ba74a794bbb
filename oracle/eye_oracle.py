"""CPU restatement of the motion-vision front end for one fly (test infrastructure only).

Restates MotionEye.see (flyzero/vision.py:137-191, default parameters: motion channels only) in
float32 numpy with the device's order: per pixel ((r + g) + b) / 3 / 255, summed over the cell
row by row, column by column, divided by cell^2.  Nothing in the product imports this module.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


class OracleEye:
    def __init__(self, key, gy, gx, py, frame_shape, cell=4, shift=1, tau_hp=4.0, tau_delay=2.0, gain=1000.0,
                 max_rate=150.0, y_min=0.0):
        self.key, self.gy, self.gx, self.py = (np.asarray(a) for a in (key, gy, gx, py))
        self.h, self.w = frame_shape
        self.c, self.k = cell, shift
        self.gh, self.gw = self.h // cell, self.w // cell
        self.a_hp, self.a_d = f32(1.0 / tau_hp), f32(1.0 / tau_delay)
        self.gain, self.max_rate, self.y_cut = f32(gain), f32(max_rate), f32(y_min * self.h)
        self.lp = None
        self.d = None

    def grid(self, frame):
        c, gh, gw = self.c, self.gh, self.gw
        img = frame[: gh * c, : gw * c].astype(f32)
        pix = (img[..., 0] + img[..., 1] + img[..., 2]) / f32(3.0) / f32(255.0)
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
        m = np.zeros(len(self.key), f32)
        for key, mp in maps.items():
            sel = self.key == key
            m[sel] = mp[self.gy[sel], self.gx[sel]]
        rates = self.max_rate * np.minimum(np.maximum(self.gain * m, f32(0.0)), f32(1.0))
        rates[self.py.astype(f32) < self.y_cut] = 0.0
        for pol in (0, 1):
            self.d[pol] = self.d[pol] + self.a_d * (sig[pol] - self.d[pol])
        return rates.astype(f32)
