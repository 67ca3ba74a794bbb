"""Seeded synthetic FlyWire-shaped connectome (the benchmark workload; real FlyWire data needs
a download the reference performs in flyzero/connectome.py:52-65, which is out of scope here).

``flywire_like()`` builds, from one seed:

* N = 138,639 neurons, E = 15.1 M directed connections (pre x post CSR, rows sorted);
* out-degree ~ lognormal(mu, sigma=1.2) with mu = ln(E/N) - sigma^2/2, rescaled to sum to E;
* targets drawn with probability proportional to a power law of a random hub rank,
  (rank + j0)^(-1/(alpha-1)) with alpha = 2.1, j0 chosen so the largest expected in-degree is
  5% of N; self-loops and duplicate connections are redrawn;
* synapses per connection ~ geometric(p=0.28) on {1, 2, ...} (mean 3.57), clipped to 2,400;
* one sign per presynaptic neuron (Dale's rule): 55% excitatory, 45% inhibitory;
* 12,246 visual input neurons with retinotopic positions in [0,1)^2 and rates U(0, 200) Hz,
  plus one drive neuron at 60 Hz;
* descending-neuron cell types with the names the reference's readout uses
  (flyzero/instruct.py:43-48: DNa02, DNg02*, DNa01 per side, DNp09, MDN), ten neurons each,
  picked among neurons whose in-degree is near 7,300 / 40 so that all inputs onto the four
  motor types DNa02, DNa01, DNg02*, DNp09 number about 7,300; and the giant fiber DNp01 (one
  per side) that the batched instructed rules also drive (flyzero/batch.py:20-22).

The matrix holds signed synapse counts like ``Connectome.weights`` (flyzero/connectome.py:68-120);
``find(type, side)`` mirrors ``Connectome.find``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

N_FLYWIRE = 138_639
E_FLYWIRE = 15_100_000
N_INPUTS = 12_246

# the 8 descending-neuron rate groups (flyzero/instruct.py:43-48)
GROUPS = {
    "a02L": ("DNa02", "left"), "a02R": ("DNa02", "right"),
    "g02L": ("DNg02*", "left"), "g02R": ("DNg02*", "right"),
    "a01L": ("DNa01", "left"), "a01R": ("DNa01", "right"),
    "gas": ("DNp09", None), "brake": ("MDN", None),
}
MOTOR_TYPES = ("DNa02", "DNa01", "DNg02*", "DNp09")
DN_TYPES = ("DNa02", "DNg02*", "DNa01", "DNp09", "MDN")


@dataclass
class SyntheticFly:
    weights: sp.csr_matrix          # signed synapse counts, pre x post
    input_idx: np.ndarray           # visual input neurons
    input_rate: np.ndarray          # Hz, U(0, 200)
    input_xy: np.ndarray            # (n_inputs, 2) retinotopic position in [0, 1)^2
    drive_idx: int
    drive_hz: float
    types: dict = field(default_factory=dict)   # cell type -> neuron ids
    side: np.ndarray | None = None               # 0 left, 1 right per neuron
    seed: int = 0

    @property
    def n(self) -> int:
        return self.weights.shape[0]

    def find(self, cell_type: str, side: str | None = None) -> np.ndarray:
        ids = self.types.get(cell_type, np.zeros(0, np.int64))
        if side is None:
            return ids
        want = 0 if side == "left" else 1
        return ids[self.side[ids] == want]

    def group_members(self) -> list[np.ndarray]:
        return [self.find(t, s) for t, s in GROUPS.values()]


def _hub_offset(n: int, n_edges: int, beta: float, cap: float) -> float:
    """j0 such that n_edges * j0^-beta / sum_r (r + j0)^-beta == cap (bisection in log space)."""
    r = np.arange(n, dtype=np.float64)
    lo, hi = 1e-3, float(n)
    for _ in range(80):
        mid = np.sqrt(lo * hi)
        w = (r + mid) ** -beta
        top = n_edges * w[0] / w.sum()
        if top > cap:
            lo = mid
        else:
            hi = mid
    return float(np.sqrt(lo * hi))


def flywire_like(n: int = N_FLYWIRE, n_edges: int = E_FLYWIRE, seed: int = 0, sigma: float = 1.2,
                 alpha: float = 2.1, p_count: float = 0.28, max_count: int = 2400, frac_exc: float = 0.55,
                 n_inputs: int | None = None, max_in_frac: float = 0.05, dn_per_type: int = 10,
                 plastic_target: int = 7_300) -> SyntheticFly:
    rng = np.random.default_rng(seed)
    # out-degrees
    mean = n_edges / n
    mu = np.log(mean) - sigma ** 2 / 2
    d = rng.lognormal(mu, sigma, n)
    d *= n_edges / d.sum()
    deg = np.floor(d).astype(np.int64)
    rem = int(n_edges - deg.sum())
    if rem > 0:
        deg[np.argsort(-(d - deg), kind="stable")[:rem]] += 1
    deg = np.minimum(deg, n - 1)
    # in-degree propensity: power law over a random hub rank
    beta = 1.0 / (alpha - 1.0)
    j0 = _hub_offset(n, int(deg.sum()), beta, max_in_frac * n)
    rank = rng.permutation(n)
    wgt = (rank.astype(np.float64) + j0) ** -beta
    cdf = np.cumsum(wgt)
    cdf /= cdf[-1]
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    cols = np.searchsorted(cdf, rng.random(len(rows)), side="right").astype(np.int64)
    np.minimum(cols, n - 1, out=cols)
    for _ in range(64):  # redraw self-loops and duplicate connections
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(len(ks), bool)
        dup[1:] = ks[1:] == ks[:-1]
        bad = np.zeros(len(key), bool)
        bad[order[dup]] = True
        bad |= rows == cols
        nb = int(bad.sum())
        if nb == 0:
            break
        cols[bad] = np.minimum(np.searchsorted(cdf, rng.random(nb), side="right"), n - 1)
    else:  # drop whatever could not be placed
        keep = ~bad
        rows, cols = rows[keep], cols[keep]
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    rows, cols = rows[order], cols[order]
    count = np.minimum(rng.geometric(p_count, len(rows)), max_count).astype(np.float32)
    sign = np.where(rng.random(n) < frac_exc, 1.0, -1.0).astype(np.float32)
    data = sign[rows] * count
    indptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int64)
    w = sp.csr_matrix((data, cols.astype(np.int32), indptr), shape=(n, n))

    # inputs (seed + 1: neurons and positions; seed + 2: rates)
    if n_inputs is None:
        n_inputs = N_INPUTS if n == N_FLYWIRE else max(1, int(round(N_INPUTS * n / N_FLYWIRE)))
    r1 = np.random.default_rng(seed + 1)
    chosen = r1.choice(n, n_inputs + 1, replace=False)
    input_idx, drive_idx = np.sort(chosen[:-1]).astype(np.int64), int(chosen[-1])
    input_xy = r1.random((n_inputs, 2))
    input_rate = np.random.default_rng(seed + 2).uniform(0.0, 200.0, n_inputs).astype(np.float32)

    # descending neurons: in-degree near plastic_target / (4 motor types * dn_per_type)
    indeg = np.bincount(cols, minlength=n)
    taken = np.zeros(n, bool)
    taken[chosen] = True
    want = plastic_target / (len(MOTOR_TYPES) * dn_per_type)
    r3 = np.random.default_rng(seed + 3)
    cand = np.flatnonzero(~taken & (indeg >= 0.9 * want) & (indeg <= 1.1 * want))
    if len(cand) < len(DN_TYPES) * dn_per_type:
        cand = np.argsort(np.abs(indeg - want) + taken * 1e9)[: len(DN_TYPES) * dn_per_type * 2]
    pick = r3.choice(cand, len(DN_TYPES) * dn_per_type, replace=False)
    types = {t: np.sort(pick[k * dn_per_type:(k + 1) * dn_per_type]).astype(np.int64)
             for k, t in enumerate(DN_TYPES)}
    side = (r3.random(n) < 0.5).astype(np.int8)
    for t, ids in types.items():   # balanced sides within a type
        side[ids] = (np.arange(len(ids)) >= len(ids) // 2).astype(np.int8)
    # the giant fiber (one per side), drawn afterwards from its own stream so that the types
    # above do not depend on it
    taken[pick] = True
    r6 = np.random.default_rng(seed + 6)
    cand = np.argsort(np.abs(indeg - want) + taken * 1e9)[:64]
    gf = np.sort(r6.choice(cand, 2, replace=False)).astype(np.int64)
    types["DNp01"] = gf
    side[gf] = (0, 1)
    return SyntheticFly(w, input_idx, input_rate, input_xy, drive_idx, 60.0, types, side, seed)


def smooth_field(xy: np.ndarray, frame: int, seed: int = 4, n_blobs: int = 6, peak_hz: float = 200.0) -> np.ndarray:
    """Input rates (Hz, float32) for one game frame: a sum of drifting Gaussian blobs over the
    retinotopic positions ``xy`` (the frame-to-frame drive that replaces the game's eye)."""
    r = np.random.default_rng(seed)
    c0 = r.random((n_blobs, 2))
    vel = (r.random((n_blobs, 2)) - 0.5) * 0.02
    width = 0.08 + 0.12 * r.random(n_blobs)
    amp = 0.4 + 0.6 * r.random(n_blobs)
    c = np.mod(c0 + vel * frame, 1.0)
    dx = np.abs(xy[:, None, :] - c[None, :, :])
    dx = np.minimum(dx, 1.0 - dx)                     # torus
    f = (amp[None] * np.exp(-(dx ** 2).sum(-1) / (2 * width[None] ** 2))).sum(-1)
    return (peak_hz * np.clip(f, 0.0, 1.0)).astype(np.float32)


def plastic_positions(fly: SyntheticFly, types=MOTOR_TYPES) -> np.ndarray:
    """CSR positions of every synapse onto the given cell types (flyzero/batch.py:58-65)."""
    targets = np.unique(np.concatenate([fly.find(t) for t in types]))
    is_t = np.zeros(fly.n, bool)
    is_t[targets] = True
    w = sp.csr_matrix(fly.weights, dtype=np.float32)
    w.sort_indices()
    return np.flatnonzero(is_t[w.indices])
