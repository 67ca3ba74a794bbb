"""FlyWire ingestion: the three files of materialization 783 (their schema is described in
flyzero/connectome.py:1-19) -> a ``Connectome`` of signed synapse counts and a neuron table, the
biological corrections of flyzero/biology.py:38-55, and an on-disk CSR cache (int64 row
pointers, int32 columns, int16 counts, memory-mapped on load) so that the 15 M-connection matrix
is parsed once.

The files are not in this build's environment (the reference downloads them,
flyzero/connectome.py:52-65); nothing here fetches anything.  ``BatchBrain(conn.weights, ...)``
and every ``find`` caller work with a ``Connectome`` as with the synthetic fly.

Design: the connectivity table is read column by column (pyarrow, no row objects) and turned into
a CSR by one stable sort of the (pre, post) key and a merge of repeated pairs, so peak memory is a
few arrays of E entries; cell-type lookups factorize the type columns once; the corrections are a
per-presynaptic-neuron rule applied to the CSR's values in place of a COO round trip.
"""

from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.sparse as sp

# file names and columns of the FlyWire v783 release used by Shiu et al. 2024
COMPLETENESS = "Completeness_783.csv"          # one row per neuron; its first column is the root id
CONNECTIVITY = "Connectivity_783.parquet"      # one row per connected pair
ANNOTATIONS = "neuron_annotations.tsv"         # Schlegel et al. 2024 neuron annotations
FILES = (COMPLETENESS, CONNECTIVITY, ANNOTATIONS)
PRE, POST, SIGNED_COUNT = "Presynaptic_Index", "Postsynaptic_Index", "Excitatory x Connectivity"
TEXT_COLUMNS = ("super_class", "cell_class", "cell_type", "hemibrain_type", "side", "top_nt")
POSITION_COLUMNS = ("pos_x", "pos_y", "pos_z")
# column order of the neuron table (flyzero/connectome.py:155-158)
ANNOTATION_COLUMNS = ("super_class", "cell_class", "cell_type", "hemibrain_type", "side", "top_nt") + POSITION_COLUMNS

# transmitters without fast ionotropic receptors in Drosophila (flyzero/biology.py:9-13)
MODULATORY = ("dopamine", "serotonin", "octopamine")
ALL_FIXES = ("modulatory", "kenyon_cholinergic")
# FlyWire photoreceptor types: R1-6, R7 / R8 and their subtypes (R7p, R8y, R7d, ...)
PHOTORECEPTOR_RE = re.compile(r"^R[1-8]([a-z]|-6)?$|^R1-6$", re.IGNORECASE)


def _lower_codes(column):
    """(codes, lowercase names) of a text column; missing entries get code -1."""
    import pandas as pd

    codes, names = pd.factorize(column.str.lower(), use_na_sentinel=True)
    return np.asarray(codes), [str(x) for x in names]


@dataclass
class Connectome:
    """``weights``: pre x post CSR of signed synapse counts; ``neurons``: one row per neuron in
    index order (root_id, then the annotation columns) -- the contract of flyzero/connectome.py:68-121."""

    weights: sp.csr_matrix
    neurons: object   # pandas.DataFrame
    name: str = "connectome"
    _index: dict = field(default_factory=dict, repr=False)

    @property
    def n(self) -> int:
        return self.weights.shape[0]

    def _codes(self, column: str):
        if column not in self._index:
            self._index[column] = _lower_codes(self.neurons[column]) if column in self.neurons else None
        return self._index[column]

    def find(self, cell_type: str, side: str | None = None) -> np.ndarray:
        """Neurons of a cell type, case-insensitive: an exact match of the cell type or of the
        hemibrain type, or with a trailing ``*`` any cell type with that prefix (``DNg02*`` for
        DNg02_a ... DNg02_h).  ``side``: keep one hemisphere ('left' / 'right')."""
        want = cell_type.lower()
        prefix = want.endswith("*")
        hit = np.zeros(self.n, bool)
        for column in ("cell_type",) if prefix else ("cell_type", "hemibrain_type"):
            entry = self._codes(column)
            if entry is None:
                continue
            codes, names = entry
            keys = [k for k, name in enumerate(names) if (name.startswith(want[:-1]) if prefix else name == want)]
            if keys:
                hit |= np.isin(codes, keys)
        if side is not None:
            hit &= self.neurons["side"].to_numpy(dtype=object, na_value=None) == side
        return np.flatnonzero(hit)

    def photoreceptors(self) -> np.ndarray:
        """Photoreceptor neurons by cell type; for annotation releases without per-type labels,
        sensory neurons whose class names a photoreceptor, retina or visual input
        (flyzero/connectome.py:103-111)."""
        types = self.neurons["cell_type"].to_numpy(dtype=object, na_value="")
        hit = np.fromiter((PHOTORECEPTOR_RE.match(t) is not None for t in types), bool, len(types))
        if not hit.any():
            cls = self.neurons["cell_class"].to_numpy(dtype=object, na_value="")
            sensory = self.neurons["super_class"].to_numpy(dtype=object, na_value="") == "sensory"
            visual = np.fromiter((re.search("photoreceptor|retin|visual", c.lower()) is not None for c in cls),
                                 bool, len(cls))
            hit = sensory & visual
        return np.flatnonzero(hit)


def _ptr(indptr: np.ndarray) -> np.ndarray:
    """Row pointers as int32 when they fit (scipy then keeps int32 columns)."""
    return indptr.astype(np.int32) if indptr[-1] < 2 ** 31 else indptr


def counts_to_csr(pre: np.ndarray, post: np.ndarray, value: np.ndarray, n: int) -> sp.csr_matrix:
    """Signed counts of (pre, post) pairs -> an n x n float32 CSR with sorted columns; a pair
    listed more than once gets the sum of its entries."""
    pre = np.asarray(pre, np.int64)
    post = np.asarray(post, np.int64)
    if pre.size and (min(pre.min(), post.min()) < 0 or max(pre.max(), post.max()) >= n):
        raise ValueError("a neuron index is outside the completeness table")
    key = pre * n + post
    order = np.argsort(key, kind="stable")
    key = key[order]
    val = np.asarray(value, np.float32)[order]
    first = np.ones(key.size, bool)
    first[1:] = key[1:] != key[:-1]
    starts = np.flatnonzero(first)
    summed = np.add.reduceat(val, starts) if starts.size else val[:0]
    rows = key[starts] // n
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    w = sp.csr_matrix((summed.astype(np.float32), (key[starts] % n).astype(np.int32), _ptr(indptr)), shape=(n, n))
    w.has_sorted_indices = True
    return w


def _best_id_column(table, root_ids: np.ndarray) -> str:
    """The annotation column of root ids (``root_id`` or ``root_<version>``) that holds the most
    of the completeness table's ids; the first such column on a tie."""
    candidates = [c for c in table.columns if c == "root_id" or c.startswith("root_")]
    if not candidates:
        raise ValueError("the annotation table has no root-id column")
    matches = [int(np.isin(table[c].to_numpy(), root_ids).sum()) for c in candidates]
    return candidates[int(np.argmax(matches))]


def load_flywire(data_dir) -> Connectome:
    """The FlyWire v783 files in ``data_dir``: neuron order from the completeness table, signed
    counts from the connectivity table (repeated pairs summed), annotations joined on the
    root-id column that matches the most ids (first row of a repeated id).  Replaces
    flyzero/connectome.py:123-165."""
    import pandas as pd
    import pyarrow.parquet as pq

    data_dir = Path(data_dir)
    missing = [f for f in FILES if not (data_dir / f).exists()]
    if missing:
        raise FileNotFoundError(f"missing {missing} in {data_dir} (the FlyWire files are not downloaded here)")
    root_ids = pd.read_csv(data_dir / COMPLETENESS, usecols=[0]).iloc[:, 0].to_numpy(np.int64)
    n = root_ids.size
    table = pq.read_table(data_dir / CONNECTIVITY, columns=[PRE, POST, SIGNED_COUNT])
    weights = counts_to_csr(table.column(PRE).to_numpy(), table.column(POST).to_numpy(),
                            table.column(SIGNED_COUNT).to_numpy(), n)
    del table

    ann = pd.read_csv(data_dir / ANNOTATIONS, sep="\t", low_memory=False)
    ids = ann[_best_id_column(ann, root_ids)].to_numpy()
    row_of = {}
    for r, rid in enumerate(ids):          # the first annotation row of each id
        row_of.setdefault(rid, r)
    at = np.fromiter((row_of.get(rid, -1) for rid in root_ids), np.int64, n)
    found = at >= 0
    cols = {"root_id": root_ids}
    for c in ANNOTATION_COLUMNS:
        if c in ann:
            src = ann[c].to_numpy()
            out = np.full(n, np.nan, dtype=object if src.dtype == object else np.float64)
            out[found] = src[at[found]]
        else:
            out = np.full(n, np.nan)
        cols[c] = out
    neurons = pd.DataFrame(cols)
    for c in TEXT_COLUMNS:
        neurons[c] = neurons[c].astype("string")
    return Connectome(weights, neurons, name="FlyWire v783")


def corrected(conn: Connectome, fixes=ALL_FIXES) -> Connectome:
    """The corrections of flyzero/biology.py:38-55 as one rule per presynaptic neuron:

    * ``modulatory``: dopaminergic, serotonergic and octopaminergic neurons get no fast weight;
    * ``kenyon_cholinergic``: Kenyon cells are cholinergic whatever their predicted transmitter,
      so their outputs become excitatory, except onto other Kenyon cells (inhibitory, mAChR-B);
      with ``modulatory`` alone their outputs get no fast weight (they stay labelled
      dopaminergic).

    Only signs change or weights vanish; zeros are dropped from the CSR."""
    if not fixes:
        return conn
    KEEP, SILENT, KENYON = 0, 1, 2
    nt = conn.neurons["top_nt"].to_numpy(dtype=object, na_value="")
    kenyon = conn.neurons["cell_class"].to_numpy(dtype=object, na_value="") == "Kenyon_Cell"
    rule = np.full(conn.n, KEEP, np.int8)
    if "modulatory" in fixes:
        rule[np.isin(nt, MODULATORY)] = SILENT
    if "kenyon_cholinergic" in fixes:
        rule[kenyon] = KENYON
    elif "modulatory" in fixes:
        rule[kenyon] = SILENT
    w = sp.csr_matrix(conn.weights, dtype=np.float32, copy=True)
    w.sort_indices()
    per_edge = np.repeat(rule, np.diff(w.indptr))
    val = w.data
    val[per_edge == SILENT] = 0.0
    kc_edge = per_edge == KENYON
    val[kc_edge] = np.where(kenyon[w.indices[kc_edge]], -np.abs(val[kc_edge]), np.abs(val[kc_edge]))
    keep = val != 0.0
    rows = np.repeat(np.arange(conn.n, dtype=np.int64), np.diff(w.indptr))[keep]
    indptr = np.zeros(conn.n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=conn.n), out=indptr[1:])
    out = sp.csr_matrix((val[keep], w.indices[keep].astype(np.int32), _ptr(indptr)), shape=w.shape)
    out.has_sorted_indices = True
    return Connectome(out, conn.neurons, name=conn.name + " + corrections(" + ",".join(fixes) + ")")


def save_csr(conn: Connectome, path) -> Path:
    """The weights as int64 row pointers, int32 columns and int16 signed counts (.npy, so a load
    can memory-map them) and the neuron table as parquet, in directory ``path``."""
    w = sp.csr_matrix(conn.weights)
    w.sort_indices()
    counts = np.rint(w.data)
    if not np.array_equal(counts, w.data):
        raise ValueError("weights are not whole synapse counts")
    if np.abs(counts).max(initial=0) > np.iinfo(np.int16).max:
        raise ValueError("a synapse count does not fit int16")
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    np.save(path / "indptr.npy", w.indptr.astype(np.int64))
    np.save(path / "indices.npy", w.indices.astype(np.int32))
    np.save(path / "counts.npy", counts.astype(np.int16))
    conn.neurons.to_parquet(path / "neurons.parquet")
    (path / "name.txt").write_text(conn.name)
    return path


def load_csr(path) -> Connectome:
    """Inverse of ``save_csr``; the arrays are memory-mapped."""
    import pandas as pd

    path = Path(path)
    indptr = np.load(path / "indptr.npy", mmap_mode="r")
    indices = np.load(path / "indices.npy", mmap_mode="r")
    counts = np.load(path / "counts.npy", mmap_mode="r")
    n = len(indptr) - 1
    w = sp.csr_matrix((np.asarray(counts, np.float32), np.asarray(indices), np.asarray(indptr)), shape=(n, n))
    neurons = pd.read_parquet(path / "neurons.parquet")
    for col in TEXT_COLUMNS:
        if col in neurons:
            neurons[col] = neurons[col].astype("string")
    name = (path / "name.txt").read_text() if (path / "name.txt").exists() else os.fspath(path)
    return Connectome(w, neurons, name=name)
