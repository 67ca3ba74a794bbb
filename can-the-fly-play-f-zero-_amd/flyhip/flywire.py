"""FlyWire ingestion: the files of flyzero/connectome.py:1-19 -> a ``Connectome`` (signed synapse
counts, pre x post CSR, and the neuron table), the biological corrections of
flyzero/biology.py:38-55, and an on-disk CSR cache (int64 row pointers, int32 columns, int16
counts, memory-mapped on load) so that the 15 M-edge matrix is read once.

The files are not in this build's environment (the reference downloads them,
flyzero/connectome.py:52-65); nothing here fetches anything.  ``BatchBrain(conn.weights, ...)``
and every ``find`` caller work with a ``Connectome`` exactly as with the synthetic fly.
"""

from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.sparse as sp

FILES = ("Completeness_783.csv", "Connectivity_783.parquet", "neuron_annotations.tsv")
ANNOTATION_COLUMNS = ("super_class", "cell_class", "cell_type", "hemibrain_type", "side", "top_nt", "pos_x", "pos_y",
                      "pos_z")
TEXT_COLUMNS = ("super_class", "cell_class", "cell_type", "hemibrain_type", "side", "top_nt")
MODULATORY = ("dopamine", "serotonin", "octopamine")
ALL_FIXES = ("modulatory", "kenyon_cholinergic")
PHOTORECEPTOR_RE = re.compile(r"^R[1-8]([a-z]|-6)?$|^R1-6$", re.IGNORECASE)


@dataclass
class Connectome:
    """flyzero/connectome.py:68-121: ``weights`` (pre x post CSR of signed synapse counts) and
    ``neurons`` (one row per neuron in index order: root_id, cell_type, hemibrain_type, side,
    top_nt, super_class, cell_class, pos_x/y/z)."""

    weights: sp.csr_matrix
    neurons: object   # pandas.DataFrame
    name: str = "connectome"
    _type_cache: dict = field(default_factory=dict, repr=False)

    @property
    def n(self) -> int:
        return self.weights.shape[0]

    def find(self, cell_type: str, side: str | None = None) -> np.ndarray:
        """Neurons whose cell type (or hemibrain type) matches exactly; a trailing ``*`` matches a
        prefix (``DNg02*``).  Optionally one side."""
        key = cell_type.lower()
        if key not in self._type_cache:
            ct = self.neurons["cell_type"].str.lower()
            if key.endswith("*"):
                m = ct.str.startswith(key[:-1])
            else:
                m = ct.eq(key)
                if "hemibrain_type" in self.neurons:
                    m = m | self.neurons["hemibrain_type"].str.lower().eq(key)
            self._type_cache[key] = m.to_numpy(dtype=bool, na_value=False)
        m = self._type_cache[key]
        if side is not None:
            m = m & self.neurons["side"].eq(side).to_numpy(dtype=bool, na_value=False)
        return np.flatnonzero(m)

    def photoreceptors(self) -> np.ndarray:
        """flyzero/connectome.py:103-111."""
        m = self.neurons["cell_type"].fillna("").str.match(PHOTORECEPTOR_RE)
        if m.sum() == 0:
            cls = self.neurons["cell_class"].fillna("").str.lower()
            m = self.neurons["super_class"].eq("sensory") & cls.str.contains("photoreceptor|retin|visual")
        return np.flatnonzero(m.to_numpy(dtype=bool, na_value=False))


def load_flywire(data_dir) -> Connectome:
    """The three files of FlyWire materialization 783 (flyzero/connectome.py:123-165): neuron order
    from the completeness table, signed counts from the connectivity table (repeated pairs
    summed), annotations joined on the root-id column that matches the most ids."""
    import pandas as pd

    data_dir = Path(data_dir)
    missing = [f for f in FILES if not (data_dir / f).exists()]
    if missing:
        raise FileNotFoundError(f"missing {missing} in {data_dir} (the FlyWire files are not downloaded here)")
    comp = pd.read_csv(data_dir / FILES[0], index_col=0)
    root_ids = comp.index.to_numpy(dtype=np.int64)
    n = len(root_ids)
    con = pd.read_parquet(data_dir / FILES[1],
                          columns=["Presynaptic_Index", "Postsynaptic_Index", "Excitatory x Connectivity"])
    weights = sp.csr_matrix((con["Excitatory x Connectivity"].to_numpy(np.float32),
                             (con["Presynaptic_Index"].to_numpy(), con["Postsynaptic_Index"].to_numpy())),
                            shape=(n, n))
    weights.sort_indices()
    del con
    ann = pd.read_csv(data_dir / FILES[2], sep="\t", low_memory=False)
    id_cols = [c for c in ann.columns if c == "root_id" or c.startswith("root_")]
    best = max(id_cols, key=lambda c: pd.Index(ann[c]).isin(root_ids).sum())
    ann = ann.drop_duplicates(best).set_index(best)
    neurons = pd.DataFrame({"root_id": root_ids})
    for col in ANNOTATION_COLUMNS:
        neurons[col] = ann[col].reindex(root_ids).to_numpy() if col in ann else np.nan
    for col in TEXT_COLUMNS:
        neurons[col] = neurons[col].astype("string")
    return Connectome(weights, neurons, name="FlyWire v783")


def corrected(conn: Connectome, fixes=ALL_FIXES) -> Connectome:
    """flyzero/biology.py:38-55: modulatory (dopamine, serotonin, octopamine) presynaptic neurons
    get zero fast weight; Kenyon cells are cholinergic (excitatory) except onto other Kenyon
    cells (inhibitory).  Only signs change or weights vanish; zeros are removed."""
    if not fixes:
        return conn
    w = conn.weights.tocoo()
    data = w.data.copy()
    nt = conn.neurons["top_nt"].fillna("").to_numpy().astype(str)
    kc = conn.neurons["cell_class"].eq("Kenyon_Cell").to_numpy(dtype=bool, na_value=False)
    pre_kc = kc[w.row]
    if "modulatory" in fixes:
        data[np.isin(nt[w.row], MODULATORY) & ~pre_kc] = 0.0
    if "kenyon_cholinergic" in fixes:
        mag = np.abs(data[pre_kc])
        data[pre_kc] = np.where(kc[w.col[pre_kc]], -mag, mag)
    elif "modulatory" in fixes:
        data[pre_kc] = 0.0
    weights = sp.csr_matrix((data, (w.row, w.col)), shape=w.shape)
    weights.eliminate_zeros()
    weights.sort_indices()
    return Connectome(weights, conn.neurons, name=conn.name + " + corrections(" + ",".join(fixes) + ")")


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
