"""FlyWire ingestion on a hand-made dataset with the real files' schema (the files themselves are
not available here): loading (flyzero/connectome.py:123-165), ``find``, the corrections of
flyzero/biology.py:38-55 and the on-disk CSR round trip."""

import numpy as np
import pandas as pd
import pytest

from flyhip.flywire import Connectome, corrected, load_csr, load_flywire, save_csr

ROOTS = [720575940600000000 + k for k in range(8)]
# index: 0 DNg02_a L, 1 DNg02_b R, 2 DNa02 L (hemibrain "DNa02"), 3 KC, 4 KC, 5 dopaminergic,
# 6 photoreceptor R7p, 7 unannotated
ANN = {
    "cell_type": ["DNg02_a", "DNg02_b", "DNa02", "KCg-m", "KCab", "PAM01", "R7p", None],
    "hemibrain_type": [None, None, "DNa02", None, None, None, None, None],
    "side": ["left", "right", "left", "left", "right", "left", "right", None],
    "top_nt": ["acetylcholine", "acetylcholine", "gaba", "dopamine", "dopamine", "dopamine", "histamine", None],
    "cell_class": [None, None, None, "Kenyon_Cell", "Kenyon_Cell", "DAN", "photoreceptor", None],
    "super_class": ["descending"] * 3 + ["central"] * 3 + ["sensory", None],
}
# (pre, post, signed count); the pair (0, 1) appears twice and is summed
EDGES = [(0, 1, 3), (0, 1, 2), (0, 2, 4), (3, 4, 5), (3, 0, 6), (4, 3, 2), (5, 0, 7), (5, 3, -2), (6, 7, 1),
         (2, 5, -3)]


@pytest.fixture
def dataset(tmp_path):
    pd.DataFrame({"Completed": [True] * 8}, index=pd.Index(ROOTS, name="")).to_csv(tmp_path / "Completeness_783.csv")
    pre, post, cnt = zip(*EDGES)
    pd.DataFrame({"Presynaptic_Index": pre, "Postsynaptic_Index": post, "Excitatory x Connectivity": cnt,
                  "Presynaptic_ID": [ROOTS[p] for p in pre]}).to_parquet(tmp_path / "Connectivity_783.parquet")
    ann = pd.DataFrame(ANN)
    ann["root_783"] = ROOTS
    ann["root_630"] = [r - 5 for r in ROOTS]         # an older materialization: fewer matches
    ann["pos_x"], ann["pos_y"], ann["pos_z"] = np.arange(8.0), np.arange(8.0) * 2, np.arange(8.0) * 3
    ann = ann.iloc[::-1]                              # row order differs from the neuron order
    ann.to_csv(tmp_path / "neuron_annotations.tsv", sep="\t", index=False)
    return tmp_path


def test_load_builds_the_signed_count_matrix_and_table(dataset):
    conn = load_flywire(dataset)
    dense = np.zeros((8, 8))
    for p, q, c in EDGES:
        dense[p, q] += c
    np.testing.assert_array_equal(conn.weights.toarray(), dense)
    assert conn.weights.has_sorted_indices and conn.n == 8
    assert list(conn.neurons["root_id"]) == ROOTS
    assert conn.neurons["cell_type"].iloc[3] == "KCg-m" and pd.isna(conn.neurons["cell_type"].iloc[7])
    np.testing.assert_array_equal(conn.neurons["pos_y"].to_numpy(), np.arange(8.0) * 2)


def test_find_prefix_side_and_hemibrain_type(dataset):
    conn = load_flywire(dataset)
    assert list(conn.find("DNg02*")) == [0, 1]
    assert list(conn.find("DNg02*", "right")) == [1]
    assert list(conn.find("dna02")) == [2]                   # case-insensitive, also the hemibrain type
    assert list(conn.find("MDN")) == []
    assert list(conn.photoreceptors()) == [6]


def test_corrections_zero_modulators_and_resign_kenyon_cells(dataset):
    conn = corrected(load_flywire(dataset))
    w = conn.weights.toarray()
    assert (w[5] == 0).all()                                 # dopaminergic, not a Kenyon cell: no fast weight
    assert w[3, 4] == -5 and w[4, 3] == -2                   # KC -> KC inhibitory (mAChR-B)
    assert w[3, 0] == 6                                      # KC -> other: excitatory
    assert w[0, 1] == 5 and w[2, 5] == -3                    # everything else unchanged
    assert conn.weights.nnz == np.count_nonzero(w)           # zeros removed
    only_mod = corrected(load_flywire(dataset), ("modulatory",)).weights.toarray()
    assert (only_mod[3] == 0).all() and (only_mod[4] == 0).all()   # KCs still labelled dopaminergic


def test_csr_cache_round_trip(dataset, tmp_path):
    conn = corrected(load_flywire(dataset))
    path = save_csr(conn, tmp_path / "cache")
    back = load_csr(path)
    assert isinstance(back, Connectome) and back.name == conn.name
    np.testing.assert_array_equal(back.weights.toarray(), conn.weights.toarray())
    assert back.weights.indices.dtype == np.int32
    assert list(back.find("DNg02*", "left")) == [0]
    big = Connectome(conn.weights * 20000, conn.neurons)
    with pytest.raises(ValueError, match="int16"):
        save_csr(big, tmp_path / "big")
    frac = Connectome(conn.weights * 0.5, conn.neurons)
    with pytest.raises(ValueError, match="whole"):
        save_csr(frac, tmp_path / "frac")


def test_missing_files_are_reported(tmp_path):
    with pytest.raises(FileNotFoundError, match="missing"):
        load_flywire(tmp_path)


def test_repeated_annotation_ids_take_the_first_row_and_bad_indices_are_refused(dataset):
    import pandas as pd

    from flyhip.flywire import counts_to_csr

    ann = pd.read_csv(dataset / "neuron_annotations.tsv", sep="\t")
    extra = ann[ann["root_783"] == ROOTS[0]].copy()
    extra["cell_type"] = "not_this_one"
    pd.concat([ann, extra]).to_csv(dataset / "neuron_annotations.tsv", sep="\t", index=False)
    conn = load_flywire(dataset)
    assert conn.neurons["cell_type"].iloc[0] == "DNg02_a"
    with pytest.raises(ValueError, match="outside"):
        counts_to_csr(np.array([0, 9]), np.array([1, 2]), np.array([1.0, 2.0]), 8)
    w = counts_to_csr(np.array([2, 0, 2, 0]), np.array([1, 3, 1, 1]), np.array([1.0, 2.0, 4.0, -1.0]), 4)
    np.testing.assert_array_equal(w.toarray(), [[0, -1, 0, 2], [0, 0, 0, 0], [0, 5, 0, 0], [0, 0, 0, 0]])
    assert w.has_sorted_indices and w.indices.dtype == np.int32
