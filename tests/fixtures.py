"""Shared test data: the reference's tiny fixture and seeded synthetic sets."""
import json
import os
import types

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture():
    with open(os.path.join(GOLDEN, "tiny_fixture.json")) as fh:
        return json.load(fh)


def dataset_from_rows(rows, n_users, n_items, name="ds"):
    """Dataset built directly from (user, item, rating, date) rows, the way the
    reference's tests build theirs (tests/helpers.rs:9-33)."""
    a = np.asarray(rows, dtype=np.int64).reshape(-1, 4)
    d = types.SimpleNamespace()
    d.name = name
    d.user_idxs = a[:, 0].astype(np.int32)
    d.item_idxs = a[:, 1].astype(np.int32)
    d.raw_ratings = a[:, 2].astype(np.int8)
    d.residuals = a[:, 2].astype(np.float32)
    d.dates = a[:, 3].astype(np.int16)
    d.is_test = np.zeros(len(a), np.int8)
    d.user_cnts = np.bincount(d.user_idxs, minlength=n_users).astype(np.int32)
    d.item_cnts = np.bincount(d.item_idxs, minlength=n_items).astype(np.int32)
    d.n_users = n_users
    d.n_items = n_items
    d.n_ratings = len(a)
    return d


def tiny():
    fx = load_fixture()
    tr = dataset_from_rows(fx["train"], fx["n_users"], fx["n_items"], "tiny_train")
    pr = dataset_from_rows(fx["probe"], fx["n_users"], fx["n_items"], "tiny_probe")
    return fx, tr, pr
