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


_SYNTH_CACHE = {}


def synthetic(name="tiny_synth", seed=7):
    """Seeded synthetic (train, heldout) sets, cached per process.

    tiny_synth: 600 users x 150 items x 30k ratings (serial-schedule checks)
    small:      the 10k x 2k x 1M benchmark shape
    """
    from netflix_prize_amd import synth
    key = (name, seed)
    if key not in _SYNTH_CACHE:
        if name == "tiny_synth":
            spec = synth.SynthSpec(n_users=600, n_items=150, n_ratings=30_000, user_median=30, user_max=150,
                                   item_sigma=1.0, item_min=3, item_max=5_000)
        elif name == "small":
            spec = synth.SMALL
        else:
            raise KeyError(name)
        _SYNTH_CACHE[key] = synth.generate(spec, seed)
    return _SYNTH_CACHE[key]


def lm_config(**kw):
    """The tsvdx4 lm hyper-parameters (src/bin/tsvdx4-new.rs:130-154) with
    per-test overrides, as a TxConfig."""
    from netflix_prize_amd.tx import TxConfig
    return TxConfig(**kw)


_FULL = {}


def full_data(seed=1):
    """The full-shape synthetic set (480,189 x 17,770, 100,480,507 ratings),
    generated once per process and shared by every test that needs it."""
    if seed not in _FULL:
        from netflix_prize_amd import synth
        _FULL[seed] = synth.generate(synth.FULL, seed)
    return _FULL[seed]
