"""Writes tests/golden/oracle_curve_<shape>_f<f>.json: per-epoch holdout RMSE of
the serial oracle (the reference trainer's CPU restatement) on the seeded
full-shape synthetic data, plain timeSVD++ form.  Slow (minutes per epoch):
run by hand, `python tests/golden/make_oracle_curves.py 64 20`.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from netflix_prize_amd import synth, tx  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main(f: int, epochs: int, seed: int = 1) -> None:
    tr, pr = synth.generate(synth.FULL, seed)
    cfg = tx.plain_form(tx.TxConfig(n_feat=f, n_epochs=epochs, n_time_bins=30, low_memory=True,
                                    reset_u_epoch=10_000))
    m = O.OracleModel(cfg, tr, pr)
    out = {"what": "serial oracle holdout RMSE per epoch", "shape": "full", "seed": seed, "f": f,
           "config": "plain form of the tsvdx4 lm set, 30 time bins, low_memory, no reset",
           "n_train": int(tr.n_ratings), "n_heldout": int(pr.n_ratings), "epochs": []}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"oracle_curve_full_f{f}.json")
    for e in range(1, epochs + 1):
        t0 = time.time()
        m.fit_epoch(e)
        out["epochs"].append({"epoch": e, "holdout_rmse": m.calc_rmse(), "sec": round(time.time() - t0, 1)})
        with open(path, "w") as fh:
            json.dump(out, fh, indent=1)
        print(out["epochs"][-1], flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]))
