"""Writes tests/golden/oracle_curve_full_f<f>.json: per-epoch holdout RMSE of
the serial oracle (the reference trainer's CPU restatement) on the seeded
full-shape synthetic data, plain timeSVD++ form; with the form argument
`job60`, oracle_curve_full_job60.json: the tsvdx4-60 job's model (the full lm
set with day terms, low_memory = false, f = 60, 29 time bins, seed 42;
src/bin/tsvdx4-new.rs:18, 118-154).  Slow (one to several minutes per epoch):
run by hand, `python tests/golden/make_oracle_curves.py 64 20` or
`python tests/golden/make_oracle_curves.py 60 20 job60`.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from netflix_prize_amd import synth, tx  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main(f: int, epochs: int, form: str = "plain", seed: int = 1) -> None:
    tr, pr = synth.generate(synth.FULL, seed)
    if form == "job60":
        cfg = tx.TxConfig(n_feat=f, n_epochs=epochs, seed=42, n_time_bins=29, low_memory=False)
        what, name = "the tsvdx4-60 job's model: lm set, day terms (low_memory false)", "oracle_curve_full_job60.json"
    else:
        cfg = tx.plain_form(tx.TxConfig(n_feat=f, n_epochs=epochs, n_time_bins=30, low_memory=True,
                                        reset_u_epoch=10_000))
        what, name = "plain form of the tsvdx4 lm set, 30 time bins, low_memory, no reset", f"oracle_curve_full_f{f}.json"
    m = O.OracleModel(cfg, tr, pr)
    out = {"what": "serial oracle holdout RMSE per epoch", "shape": "full", "seed": seed, "f": f,
           "config": what, "n_train": int(tr.n_ratings), "n_heldout": int(pr.n_ratings), "epochs": []}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), name)
    for e in range(1, epochs + 1):
        t0 = time.time()
        m.fit_epoch(e)
        out["epochs"].append({"epoch": e, "holdout_rmse": m.calc_rmse(), "sec": round(time.time() - t0, 1)})
        with open(path, "w") as fh:
            json.dump(out, fh, indent=1)
        print(out["epochs"][-1], flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]), *(sys.argv[3:4]))
