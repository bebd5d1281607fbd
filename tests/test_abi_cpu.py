"""CPU-side checks of libtsvd_tx.so: it loads, exports every entry point the
header declares, rejects bad input before touching a GPU, and its host-only
RNG helpers reproduce the reference streams (checked against the oracle)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from netflix_prize_amd import tx
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "tsvd_tx.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(tx_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    names = _declared()
    assert len(names) >= 15
    L = C.CDLL(tx.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_library_exports_every_declared_knn_entry_point():
    text = open(os.path.join(ROOT, "include", "tsvd_knn.h")).read()
    names = sorted(set(re.findall(r"^int\s+(knn\w+)\(", text, re.M)))
    assert {"knnf_create", "knns_create", "knn3_create", "knn3_predict", "knn3_weights"} <= set(names)
    L = C.CDLL(tx.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_encoder_library_exports_every_declared_entry_point():
    text = open(os.path.join(ROOT, "include", "tsvd_enc.h")).read()
    names = sorted(set(re.findall(r"^(?:int|const char\*)\s+(enc_\w+)\(", text, re.M)))
    assert len(names) >= 7
    from netflix_prize_amd import encoder
    L = C.CDLL(encoder.ENC_LIB_PATH)
    for n in names:
        assert hasattr(L, n), n


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors have the sizes and field offsets the C header gives."""
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "tsvd_tx.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(tx_config), sizeof(tx_ratings),'
        ' sizeof(tx_schedule), sizeof(tx_epoch_stats), offsetof(tx_config, beta),'
        ' offsetof(tx_config, th_init), offsetof(tx_config, seq_epochs)); return 0;}\n')
    exe = tmp_path / "probe"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(tx.CConfig), C.sizeof(tx.CRatings), C.sizeof(tx.CSchedule), C.sizeof(tx.CEpochStats),
            tx.CConfig.beta.offset, tx.CConfig.th_init.offset, tx.CConfig.seq_epochs.offset]
    assert got == want


@pytest.mark.parametrize("seed,sigma", [(42, 0.005), (8003, 0.0), (2400, 1.0)])
def test_rand_normal_matches_reference_stream(seed, sigma):
    np.testing.assert_array_equal(tx.rand_array(seed, sigma, 50_000), O.normal_fill(seed, sigma, 50_000))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 13, 14, 1000, 480_189])
def test_get_users_matches_reference_shuffle(n):
    for epoch in (1, 7):
        np.testing.assert_array_equal(tx.get_users(n, True, 42, epoch), O.get_users(n, True, 42, epoch))
    np.testing.assert_array_equal(tx.get_users(n, False, 42, 1), np.arange(n))


def test_last_error_and_bad_arguments():
    L = tx.lib()
    assert L.tx_get_users(-1, 1, 0, 1, None) != 0
    assert "bad arguments" in L.tx_last_error().decode()
    assert L.tx_rand_normal(1, C.c_float(-1.0), 4, None) != 0


def test_calc_user_offsets_rejects_unsorted():
    from tests.fixtures import tiny
    _, tr, pr = tiny()
    off = tx.calc_user_offsets(tr)
    assert np.array_equal(np.diff(off), tr.user_cnts)
    with pytest.raises(tx.TxError, match="not user-sorted"):
        tx.calc_user_offsets(pr)


def test_gbias_equals_fp64_mean():
    from tests.fixtures import tiny
    _, tr, _ = tiny()
    assert tx.calc_gbias(tr) == np.float32(np.mean(tr.residuals.astype(np.float64)))
