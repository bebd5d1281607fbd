"""GPU KnnfModel (libtsvd_tx.so, include/tsvd_knn.h) against the reference's
KAT and the CPU oracle (needs an MI355X)."""
import numpy as np
import pytest

from netflix_prize_amd import knn
from oracle.knnf_ref import KnnfRef, sim_matrix
from tests.fixtures import synthetic, tiny
from tests.test_knnf_cpu import KNNF_KAT_RMSE, kat_factors, rmse

pytestmark = pytest.mark.gpu


def test_reference_kat_on_gpu(gpu):
    """tests/regressors_test.rs:403-422: probe RMSE 1.9509877627957246 on the
    tiny fixture with the 10 x 4 formula factors, k=3 (FP32 on both sides;
    within 1e-6)."""
    _, tr, pr = tiny()
    with knn.KnnfModel(tr, kat_factors(), knn.KnnfConfig(k=3, scaling=1.0, tau=0.025)) as m:
        r = rmse(m.predict(pr.user_idxs, pr.item_idxs, pr.dates), pr)
    assert r == pytest.approx(KNNF_KAT_RMSE, abs=1e-6), r


@pytest.mark.parametrize("f", [64, 1000])
def test_similarity_matches_fp64(gpu, f):
    """The MFMA cosine matrix against FP64 on 2,000 items (zero rows and
    anti-parallel rows included): |diff| <= 2e-6."""
    rng = np.random.default_rng(f)
    F = rng.normal(size=(2000, f)).astype(np.float32)
    F[7] = 0.0
    F[8] = -F[9]
    _, tr, _ = tiny()
    S = sim_matrix(F)
    with knn.KnnfModel(tr, F) as m:
        i = rng.integers(0, 2000, 4000)
        j = rng.integers(0, 2000, 4000)
        i[:3], j[:3] = [7, 8, 9], [3, 9, 9]
        g = m.similarity(i, j)
    np.testing.assert_allclose(g, S[i, j], rtol=0, atol=2e-6)
    assert g[0] == 0.0 and g[1] == 0.0


@pytest.mark.parametrize("k,scaling,tau", [(25, 1.0, 0.025), (40, 3.0, 0.05), (10, 0.5, 0.0)])
def test_predictions_match_oracle(gpu, k, scaling, tau):
    """10k x 2k x 1M set, random factors (f=32), 20,000 held-out queries:
    within 1e-5 of the FP64 oracle wherever the k-th and (k+1)-th weights are
    more than 1e-4 apart (relative; closer ties are open in the reference too,
    whose selection is unstable), and such ties are rare."""
    tr, pr = synthetic("small")
    rng = np.random.default_rng(3)
    F = rng.normal(size=(tr.n_items, 32)).astype(np.float32)
    q = rng.choice(pr.n_ratings, 20_000, replace=False)
    u, i, d = pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]
    ref = KnnfRef(tr, F, k=k, scaling=scaling, tau=tau)
    want = ref.predict(u, i, d)
    margin = ref.margins(u, i, d)
    with knn.KnnfModel(tr, F, knn.KnnfConfig(k=k, scaling=scaling, tau=tau)) as m:
        got = m.predict(u, i, d)
    clear = margin > 1e-4
    assert clear.mean() > 0.99, clear.mean()
    err = np.abs(got.astype(np.float64) - want)
    print(f"k={k}: max |diff| {err[clear].max():.2e} over {clear.sum()} clear queries")
    assert err[clear].max() <= 1e-5


def test_knns_reference_kat_on_gpu(gpu):
    """tests/regressors_test.rs:218-233: KnnsModel (support computed from the
    tiny fixture's training and probe sets, k=3, shrinkage 5) probe RMSE
    1.2349219913552398, within 1e-6."""
    from tests.test_knnf_cpu import KNNS_KAT_RMSE
    _, tr, pr = tiny()
    with knn.KnnsModel(tr, pr, knn.KnnsConfig(k=3, shrinkage=5.0, scaling=1.0, tau=0.025)) as m:
        r = rmse(m.predict(pr.user_idxs, pr.item_idxs, pr.dates), pr)
    assert r == pytest.approx(KNNS_KAT_RMSE, abs=1e-6), r


@pytest.mark.parametrize("k,scaling,tau", [(25, 1.0, 0.025), (40, 3.0, 0.05)])
def test_knns_matches_oracle(gpu, k, scaling, tau):
    """10k x 2k x 1M set: the GPU support matrix (atomic pair counts over
    training and held-out items) and shrunk-lift similarity within 1e-6
    relative of the FP64 oracle; predictions as for knnf."""
    from oracle.knnf_ref import KnnsRef
    tr, pr = synthetic("small")
    ref = KnnsRef(tr, pr, tr.n_items, k=k, shrinkage=100.0, scaling=scaling, tau=tau)
    rng = np.random.default_rng(5)
    q = rng.choice(pr.n_ratings, 20_000, replace=False)
    u, i, d = pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]
    with knn.KnnsModel(tr, pr, knn.KnnsConfig(k=k, shrinkage=100.0, scaling=scaling, tau=tau)) as m:
        a, b = rng.integers(0, tr.n_items, 5000), rng.integers(0, tr.n_items, 5000)
        s = m.similarity(a, b)
        got = m.predict(u, i, d)
    np.testing.assert_allclose(s, ref.S[a, b], rtol=1e-6, atol=1e-9)
    want = ref.predict(u, i, d)
    clear = ref.margins(u, i, d) > 1e-4
    assert clear.mean() > 0.99, clear.mean()
    err = np.abs(got.astype(np.float64) - want)
    print(f"knns k={k}: max |diff| {err[clear].max():.2e} over {clear.sum()} clear queries")
    assert err[clear].max() <= 1e-5


@pytest.mark.parametrize("scaling,tau", [(3.5, 0.0), (1.0, 0.025)])
def test_knn_stat_matches_oracle(gpu, scaling, tau):
    """KnnModel in the `*__knn-d` form over given stat tables (10k x 2k x 1M
    set; stat0 = the training support, stat1 = seeded values): similarities
    within 1e-6 relative and predictions within 1e-4 (values up to ~1) of the
    FP64 oracle on clear queries."""
    from oracle.knnf_ref import KnnStatRef
    tr, pr = synthetic("small")
    B = np.zeros((tr.n_users, tr.n_items), np.float32)
    B[tr.user_idxs, tr.item_idxs] = 1.0
    s0 = (B.T @ B).astype(np.float32)
    s1 = np.random.default_rng(9).normal(0, 50, s0.shape).astype(np.float32)
    ref = KnnStatRef(tr, s0, s1, k=15, shrinkage=100.0, scaling=scaling, tau=tau)
    rng = np.random.default_rng(6)
    q = rng.choice(pr.n_ratings, 20_000, replace=False)
    u, i, d = pr.user_idxs[q], pr.item_idxs[q], pr.dates[q]
    with knn.KnndModel(tr, s0, s1, knn.KnndConfig(k=15, shrinkage=100.0, scaling=scaling, tau=tau)) as m:
        a, b = rng.integers(0, tr.n_items, 5000), rng.integers(0, tr.n_items, 5000)
        s = m.similarity(a, b)
        got = m.predict(u, i, d)
    np.testing.assert_allclose(s, ref.S[a, b], rtol=1e-6, atol=1e-9)
    want = ref.predict(u, i, d)
    clear = ref.margins(u, i, d) > 1e-4
    assert clear.mean() > 0.99, clear.mean()
    err = np.abs(got.astype(np.float64) - want)
    print(f"knn-d scaling={scaling}: max |diff| {err[clear].max():.2e} over {clear.sum()} clear queries")
    assert err[clear].max() <= 1e-4
