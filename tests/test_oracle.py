"""Pins the CPU oracle against the reference's own known answers and fixtures, and
against independent implementations (scipy, scikit-learn) and invariants. CPU only."""
import numpy as np
import pytest
from scipy.spatial.distance import cdist

EPS = np.finfo(np.float64).eps


@pytest.mark.parametrize("case", ["pairwise_fixed_xy", "pairwise_fixed_xx"])
@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "manhattan", "cosine",
                                    "minkowski", "inner_product", "euclidean_gemm",
                                    "sqeuclidean_gemm"])
def test_pairwise_known_answers(known, oracle, case, metric):
    c = known[case]
    X = np.array(c["X_row"])
    Y = np.array(c["Y_row"]) if "Y_row" in c else None
    key = {"euclidean": "sqeuclidean", "euclidean_gemm": "sqeuclidean",
           "sqeuclidean_gemm": "sqeuclidean"}.get(metric, metric)
    exp = np.array(c["expected"][key])
    if metric in ("euclidean", "euclidean_gemm"):
        exp = np.sqrt(exp)
    D = oracle.pairwise(X, Y, metric, p=c["p"])
    np.testing.assert_allclose(D, exp, rtol=0, atol=100 * EPS * max(1, np.abs(exp).max()))


@pytest.mark.parametrize("metric,sp", [("euclidean", "euclidean"), ("sqeuclidean", "sqeuclidean"),
                                       ("manhattan", "cityblock"), ("cosine", "cosine"),
                                       ("minkowski", "minkowski")])
def test_pairwise_vs_scipy(oracle, metric, sp):
    rng = np.random.default_rng(7)
    X = rng.uniform(-10, 10, (37, 13))
    Y = rng.uniform(-10, 10, (29, 13))
    kw = {"p": 3.5} if sp == "minkowski" else {}
    exp = cdist(X, Y, sp, **kw)
    D = oracle.pairwise(X, Y, metric, p=3.5)
    np.testing.assert_allclose(D, exp, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.pairwise(X, None, metric, p=3.5),
                               cdist(X, X, sp, **kw), rtol=1e-12, atol=1e-12)


def test_pairwise_invariants(oracle):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((20, 6))
    for metric in ("euclidean", "manhattan", "cosine", "sqeuclidean_gemm"):
        D = oracle.pairwise(X, None, metric)
        assert np.all(np.diag(D) == 0)
        np.testing.assert_allclose(D, D.T, rtol=0, atol=1e-13)
    np.testing.assert_allclose(oracle.pairwise(X, X, "inner_product"), X @ X.T, rtol=1e-13)


def test_knn_example(known, oracle):
    c = known["knn_example"]
    Xt = np.array(c["X_train_colmajor"]).reshape((c["n_features"], c["n_samples"])).T
    Xq = np.array(c["X_test_colmajor"]).reshape((c["n_features"], c["n_queries"])).T
    exp_d = np.array(c["k4_dist_colmajor"]).reshape((4, 3)).T
    for k, key in ((3, "k3_ind_colmajor"), (4, "k4_ind_colmajor")):
        ind, dist = oracle.knn(Xt, Xq, k)
        exp = np.array(c[key]).reshape((k, c["n_queries"])).T
        # Query 2 has training rows 0 and 2 tied at squared distance 38 in slot 4: the
        # reference keeps the later index, the oracle orders ties by index.
        _assert_knn_match(ind, dist, exp, exp_d[:, :k], 500 * EPS, Xt, Xq)
    assert ind[2, 3] == 0 and exp[2, 3] == 2


def _assert_knn_match(ind, dist, exp_ind, exp_dist, atol, Xt, Xq):
    """Distances must match; an index may differ only where the two candidates are at the
    same distance from the query (a tie, whose order the reference leaves to its schedule)."""
    np.testing.assert_allclose(dist, exp_dist, rtol=0, atol=atol)
    for q, j in np.argwhere(ind != exp_ind):
        da = np.linalg.norm(Xt[ind[q, j]] - Xq[q])
        db = np.linalg.norm(Xt[exp_ind[q, j]] - Xq[q])
        assert abs(da - db) <= atol, f"query {q} slot {j}: index {ind[q, j]} vs {exp_ind[q, j]}"


@pytest.mark.parametrize("name", ["2092x5", "2125x20"])
def test_knn_csv_fixtures(nn_csv, oracle, name):
    z = nn_csv[name]
    ind, dist = oracle.knn(z["X_train"], z["X_test"], 5)
    _assert_knn_match(ind, dist, z["k_ind"], z["k_dist"], 500 * EPS * 10, z["X_train"],
                      z["X_test"])


def test_knn_vs_sklearn(oracle):
    from sklearn.neighbors import NearestNeighbors
    rng = np.random.default_rng(11)
    Xt = rng.standard_normal((500, 9))
    Xq = rng.standard_normal((60, 9))
    for metric, skm in (("euclidean", "euclidean"), ("manhattan", "manhattan"),
                        ("cosine", "cosine")):
        nn = NearestNeighbors(n_neighbors=7, algorithm="brute", metric=skm).fit(Xt)
        sd, si = nn.kneighbors(Xq)
        ind, dist = oracle.knn(Xt, Xq, 7, metric)
        np.testing.assert_array_equal(ind, si)
        np.testing.assert_allclose(dist, sd, rtol=1e-10, atol=1e-12)


def _colmajor(v, rows, cols):
    return np.array(v).reshape((cols, rows)).T


@pytest.mark.parametrize("idx", range(5))
def test_kmeans_known_answers(known, oracle, idx):
    c = known["kmeans"][idx]
    n, d, k = c["n_samples"], c["n_features"], c["n_clusters"]
    A = _colmajor(c["A_colmajor"], n, d)
    C0 = _colmajor(c["C_colmajor"], k, d)
    init = c.get("init", "supplied")
    r = oracle.kmeans(A, k, init, C0 if init == "supplied" else None, max_iter=c["max_iter"],
                      tol=c["tol"], empty_clusters=c["empty_clusters"], seed=78)
    assert r["status"] == 0
    tol = {"100*sqrt(eps)": 100 * np.sqrt(EPS), "10000*sqrt(eps)": 1e4 * np.sqrt(EPS),
           "10*eps": 10 * EPS}[c["tol_check"]]
    tol = max(tol, 1e-6)  # the expected values are printed to ~7 significant digits
    assert r["n_iter"] == c["rinfo"][3]
    assert abs(r["inertia"] - c["rinfo"][4]) <= tol
    np.testing.assert_allclose(r["centres"], _colmajor(c["centres_colmajor"], k, d), rtol=0,
                               atol=tol)
    np.testing.assert_array_equal(r["labels"], c["labels"])
    X = _colmajor(c["X_colmajor"], c["m_samples"], d)
    np.testing.assert_allclose(oracle.pairwise(X, r["centres"], "euclidean"),
                               _colmajor(c["X_transform_colmajor"], c["m_samples"], k), rtol=0,
                               atol=tol)
    Y = _colmajor(c["Y_colmajor"], c["k_samples"], d)
    pred = np.argmin(oracle.pairwise(Y, r["centres"], "sqeuclidean"), axis=1)
    np.testing.assert_array_equal(pred, c["Y_labels"])


def test_kmeans_error_mode(known, oracle):
    c = known["kmeans"][2]
    A = _colmajor(c["A_colmajor"], 5, 2)
    C0 = _colmajor(c["C_colmajor"], 2, 2)
    assert oracle.kmeans(A, 2, "supplied", C0, max_iter=30, empty_clusters="error")["status"] == 900


def test_kmeans_vs_sklearn(oracle):
    from sklearn.cluster import KMeans
    rng = np.random.default_rng(5)
    centres = rng.uniform(-10, 10, (6, 4))
    X = np.concatenate([c + rng.standard_normal((200, 4)) for c in centres])
    C0 = X[rng.choice(len(X), 6, replace=False)]
    r = oracle.kmeans(X, 6, "supplied", C0, max_iter=100, tol=0.0)
    sk = KMeans(n_clusters=6, init=C0, n_init=1, algorithm="lloyd", max_iter=100, tol=0.0).fit(X)
    np.testing.assert_array_equal(r["labels"], sk.labels_)
    np.testing.assert_allclose(r["centres"], sk.cluster_centers_, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(r["inertia"], sk.inertia_, rtol=1e-10)


def test_kmeanspp_quality_and_determinism(oracle):
    rng = np.random.default_rng(1)
    centres = rng.uniform(-20, 20, (8, 3))
    X = np.concatenate([c + 0.3 * rng.standard_normal((100, 3)) for c in centres])
    a = oracle.kmeans(X, 8, "k-means++", seed=42, max_iter=50)
    b = oracle.kmeans(X, 8, "k-means++", seed=42, max_iter=50)
    np.testing.assert_array_equal(a["centres"], b["centres"])
    # well separated blobs: k-means++ finds every blob, so the inertia is close to n*d*0.09
    assert a["inertia"] < 1.5 * 800 * 3 * 0.09

def _random_case(known, c):
    from _kmeans_cases import base_data
    A, _, _, _ = base_data(known, c["base"])
    return A


@pytest.mark.parametrize("i", range(17))
def test_kmeans_reference_random_inits(known, oracle, i):
    """The reference's checks of random initialisations (kmeans_public.cpp:272-276): the
    best inertia of n_init runs is at most the stated bound."""
    from _kmeans_cases import check_tol
    c = known["kmeans_random"][i]
    A = _random_case(known, c)
    r = oracle.kmeans(A, c["n_clusters"], c["init"], n_init=c["n_init"],
                      max_iter=c["max_iter"], seed=c["seed"], algorithm=c["algorithm"])
    assert r["status"] in (0, 21)
    assert r["inertia"] <= c["max_inertia"] + check_tol(c.get("check_tol", "100*sqrt(eps)"),
                                                        np.float64)


@pytest.mark.parametrize("i", range(7))
def test_kmeans_reference_fixed_random_inits(known, oracle, i):
    """Random initialisations whose answers the reference fixes: all-zero data (every
    initialisation coincides) and the 1 x 1 matrix."""
    c = known["kmeans_fixed_random_init"][i]
    n, d, k = c["n_samples"], c["n_features"], c["n_clusters"]
    A = _colmajor(c["A_colmajor"], n, d)
    r = oracle.kmeans(A, k, c["init"], n_init=c["n_init"], max_iter=c["max_iter"],
                      tol=c["tol"], seed=abs(c["seed"]), algorithm=c["algorithm"])
    assert r["status"] == 0
    assert [n, d, k, r["n_iter"], r["inertia"], r["lp_n_iter"]] == c["rinfo"]
    np.testing.assert_array_equal(r["centres"], _colmajor(c["centres_colmajor"], k, d))
    np.testing.assert_array_equal(r["labels"], c["labels"])


@pytest.mark.parametrize("i", range(3))
def test_kmeans_reference_elkan(known, oracle, i):
    """Elkan's algorithm returns the answers the reference expects of its base cases."""
    from _kmeans_cases import base_data, check_tol
    c = known["kmeans_elkan"][i]
    A, C0, _, b = base_data(known, c["base"])
    k, d = C0.shape
    r = oracle.kmeans(A, k, "supplied", C0, max_iter=b["max_iter"], tol=b["tol"],
                      empty_clusters=b["empty_clusters"], algorithm="elkan")
    assert r["status"] == 0
    tol = max(check_tol(b["tol_check"], np.float64), 1e-6)
    assert r["n_iter"] == b["rinfo"][3]
    assert abs(r["inertia"] - b["rinfo"][4]) <= tol
    np.testing.assert_allclose(r["centres"], _colmajor(b["centres_colmajor"], k, d), atol=tol)
    np.testing.assert_array_equal(r["labels"], b["labels"])


@pytest.mark.parametrize("i", range(2))
@pytest.mark.parametrize("dtype", ["double", "float"])
def test_kmeans_reference_mixed_precision(known, oracle, i, dtype):
    """Iteration counts of mixed-precision refinement (kmeans_test_data.hpp:944-969): the
    lower-precision phase runs in float for double data and in half for float data."""
    from _kmeans_cases import base_data
    c = known["kmeans_mixed"][i]
    A, C0, _, b = base_data(known, c["base"])
    if dtype == "float":
        A, C0 = A.astype(np.float32).astype(np.float64), C0.astype(np.float32).astype(np.float64)
    r = oracle.kmeans(A, C0.shape[0], "supplied", C0, max_iter=b["max_iter"], tol=b["tol"],
                      mixed_precision={"double": "float", "float": "half"}[dtype],
                      lp_tol=c["lp_tol"], lp_max_iter=c["lp_max_iter"])
    assert r["status"] == 0
    assert r["lp_n_iter"] == c["lp_n_iter"]
    assert r["n_iter"] == c["n_iter"][dtype]
    # the refinement ends at the answer of the working precision alone
    np.testing.assert_array_equal(r["labels"], b["labels"])


def test_kmeans_mixed_precision_incompatible_tolerance(oracle):
    X = np.arange(12.0).reshape(6, 2)
    assert oracle.kmeans(X, 2, "supplied", X[:2], tol=1e-2, mixed_precision="float",
                         lp_tol=1e-2)["status"] == 41


def test_round_lp_matches_ieee(oracle):
    """Rounding to float and to half (through float), round to nearest even, against numpy."""
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-9, 6, 2000),
                        [65504.0, 65519.99, 65520.0, 1e5, 6.1e-5, 2.0 ** -24, 3 * 2.0 ** -25,
                         2.0 ** -26, 0.0, -0.0, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11]])
    np.testing.assert_array_equal(oracle.round_lp(v, 1), v.astype(np.float32).astype(np.float64))
    with np.errstate(over="ignore"):
        half = v.astype(np.float32).astype(np.float16).astype(np.float64)
    np.testing.assert_array_equal(oracle.round_lp(v, 2), half)


@pytest.mark.parametrize("seed", [0, 7, 78])
@pytest.mark.parametrize("k", [2, 5, 12])
def test_kmeanspp_draws_equal_independent_restatement(oracle, seed, k):
    """On integer data (exact squared distances, as the reference's GEMM form also gives
    there) the oracle's k-means++ picks exactly the rows of an independent restatement of
    kmeans.cpp:1524-1664 over libstdc++'s engine and distributions: D^2 weights, 2 + ln k
    trials, the cheapest candidate, first centre uniform."""
    import _stdrng
    rng = np.random.default_rng(seed + 100 * k)
    X = rng.integers(-6, 7, size=(40, 3)).astype(np.float64)
    ref = _stdrng.rows(X, _stdrng.kmeanspp(X, k, seed))
    np.testing.assert_array_equal(oracle.kmeans_init(X, k, "k-means++", seed=seed), ref)


@pytest.mark.parametrize("seed", [0, 78])
@pytest.mark.parametrize("k,m", [(3, 1), (6, 5), (9, 50)])
def test_afkmc2_draws_equal_independent_restatement(oracle, seed, k, m):
    """The same check for AFK-MC2 (kmeans.cpp:1667-1811): the proposal, the chain of m steps
    and its acceptance rule, including the acceptance without a draw when d(x) is zero."""
    import _stdrng
    rng = np.random.default_rng(seed + 10 * k + m)
    X = rng.integers(-5, 6, size=(30, 2)).astype(np.float64)
    X[5] = X[6]  # a duplicate row: a chain at an existing centre has d(x) = 0
    ref = _stdrng.rows(X, _stdrng.afkmc2(X, k, seed, m))
    np.testing.assert_array_equal(oracle.kmeans_init(X, k, "afk-mc2", seed=seed,
                                                     afk_samples=m), ref)


def test_kmeanspp_picks_every_distinct_point(oracle):
    """With k equal to the number of distinct rows, k-means++ picks each of them once (a
    chosen row has weight zero), so the centres are the distinct rows."""
    rng = np.random.default_rng(11)
    P = rng.integers(-20, 20, size=(7, 4)).astype(np.float64)
    X = P[rng.integers(0, 7, size=60)]
    X[:7] = P
    C = oracle.kmeans_init(X, 7, "k-means++", seed=3)
    assert sorted(map(tuple, C)) == sorted(map(tuple, P))


def test_elkan_equals_lloyd_without_ties(oracle):
    """Elkan's bounds only skip distances that cannot change the label, so on data without
    equidistant centres it follows Lloyd's iterates exactly."""
    for seed, (n, d, k) in enumerate([(400, 3, 7), (300, 8, 20), (250, 2, 40)]):
        rng = np.random.default_rng(seed)
        X = rng.standard_normal((n, d)) + rng.integers(0, 4, size=(n, 1)) * 3.0
        C0 = X[rng.choice(n, k, replace=False)]
        a = oracle.kmeans(X, k, "supplied", C0, max_iter=60, tol=1e-6, algorithm="lloyd")
        b = oracle.kmeans(X, k, "supplied", C0, max_iter=60, tol=1e-6, algorithm="elkan")
        np.testing.assert_array_equal(a["labels"], b["labels"])
        np.testing.assert_array_equal(a["centres"], b["centres"])
        assert a["n_iter"] == b["n_iter"]


def _split(counts, flat):
    out, o = [], 0
    for c in counts:
        out.append(list(flat[o:o + c]))
        o += c
    return out


@pytest.mark.parametrize("single", [False, True])
def test_radius_known_answers(known, oracle, single):
    """The reference's radius-neighbour answers on the 6x3 example (radius 5, sorted): the
    euclidean search compares squared distances with 25, the squared one with 5."""
    c = known["knn_example"]
    Xt = np.array(c["X_train_colmajor"]).reshape((c["n_features"], c["n_samples"])).T
    Xq = np.array(c["X_test_colmajor"]).reshape((c["n_features"], c["n_queries"])).T
    for metric in ("euclidean", "sqeuclidean"):
        counts, ind, dist = oracle.radius(Xt, Xq, c["radius"], metric, single=single, sort=True)
        assert _split(counts, ind) == c[f"radius_{metric}_ind"]
        for got, exp in zip(_split(counts, dist), c[f"radius_{metric}_dist"]):
            np.testing.assert_allclose(got, exp, rtol=1e-6 if single else 1e-14)
    # Unsorted: the same sets in ascending index.
    counts, ind, _ = oracle.radius(Xt, Xq, c["radius"], "euclidean", single=single)
    assert _split(counts, ind) == [sorted(v) for v in c["radius_euclidean_ind"]]


def test_radius_vs_sklearn_and_brute_force(oracle):
    """Radius neighbours against scikit-learn's brute force (sorted) and against a direct
    numpy restatement of the inclusive threshold (unsorted, ascending index)."""
    from sklearn.neighbors import NearestNeighbors
    rng = np.random.default_rng(12)
    Xt = rng.standard_normal((400, 7))
    Xq = rng.standard_normal((50, 7))
    for metric, r in (("euclidean", 2.6), ("manhattan", 6.0), ("cosine", 0.4)):
        nn = NearestNeighbors(radius=r, algorithm="brute", metric=metric).fit(Xt)
        sd, si = nn.radius_neighbors(Xq, sort_results=True)
        counts, ind, dist = oracle.radius(Xt, Xq, r, metric, sort=True)
        got_i, got_d = _split(counts, ind), _split(counts, dist)
        for q in range(len(Xq)):
            np.testing.assert_array_equal(got_i[q], si[q])
            np.testing.assert_allclose(got_d[q], sd[q], rtol=1e-12, atol=1e-12)
    D = np.sqrt(((Xq[:, None, :] - Xt[None, :, :]) ** 2).sum(-1))
    counts, ind, dist = oracle.radius(Xt, Xq, 2.6, "sqeuclidean")
    for q, (iv, dv) in enumerate(zip(_split(counts, ind), _split(counts, dist))):
        exp = np.nonzero(D[q] ** 2 <= 2.6)[0]
        np.testing.assert_array_equal(iv, exp)
        np.testing.assert_allclose(dv, D[q, exp] ** 2, rtol=1e-12)


def _first_seen(labels):
    """Renames cluster labels in order of first appearance (noise stays -1), as the
    reference's DBSCAN tests compare them (dbscan_public.cpp:42-63)."""
    m, out = {}, []
    for v in labels:
        if v == -1:
            out.append(-1)
        else:
            out.append(m.setdefault(v, len(m)))
    return out


def _dbscan_case_data(c):
    A = np.array(c["A"], dtype=np.float64)
    n, d, lda = c["n_samples"], c["n_features"], c["lda"]
    if c["order"] == "row-major":
        return A.reshape(n, lda)[:, :d]
    return A.reshape(d, lda).T[:n, :]


@pytest.mark.parametrize("single", [False, True])
def test_dbscan_known_answers(known, oracle, single):
    """The reference's DBSCAN cases (dbscan_test_data.hpp): labels up to renaming, core
    samples, cluster counts; tree algorithms give the same neighbourhoods as brute force."""
    for c in known["dbscan"]:
        X = _dbscan_case_data(c)
        labels, core, nc = oracle.dbscan(X, c["eps"], c["min_samples"], c["metric"], c["power"],
                                         single=single)
        assert nc == c["n_clusters"], c["name"]
        assert _first_seen(labels) == _first_seen(c["labels"]), c["name"]
        assert list(np.nonzero(core)[0]) == c["core"], c["name"]


def test_dbscan_vs_sklearn(oracle):
    """Labels equal to scikit-learn's DBSCAN (brute force), including its numbering: clusters
    in the order of their first core point, a border point in the first cluster reaching it."""
    from sklearn.cluster import DBSCAN
    rng = np.random.default_rng(21)
    centres = rng.uniform(-6, 6, (6, 4))
    X = np.vstack([c + rng.normal(0, 0.6, (120, 4)) for c in centres] +
                  [rng.uniform(-9, 9, (60, 4))])
    X = X[rng.permutation(len(X))]
    for metric, eps, p in (("euclidean", 0.9, 2.0), ("manhattan", 1.6, 2.0),
                           ("cosine", 0.02, 2.0), ("minkowski", 0.8, 3.0)):
        for ms in (1, 5, 12):
            sk = DBSCAN(eps=eps, min_samples=ms, metric=metric, p=p, algorithm="brute").fit(X)
            labels, core, nc = oracle.dbscan(X, eps, ms, metric, p)
            np.testing.assert_array_equal(labels, sk.labels_, err_msg=f"{metric} {ms}")
            np.testing.assert_array_equal(np.nonzero(core)[0], sk.core_sample_indices_)
            assert nc == len(set(sk.labels_) - {-1})
    # "sqeuclidean" compares squared distances with eps; "l2" the distance; minkowski p = 2
    # is evaluated as the squared distance and compared with eps itself, as the reference does
    sk = DBSCAN(eps=0.9, min_samples=5, algorithm="brute").fit(X)
    np.testing.assert_array_equal(oracle.dbscan(X, 0.81, 5, "sqeuclidean")[0], sk.labels_)
    np.testing.assert_array_equal(oracle.dbscan(X, 0.9, 5, "l2")[0], sk.labels_)
    np.testing.assert_array_equal(oracle.dbscan(X, 0.81, 5, "minkowski", 2.0)[0], sk.labels_)


def _knn_predict_data(c):
    Xt = np.array(c["X_train_colmajor"]).reshape(3, 6).T
    Xq = np.array(c["X_test_colmajor"]).reshape(3, 3).T
    return Xt, Xq


@pytest.mark.parametrize("single", [False, True])
def test_nn_predict_known_answers(known, oracle, single):
    """The reference's k-NN and radius classification / regression answers on its 6 x 3
    example (k 3 and 5, uniform and 1/d weights, euclidean and squared euclidean, radius 21)."""
    c = known["knn_predict"]
    Xt, Xq = _knn_predict_data(c)
    tol = 1e-6 if single else 1e-12
    for metric in ("euclidean", "sqeuclidean"):
        for w in ("uniform", "distance"):
            for k in ("3", "5"):
                proba, labels, _ = oracle.nn_classify(Xt, c["y_class"], Xq, k=int(k), metric=metric,
                                                      weights=w, single=single)
                exp = np.array(c["proba_knn"][k][w][metric]).reshape(3, 3).T
                np.testing.assert_allclose(proba, exp, rtol=tol, atol=tol)
                assert list(labels) == c["labels_knn"][k][w]
                t, _ = oracle.nn_regress(Xt, c["y_reg"], Xq, k=int(k), metric=metric, weights=w,
                                         single=single)
                np.testing.assert_allclose(t, c["targets_knn"][k][w][metric], rtol=tol)
            proba, labels, first = oracle.nn_classify(Xt, c["y_class"], Xq, r=c["radius"],
                                                      metric=metric, weights=w, single=single)
            exp = np.array(c["proba_radius"][w][metric]).reshape(3, 3).T
            np.testing.assert_allclose(proba, exp, rtol=tol, atol=tol)
            assert list(labels) == c["labels_radius"][metric] and first == -1
            t, _ = oracle.nn_regress(Xt, c["y_reg"], Xq, r=c["radius"], metric=metric, weights=w,
                                     single=single)
            np.testing.assert_allclose(t, c["targets_radius"][w][metric], rtol=tol)


def test_nn_predict_vs_sklearn(oracle):
    """k-NN and radius classifiers / regressors against scikit-learn's brute-force estimators
    (no query at distance 0, where scikit-learn's 1/d weighting differs by design)."""
    from sklearn.neighbors import (KNeighborsClassifier, KNeighborsRegressor,
                                   RadiusNeighborsClassifier, RadiusNeighborsRegressor)
    rng = np.random.default_rng(31)
    Xt = rng.standard_normal((500, 5))
    Xq = rng.standard_normal((80, 5))
    y = rng.choice([-3, 7, 10, 42], 500)
    t = rng.standard_normal(500)
    for w in ("uniform", "distance"):
        for metric in ("euclidean", "manhattan"):
            for k in (1, 6, 25):
                sk = KNeighborsClassifier(k, weights=w, metric=metric, algorithm="brute").fit(Xt, y)
                proba, labels, _ = oracle.nn_classify(Xt, y, Xq, k=k, metric=metric, weights=w)
                np.testing.assert_allclose(proba, sk.predict_proba(Xq), rtol=1e-12, atol=1e-12)
                np.testing.assert_array_equal(labels, sk.predict(Xq))
                skr = KNeighborsRegressor(k, weights=w, metric=metric, algorithm="brute").fit(Xt, t)
                np.testing.assert_allclose(oracle.nn_regress(Xt, t, Xq, k=k, metric=metric,
                                                             weights=w)[0], skr.predict(Xq),
                                           rtol=1e-12, atol=1e-12)
            r = 2.2 if metric == "euclidean" else 4.0
            sk = RadiusNeighborsClassifier(r, weights=w, metric=metric, algorithm="brute",
                                           outlier_label=None).fit(Xt, y)
            proba, labels, first = oracle.nn_classify(Xt, y, Xq, r=r, metric=metric, weights=w)
            assert first == -1
            np.testing.assert_allclose(proba, sk.predict_proba(Xq), rtol=1e-12, atol=1e-12)
            np.testing.assert_array_equal(labels, sk.predict(Xq))
            skr = RadiusNeighborsRegressor(r, weights=w, metric=metric, algorithm="brute").fit(Xt, t)
            np.testing.assert_allclose(oracle.nn_regress(Xt, t, Xq, r=r, metric=metric,
                                                         weights=w)[0], skr.predict(Xq),
                                       rtol=1e-12, atol=1e-12)
    # outliers: the manual label / target, or the most frequent label / mean target
    Xfar = np.vstack([Xq[:3], np.full((2, 5), 50.0)])
    proba, labels, first = oracle.nn_classify(Xt, y, Xfar, r=2.2, outlier="manual",
                                              outlier_label=99)
    assert first == 3 and list(labels[3:]) == [99, 99] and not proba[3:].any()
    vals, cnt = np.unique(y, return_counts=True)
    proba, labels, _ = oracle.nn_classify(Xt, y, Xfar, r=2.2, outlier="most_frequent")
    assert list(labels[3:]) == [vals[np.argmax(cnt)]] * 2
    tv, first = oracle.nn_regress(Xt, t, Xfar, r=2.2, outlier="most_frequent")
    assert first == 3 and np.allclose(tv[3:], t.mean())


def test_ivf_oracle(known, oracle):
    """IVF-Flat: the reference's fixed cases; all lists probed = exact k-NN; one list probed =
    brute force restricted to the query's nearest list (a numpy statement)."""
    for c in known["ivf"]:
        Xt = np.array(c["X_train"]).reshape(
            (c["n_features"], c["n_samples"]) if c["order"] == "column-major"
            else (c["n_samples"], c["n_features"]))
        Xq = np.array(c["X_test"]).reshape(
            (c["n_features"], c["n_queries"]) if c["order"] == "column-major"
            else (c["n_queries"], c["n_features"]))
        if c["order"] == "column-major":
            Xt, Xq = Xt.T, Xq.T
        C = Xt[:c["n_list"]]  # k-means of these cases returns rows of the data
        ind, dist, _ = oracle.ivf_search(C, Xt, Xq, c["k"], c["n_probe"])
        got = dist.T.ravel() if c["order"] == "column-major" else dist.ravel()
        np.testing.assert_allclose(got, c["kdist"])
        if c["kind"] is not None:
            assert (ind.T.ravel() if c["order"] == "column-major" else ind.ravel()).tolist() == c["kind"]
    rng = np.random.default_rng(8)
    X = rng.standard_normal((600, 6))
    Q = rng.standard_normal((40, 6))
    C = X[rng.choice(600, 12, replace=False)]
    ind, dist, sizes = oracle.ivf_search(C, X, Q, 9, 12, metric="euclidean")
    ei, ed = oracle.knn(X, Q, 9)
    np.testing.assert_array_equal(ind, ei)
    np.testing.assert_allclose(dist, ed, rtol=1e-12)
    assert sizes.sum() == 600
    ind, dist, _ = oracle.ivf_search(C, X, Q, 5, 1)
    lab = np.argmin(((X[:, None] - C[None]) ** 2).sum(-1), 1)
    for q in range(len(Q)):
        cq = np.argmin(((C - Q[q]) ** 2).sum(-1))
        members = np.nonzero(lab == cq)[0]
        d2 = ((X[members] - Q[q]) ** 2).sum(-1)
        best = members[np.lexsort((members, d2))][:5]
        got = ind[q][ind[q] >= 0]
        np.testing.assert_array_equal(got, best)


def ann_case_arrays(c, dtype=np.float64):
    """Training and query matrices of an ann_tests.hpp case from its storage order and
    leading dimensions."""
    def mat(v, rows, cols, ld):
        v = np.asarray(v, dtype=dtype)
        if c["order"] == "column-major":
            return np.array([[v[i + j * ld] for j in range(cols)] for i in range(rows)],
                            dtype=dtype)
        return np.array([[v[i * ld + j] for j in range(cols)] for i in range(rows)],
                        dtype=dtype)
    return (mat(c["X_train"], c["n_samples"], c["n_features"], c["ldx_train"]),
            mat(c["X_test"], c["n_queries"], c["n_features"], c["ldx_test"]))


def _as_order(a, order):
    return a.T.ravel() if order == "column-major" else a.ravel()


@pytest.mark.parametrize("idx", range(52))
def test_ivf_reference_cases_end_to_end(known, oracle, idx):
    """Every fixed IVF-Flat case of ann_tests.hpp (squared euclidean, euclidean, inner product
    and cosine) through the oracle: the k-means training of train_ivfflat (random initial
    centres, one initialisation, tolerance sqrt(eps); spherical k-means without data
    normalisation for inner product and cosine, on normalised rows for cosine), then add and
    search, against the expected indices and distances (3000 eps, as the reference's test).
    Cases that subsample the training rows draw from std::mt19937 and are left to the GPU
    tests."""
    cases = known["ivf_cases"]
    if idx >= len(cases):
        pytest.skip("fewer cases")
    c = cases[idx]
    if c["train_fraction"] < 1.0:
        pytest.skip("training sample drawn with std::mt19937: checked on the GPU")
    Xt, Xq = ann_case_arrays(c)
    metric = c["metric"]
    seed = max(c["seed"], 0)
    tol = np.sqrt(np.finfo(np.float64).eps)
    if metric in ("inner product", "cosine"):
        Xk = oracle._unit_rows(Xt, np.float64) if metric == "cosine" else Xt
        r = oracle.kmeans_sph(Xk, c["n_list"], "random", seed=seed, max_iter=c["kmeans_iter"],
                              tol=tol, normalize=False)
    else:
        r = oracle.kmeans(Xt, c["n_list"], "random", seed=seed, max_iter=c["kmeans_iter"],
                          tol=tol)
    ind, dist, sizes = oracle.ivf_search(r["centres"], Xt, Xq, c["k"], c["n_probe"], metric)
    np.testing.assert_allclose(_as_order(dist, c["order"]), c["expected_kdist"], rtol=0,
                               atol=3e3 * np.finfo(np.float64).eps)
    if c["expected_kind"] is not None:
        assert _as_order(ind, c["order"]).tolist() == c["expected_kind"]
    assert sizes.sum() == c["n_samples"]


@pytest.mark.parametrize("metric", ["inner product", "cosine"])
def test_ivf_oracle_ip_cosine_all_lists_is_exact(oracle, metric):
    """With every list probed the IVF search is the exact search: the largest inner products
    (ties to the lower index) or the smallest cosine distances of numpy's brute force."""
    rng = np.random.default_rng(4)
    X = rng.standard_normal((500, 7)) * rng.uniform(0.5, 2, (500, 1))
    Q = rng.standard_normal((30, 7))
    C = X[rng.choice(500, 9, replace=False)]
    ind, dist, _ = oracle.ivf_search(C, X, Q, 8, 9, metric)
    if metric == "cosine":
        Xn = X / np.linalg.norm(X, axis=1, keepdims=True)
        Qn = Q / np.linalg.norm(Q, axis=1, keepdims=True)
        D = 1.0 - Qn @ Xn.T
    else:
        D = -(Q @ X.T)
    for q in range(len(Q)):
        best = np.lexsort((np.arange(500), D[q]))[:8]
        np.testing.assert_array_equal(ind[q], best)
        np.testing.assert_allclose(dist[q], D[q, best] if metric == "cosine" else -D[q, best],
                                   rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 8, 78])
@pytest.mark.parametrize("n,k", [(12, 4), (40, 7), (300, 64)])
def test_random_init_is_the_references_selection_sampling(oracle, seed, n, k):
    """The random initial centres are the rows da_std::sample selects (one engine draw per
    row visited), not those of std::sample (which draws two indices from one engine call):
    the two differ, so a mix-up fails here."""
    import _stdrng
    X = np.arange(n * 2, dtype=np.float64).reshape(n, 2)
    want = _stdrng.rows(X, _stdrng.selection_sample(n, k, seed))
    got = oracle.kmeans(X, k, "random", seed=seed, max_iter=0)["centres"]
    np.testing.assert_array_equal(got, want)
    gs = oracle.kmeans_sph(X + 1.0, k, "random", seed=seed, max_iter=0, normalize=False)
    np.testing.assert_allclose(gs["centres"],
                               want / np.linalg.norm(want + 1.0, axis=1, keepdims=True) +
                               1.0 / np.linalg.norm(want + 1.0, axis=1, keepdims=True),
                               rtol=1e-15)
