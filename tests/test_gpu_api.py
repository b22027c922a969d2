"""The reference's end-to-end API tests (tests/test_turbocat.py:18-417), run
against this package on the GPU: same constructor arguments, same
assertions on shapes, dtypes, probability validity, toy accuracy, tree
counts, sampling options, feature importance, reproducibility and errors;
plus TCAT save/load round trips through the GPU predictor."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def simple_data():
    X = np.array([[-2], [-1], [1], [2], [-1.5], [-0.5], [0.5], [1.5]], dtype=np.float32)
    y = np.array([0, 0, 1, 1, 0, 0, 1, 1], dtype=np.float32)
    return X, y


@pytest.fixture
def random_data():
    rs = np.random.RandomState(42)
    X = rs.randn(100, 10).astype(np.float32)
    y = (X[:, 0] + X[:, 1] > 0).astype(np.float32)
    return X, y


def toy_clf():
    from turbocat import TurboCatClassifier
    return TurboCatClassifier(n_estimators=1, max_depth=3, learning_rate=1.0, use_goss=False,
                              subsample=1.0, min_child_weight=0.01, verbosity=0)


def test_fit_predict_toy(simple_data):
    X, y = simple_data
    clf = toy_clf().fit(X, y)
    assert clf.n_trees == 1
    proba = clf.predict_proba(X)
    assert proba.shape == (8, 2) and proba.dtype == np.float32
    assert ((proba >= 0) & (proba <= 1)).all() and np.allclose(proba.sum(1), 1.0)
    pred = clf.predict(X)
    assert pred.shape == (8,) and pred.dtype == np.int32
    assert (pred == y).mean() == 1.0
    # the known answer: p = sigmoid(+-1) for the two leaves
    assert np.allclose(proba[:, 1], np.where(X[:, 0] >= 0.5, 0.7310586, 0.26894143), atol=1e-6)


def test_multiple_trees_and_sampling(random_data):
    from turbocat import TurboCatClassifier
    X, y = random_data
    for kw in (dict(n_estimators=10), dict(n_estimators=10, subsample=0.7),
               dict(n_estimators=10, colsample_bytree=0.5),
               dict(n_estimators=10, use_goss=True, goss_top_rate=0.3, goss_other_rate=0.3)):
        clf = TurboCatClassifier(learning_rate=0.3, verbosity=0, **kw).fit(X, y)
        assert clf.n_trees == 10
        assert clf.predict_proba(X).shape == (100, 2)
        assert (clf.predict(X) == y).mean() > 0.8


def test_feature_importance_and_params(random_data):
    from turbocat import TurboCatClassifier
    X, y = random_data
    clf = TurboCatClassifier(n_estimators=10, learning_rate=0.3, verbosity=0).fit(X, y)
    imp = clf.feature_importance()
    assert imp["gain"].shape == (256,) and imp["gain"][:2].sum() > 0
    assert abs(float(imp["gain_normalized"].sum()) - 1.0) < 1e-5
    p = clf.get_params()
    assert p["n_estimators"] == 10 and p["learning_rate"] == 0.3


def test_same_seed_same_result():
    from turbocat import TurboCatClassifier
    rs = np.random.RandomState(42)
    X = rs.randn(50, 5).astype(np.float32)
    y = (X[:, 0] > 0).astype(np.float32)
    a = TurboCatClassifier(n_estimators=5, learning_rate=0.3, seed=123, verbosity=0).fit(X, y)
    b = TurboCatClassifier(n_estimators=5, learning_rate=0.3, seed=123, verbosity=0).fit(X, y)
    assert np.allclose(a.predict_proba(X), b.predict_proba(X))


def test_single_sample_and_many_features():
    from turbocat import TurboCatClassifier
    rs = np.random.RandomState(0)
    Xt = rs.randn(50, 5).astype(np.float32)
    clf = TurboCatClassifier(n_estimators=5, verbosity=0).fit(Xt, (Xt[:, 0] > 0).astype(np.float32))
    assert clf.predict_proba(rs.randn(1, 5).astype(np.float32)).shape == (1, 2)
    X = rs.randn(100, 100).astype(np.float32)
    clf = TurboCatClassifier(n_estimators=5, verbosity=0).fit(X, (X[:, 0] + X[:, 1] > 0).astype(np.float32))
    assert clf.predict_proba(X).shape == (100, 2)


def test_regressor_and_save_load(tmp_path):
    from turbocat import TurboCatRegressor, TurboCatClassifier
    rs = np.random.RandomState(1)
    X = rs.randn(3000, 8).astype(np.float32)
    y = (X[:, 0] + 2 * np.sin(X[:, 1])).astype(np.float32)
    reg = TurboCatRegressor(n_estimators=40, verbosity=0).fit(X, y)
    p = reg.predict(X)
    assert p.dtype == np.float32 and np.corrcoef(p, y)[0, 1] > 0.9
    path = str(tmp_path / "m.tcat")
    reg.save(path)
    reg2 = TurboCatRegressor().load(path)
    assert np.array_equal(reg2.predict(X), p)
    yb = (y > 0).astype(np.float32)
    clf = TurboCatClassifier(n_estimators=20, verbosity=0).fit(X, yb)
    clf.save(path)
    clf2 = TurboCatClassifier().load(path)
    assert np.array_equal(clf2.predict_proba(X), clf.predict_proba(X))


@pytest.mark.parametrize("name", ["reg_small", "bin_small", "reg_deep", "pred_tail"])
def test_reference_model_file_predicts_like_the_reference(tmp_path, name):
    """(f)1 TCAT v3 interoperability on the device path: each golden model the
    reference trained (tests/golden/*.npz), written in the reference's file
    layout (booster.cpp:1049-1109), loaded through TurboCatRegressor.load and
    predicted on the GPU equals the reference's own predictions bit for bit
    (booster.cpp:899-925), tail rows included; saving it again reproduces the
    file byte for byte."""
    import os
    import sys
    import turbocat
    sys.path.insert(0, os.path.dirname(__file__))
    from test_capi_host import write_tcat_v3
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", f"{name}.npz"), allow_pickle=False)
    if name == "pred_tail":
        n, F, seed, T, D = (int(v) for v in z["meta"])
        gain = np.zeros(T * D, np.float32)
        cases = [(z[f"X{r}"], z[f"pred{r}"]) for r in (517, 13, 7)]
        edges = np.zeros((F, 255), np.float32)
        ne = np.zeros(F, np.uint32)
        task, lam = 2, 0.0
    else:
        n, F, seed, loss, T, D = (int(v) for v in z["meta"])
        gain = z["m_gain"]
        cases = [(z["Xp"], z["pred"]), (z["Xp"][:509], None)]
        edges, ne = z["edges"], z["n_edges"]
        task, lam = (2 if loss == 0 else 0), float(z["lam"])
    model = dict(T=T, D=D, feature=z["m_feature"], bin=z["m_bin"], fthr=z["m_fthr"], gain=gain,
                 depth=z["m_depth"], leaves=z["m_leaves"], base=float(z["m_base"]))
    path = tmp_path / f"{name}.tcat"
    write_tcat_v3(path, model, edges, ne, task=task, lr=float(z["lr"]), max_depth=D, lam=lam)
    reg = turbocat.TurboCatRegressor().load(str(path))
    assert reg.n_trees == T
    for X, ref in cases:
        out = reg.predict(X)
        if ref is not None:
            assert np.array_equal(out, ref)
    if name != "pred_tail":       # 509 rows: the last 5 take the tree-by-tree order
        from oracle import Oracle
        m = dict(model, weight=np.full(T, z["lr"], np.float32))
        assert np.array_equal(reg.predict(z["Xp"][:509]), Oracle().predict_f32(m, z["Xp"][:509]))
    again = tmp_path / "again.tcat"
    reg.save(str(again))
    assert open(path, "rb").read() == open(again, "rb").read()
