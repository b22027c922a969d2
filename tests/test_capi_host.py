"""The drop-in boundary and host API without a GPU: the C-ABI library loads
and exports every entry point include/turbocat_hip.h declares; host-only
entry points and error paths behave like the reference's; the Python
classes keep the reference's names, keywords and defaults; TCAT v3 model
files round-trip byte for byte."""
import inspect
import os
import re
import struct

import numpy as np
import pytest

from conftest import ROOT, regression_data

HEADER = os.path.join(ROOT, "include", "turbocat_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from turbocat import capi
    L = capi.load()
    decl = declared_functions()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(capi.EXPORTS) == decl


def test_init_prediction_host_entry_matches_oracle(oracle):
    import ctypes
    from turbocat import capi
    L = capi.load()
    rng = np.random.default_rng(0)
    for loss, y in ((0, rng.standard_normal(1001).astype(np.float32)),
                    (1, (rng.random(999) < 0.3).astype(np.float32))):
        out = ctypes.c_float()
        capi.check(L.tc_init_prediction(loss, y.ctypes.data_as(ctypes.c_void_p), len(y),
                                        ctypes.byref(out)))
        assert np.float32(out.value) == np.float32(oracle.init_prediction(loss, y))
    rc = L.tc_init_prediction(7, None, 0, None)
    assert rc == -1 and b"bad arguments" in L.tc_last_error()


def test_context_without_gpu_is_a_clean_error():
    from turbocat import capi
    import ctypes
    L = capi.load()
    h = ctypes.c_void_p()
    rc = L.tc_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        L.tc_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == -3 and len(L.tc_last_error()) > 0
    with pytest.raises(capi.TurboCatError):
        capi.Context(0)


def test_python_api_surface_matches_reference_defaults():
    import turbocat
    c = inspect.signature(turbocat.TurboCatClassifier.__init__).parameters
    r = inspect.signature(turbocat.TurboCatRegressor.__init__).parameters
    # python/turbocat/__init__.py:340-363 and :628-646
    for k, v in dict(n_estimators=1000, learning_rate=0.1, max_depth=6, max_bins=255,
                     subsample=0.8, colsample_bytree=0.8, min_child_weight=1.0, lambda_l2=1.0,
                     loss="logloss", mode="auto", seed=42, n_jobs=-1).items():
        assert c[k].default == v, k
    for k, v in dict(n_estimators=1000, learning_rate=0.1, max_depth=6, loss="mse",
                     subsample=0.8, colsample_bytree=0.8, lambda_l2=0.0, min_child_weight=1.0,
                     seed=42).items():
        assert r[k].default == v, k
    for m in ("fit", "predict", "predict_proba", "predict_proba_nobinning_fast", "save", "load",
              "feature_importance", "get_params", "set_params"):
        assert hasattr(turbocat.TurboCatClassifier, m)
    for m in ("fit", "predict", "save", "load", "feature_importance"):
        assert hasattr(turbocat.TurboCatRegressor, m)
    # the pybind11 classes' own defaults (python/bindings.cpp:1048-1073, :1099-1118)
    from turbocat import _turbocat

    def defaults(cls):
        doc = cls.__init__.__doc__
        return {k: v.strip() for k, v in re.findall(r"(\w+): [^=,]+ = ('[^']*'|[^,)]+)", doc)}

    c = defaults(_turbocat.TurboCatClassifier)
    for k, v in dict(n_estimators=500, max_depth=6, max_bins=255, subsample=1.0,
                     colsample_bytree=1.0, lambda_l2=1.0, seed=42, verbosity=1).items():
        assert float(c[k]) == v, k
    assert float(np.float32(float(c["learning_rate"]))) == float(np.float32(0.2))
    for k, v in dict(loss="logloss", mode="auto", early_stopping_metric="loss",
                     grow_policy="depthwise", cat_encoding="ordered").items():
        assert c[k] == repr(v), k
    r = defaults(_turbocat.TurboCatRegressor)
    assert r["loss"] == "'mse'" and float(r["lambda_l2"]) == 3.0 and int(r["max_leaves"]) == 64


def test_not_fitted_and_unsupported_options():
    import turbocat
    X, y = regression_data(50, 4, 0)
    clf = turbocat.TurboCatClassifier(n_estimators=5, subsample=1.0)
    with pytest.raises(RuntimeError, match="not fitted"):
        clf.predict_proba(X)
    with pytest.raises(RuntimeError, match="not fitted"):
        turbocat.TurboCatRegressor().predict(X)
    with pytest.raises(ValueError, match="use_ordered_boosting"):
        turbocat.TurboCatClassifier(use_ordered_boosting=True, subsample=1.0).fit(X, y > 0)
    with pytest.raises(ValueError, match="use_gradtree"):
        turbocat.TurboCatClassifier(use_gradtree=True, subsample=1.0).fit(X, y > 0)
    with pytest.raises(ValueError, match="loss='huber'"):
        turbocat.TurboCatRegressor(loss="huber").fit(X, y)
    with pytest.raises(ValueError, match="goss_top_rate"):
        turbocat.TurboCatRegressor(use_goss=True, goss_top_rate=1.5).fit(X, y)
    assert clf.n_trees == 0


def write_tcat_v3(path, model, edges, n_edges, task=2, lr=0.1, max_depth=6, lam=0.0):
    """TCAT v3 writer from the documented layout (booster.cpp:1049-1109,
    symmetric_tree.cpp:1199-1215): little-endian, no padding, size_t = 8 B."""
    T, D = model["T"], model["D"]
    with open(path, "wb") as f:
        f.write(b"TCAT")
        f.write(struct.pack("<IfBIfHfB", 3, model["base"], task, 2, lr, max_depth, lam, 1))
        f.write(struct.pack("<Q", len(n_edges)))
        for k, ne in enumerate(n_edges):
            f.write(struct.pack("<Q", int(ne)))
            f.write(np.asarray(edges[k, :ne], "<f4").tobytes())
        f.write(struct.pack("<Q", T))
        f.write(np.full(T, lr, "<f4").tobytes())
        for t in range(T):
            d = int(model["depth"][t])
            f.write(struct.pack("<H", d))
            for l in range(d):
                f.write(struct.pack("<HBff", int(model["feature"][t * D + l]),
                                    int(model["bin"][t * D + l]), model["fthr"][t * D + l],
                                    model["gain"][t * D + l]))
            f.write(struct.pack("<Q", 1 << d))
            f.write(np.asarray(model["leaves"][t << D:(t << D) + (1 << d)], "<f4").tobytes())


def test_tcat_v3_round_trip(tmp_path):
    import turbocat
    z = np.load(os.path.join(ROOT, "tests", "golden", "reg_small.npz"), allow_pickle=False)
    n, F, seed, loss, T, D = (int(v) for v in z["meta"])
    model = dict(T=T, D=D, feature=z["m_feature"], bin=z["m_bin"], fthr=z["m_fthr"],
                 gain=z["m_gain"], depth=z["m_depth"], leaves=z["m_leaves"], base=float(z["m_base"]))
    src = tmp_path / "ref.tcat"
    write_tcat_v3(src, model, z["edges"], z["n_edges"], lr=float(z["lr"]), max_depth=D)
    reg = turbocat.TurboCatRegressor().load(str(src))
    assert reg.n_trees == T
    assert np.float32(reg.base_prediction) == z["m_base"]
    out = tmp_path / "ours.tcat"
    reg.save(str(out))
    assert open(src, "rb").read() == open(out, "rb").read()
    imp = reg.feature_importance()
    exp = np.zeros(256, np.float32)
    for t in range(T):
        for l in range(int(model["depth"][t])):
            exp[model["feature"][t * D + l]] += model["gain"][t * D + l]
    assert np.allclose(imp["gain"], exp, rtol=1e-6)
    with pytest.raises(RuntimeError, match="Invalid TurboCat model file"):
        bad = tmp_path / "bad.tcat"
        bad.write_bytes(b"NOPE" + b"\0" * 16)
        turbocat.TurboCatRegressor().load(str(bad))


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Sizes and field offsets of the parameter / model structs as a C
    compiler lays them out from include/turbocat_hip.h equal the ctypes
    mirrors the Python layer passes across the boundary."""
    import ctypes
    import shutil
    import subprocess
    from turbocat import capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"tc_tree_params": capi.TreeParams, "tc_train_params": capi.TrainParams,
               "tc_model": capi.Model, "tc_serve_params": capi.ServeParams,
               "tc_rtree_params": capi.RTreeParams, "tc_rtrain_params": capi.RTrainParams,
               "tc_rmodel": capi.RModel}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "turbocat_hip.h"',
             'int main(void) {']
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        parts = line.split()
        got[tuple(parts[:-1])] = int(parts[-1])
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def write_tcat_v3_regular(path, trees, edges, n_edges, task, n_classes, lr, max_depth, lam, base):
    """TCAT v3 writer for regular trees from the documented layouts
    (booster.cpp:1049-1109 with use_symmetric 0; Tree::save, tree.cpp:
    1748-1774: u16 depth, u32 n_classes, u32 n_nodes, u16 n_leaves, u16
    max_depth, f32 lambda_l2, then per node u16 feature, u8 bin, u8 flags,
    u16 left, u16 right (TreeIndex, types.hpp:47), f32 value, f32 gain)."""
    with open(path, "wb") as f:
        f.write(b"TCAT")
        f.write(struct.pack("<IfBIfHfB", 3, base, task, n_classes, lr, max_depth, lam, 0))
        f.write(struct.pack("<Q", len(n_edges)))
        for k, ne in enumerate(n_edges):
            f.write(struct.pack("<Q", int(ne)))
            f.write(np.asarray(edges[k, :ne], "<f4").tobytes())
        f.write(struct.pack("<Q", len(trees)))
        f.write(np.full(len(trees), lr, "<f4").tobytes())
        for t in trees:
            n = len(t["feature"])
            f.write(struct.pack("<HIIHHf", t["depth"], 1, n, t["n_leaves"], max_depth, lam))
            for k in range(n):
                f.write(struct.pack("<HBBHHff", t["feature"][k], t["bin"][k], t["flags"][k],
                                    t["left"][k], t["right"][k], t["value"][k], t["gain"][k]))


def _regular_trees(seed, n_trees, F):
    """Random well-formed regular trees in the reference's numbering: a split
    node's children get the next two ids, and the subtree visited first (left
    or right at random) takes the ids after them (tree.cpp:357-390)."""
    rng = np.random.default_rng(seed)
    trees = []
    for _ in range(n_trees):
        recs, nxt, depth = {}, [1], [0]

        def build(i, d):
            depth[0] = max(depth[0], d)
            if d < 4 and rng.random() < 0.6:
                lc, rc = nxt[0], nxt[0] + 1
                nxt[0] += 2
                right_first = bool(rng.random() < 0.5)
                recs[i] = ("split", int(rng.integers(0, F)), int(rng.integers(0, 30)),
                           float(abs(rng.standard_normal())), lc, rc,
                           2 if rng.random() < 0.7 else 0)
                for c in ((rc, lc) if right_first else (lc, rc)):
                    build(c, d + 1)
            else:
                recs[i] = ("leaf", float(rng.standard_normal()))
        build(0, 0)
        t = dict(feature=[], bin=[], flags=[], left=[], right=[], value=[], gain=[])
        for k in range(nxt[0]):
            r = recs[k]
            split = r[0] == "split"
            for key, v in zip(("feature", "bin", "flags", "left", "right", "value", "gain"),
                              (r[1], r[2], r[6], r[4], r[5], 0.0, r[3]) if split else
                              (0, 0, 3, 0, 0, r[1], 0.0)):
                t[key].append(v)
        t["value"] = np.float32(t["value"])
        t["gain"] = np.float32(t["gain"])
        t["depth"] = depth[0]
        t["n_leaves"] = sum(1 for x in t["flags"] if x & 1)
        trees.append(t)
    return trees


@pytest.mark.parametrize("task,n_classes", [(2, 2), (0, 2), (1, 3)])
def test_tcat_v3_regular_round_trip(tmp_path, task, n_classes):
    """A regular-tree TCAT v3 file (regression, binary, and a 3-class model
    whose trees are stored iteration-major, class-minor) loads and saves back
    byte for byte; the loaded model reports its classes and tree count."""
    import turbocat
    F = 5
    trees = _regular_trees(11 + task, 9, F)
    edges = np.sort(np.random.default_rng(3).standard_normal((F, 255)).astype(np.float32), axis=1)
    n_edges = np.array([30, 31, 16, 17, 0], np.uint32)
    src = tmp_path / "ref.tcat"
    write_tcat_v3_regular(src, trees, edges, n_edges, task, n_classes, 0.1, 6, 1.0,
                          0.0 if task == 1 else 0.25)
    model = turbocat.TurboCatRegressor() if task == 2 else turbocat.TurboCatClassifier()
    model.load(str(src))
    assert model.n_trees == len(trees)
    if task != 2:
        assert model.n_classes_ == n_classes
    out = tmp_path / "ours.tcat"
    model.save(str(out))
    assert open(src, "rb").read() == open(out, "rb").read()
    imp = model.feature_importance()["gain"]
    exp = np.zeros(256, np.float64)
    for t in trees:
        for k in range(len(t["feature"])):
            if not t["flags"][k] & 1 and t["gain"][k] > 0:
                exp[t["feature"][k]] += t["gain"][k]
    assert np.allclose(imp, exp / exp.sum(), rtol=1e-5)
