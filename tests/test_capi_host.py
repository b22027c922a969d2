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
    with pytest.raises(ValueError, match="mode='small'"):
        turbocat.TurboCatRegressor(mode="small", subsample=1.0).fit(X, y)
    with pytest.raises(ValueError, match="multiclass"):
        turbocat.TurboCatClassifier(subsample=1.0).fit(X, np.arange(50) % 3)
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
