"""CPU checks: golden fixtures against the oracle, and the C-ABI library's
exports against include/adtomo_hip.h (no device calls)."""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


@pytest.mark.parametrize("name", ["gradtest2d", "config1_2d"])
def test_golden_2d_oracle(name):
    d = _load(name)
    u, it, err, conv = oracle.eikonal2d(d["f"], int(d["srcx"]), int(d["srcy"]), float(d["h"]))
    assert it == int(d["iters"]) and conv == bool(d["converged"])
    np.testing.assert_array_equal(u, d["u"])
    gf, lam, ns = oracle.eikonal2d_grad(d["grad_u"], u, d["f"], int(d["srcx"]), int(d["srcy"]),
                                        float(d["h"]))
    np.testing.assert_array_equal(gf, d["grad_f"])


@pytest.mark.parametrize("name", ["gradtest3d", "small3d"])
def test_golden_3d_oracle(name):
    d = _load(name)
    u, it, err = oracle.eikonal3d(d["u0"], d["f"], float(d["h"]), float(d["tol"]))
    assert it == int(d["iters"]) and err == float(d["err"])
    np.testing.assert_array_equal(u, d["u"])
    gu0, gf, lam = oracle.eikonal3d_grad(d["grad_u"], u, d["u0"], d["f"], float(d["h"]))
    np.testing.assert_array_equal(gu0, d["grad_u0"])
    np.testing.assert_array_equal(gf, d["grad_f"])


def test_golden_grads_vs_sparse_lu():
    d = _load("config1_2d")
    gf_lu, _ = oracle.lu_grad_2d(d["grad_u"], d["u"], d["f"], int(d["srcx"]), int(d["srcy"]),
                                 float(d["h"]))
    assert np.abs(gf_lu - d["grad_f"]).max() <= 1e-10 * np.abs(d["grad_f"]).max()
    d = _load("gradtest3d")
    gu0, gf_lu, _ = oracle.lu_grad_3d(d["grad_u"], d["u"], d["u0"], d["f"], float(d["h"]))
    assert np.abs(gf_lu - d["grad_f"]).max() <= 1e-10 * np.abs(d["grad_f"]).max()


def test_abi_library_exports_header_symbols():
    from adtomo_hip import _lib as L
    lib = ctypes.CDLL(L.LIB_PATH)
    syms = L.declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_argument_errors_without_device():
    # Argument validation happens before any device call.
    from adtomo_hip import _lib as L
    lib = L.lib()
    info = L.AdtInfo()
    buf = np.zeros(8)
    rc = lib.adt_eikonal3d_fwd_f64(buf.ctypes.data, buf.ctypes.data, 1.0, 1, 2, 4, 1e-3, 0,
                                   buf.ctypes.data, ctypes.byref(info), 0, None)
    assert rc == -1 and b"dimension" in lib.adt_last_error()
    rc = lib.adt_eikonal2d_fwd_f64(buf.ctypes.data, 3, 1, 9, 1, 1.0, buf.ctypes.data,
                                   ctypes.byref(info), 0, None)
    assert rc == -1 and b"outside" in lib.adt_last_error()


def test_strided_argument_errors_without_device():
    # Stride validation happens before any device call: non-positive strides,
    # strides that map two elements to one address, and NULL strides.
    from adtomo_hip import _lib as L
    lib = L.lib()
    info = L.AdtInfo()
    buf = np.zeros(64)
    p = buf.ctypes.data
    S3, S2 = ctypes.c_int64 * 3, ctypes.c_int64 * 2
    for st, msg in (((0, 4, 1), b"positive"), ((-8, 4, 1), b"positive"),
                    ((4, 4, 1), b"one address"), ((8, 4, 2), b"one address")):
        rc = lib.adt_eikonal3d_fwd_strided_f64(p, p, 1.0, 2, 2, 4, 1e-3, 0, S3(*st), p,
                                               ctypes.byref(info), 0, None)
        assert rc == -1 and msg in lib.adt_last_error(), (st, lib.adt_last_error())
    # the unit axis of a 2D call and an axis of extent 1 carry no stride
    rc = lib.adt_eikonal3d_fwd_strided_f64(p, p, 1.0, 2, 2, 4, 1e-3, 0, None, p,
                                           ctypes.byref(info), 0, None)
    assert rc == -1 and b"NULL" in lib.adt_last_error()
    rc = lib.adt_eikonal3d_bwd_strided_f64(p, p, p, p, 1.0, 2, 2, 4, S3(1, 1, 4), p, p,
                                           ctypes.byref(info), 0, None)
    assert rc == -1 and b"one address" in lib.adt_last_error()
    rc = lib.adt_eikonal2d_fwd_strided_f64(p, 3, 3, 1, 1, 1.0, S2(1, 2), p, ctypes.byref(info),
                                           0, None)
    assert rc == -1 and b"one address" in lib.adt_last_error()
    rc = lib.adt_eikonal2d_bwd_strided_f32(p, p, p, 3, 3, 9, 1, 1.0, S2(1, 4), p,
                                           ctypes.byref(info), 0, None)
    assert rc == -1 and b"outside" in lib.adt_last_error()


def test_strided_layout_checks_in_python():
    from adtomo_hip import ops
    a = np.zeros((4, 5, 6))
    assert ops._layout([a, a.copy()], 3) == ((30, 6, 1), False)
    fa = np.asfortranarray(a)
    assert ops._layout([fa], 3) == ((1, 4, 20), False)
    with pytest.raises(ValueError, match="share their strides"):
        ops._layout([a, fa], 3)
    with pytest.raises(ValueError, match="positive"):
        ops._layout([a[::-1]], 3)
    with pytest.raises(ValueError, match="3-D"):
        ops._layout([a[0]], 3)
    b = ops._empty_like_strided(fa, (1, 4, 20))
    assert b.shape == fa.shape and b.flags.f_contiguous
    big = np.zeros((7, 8, 9))
    v = big[:4, :5, :6]
    st, _ = ops._layout([v], 3)
    assert st == (72, 9, 1) and ops._empty_like_strided(v, st).strides == v.strides


def test_product_package_does_not_import_oracle():
    import glob
    pkg = os.path.join(os.path.dirname(GOLD), "..", "adtomo.jl_amd")
    for p in glob.glob(os.path.join(pkg, "**", "*.py"), recursive=True):
        txt = open(p).read()
        assert "from oracle" not in txt and "import oracle" not in txt, p


def test_tune_get_set_validation_without_device():
    # the launch setting is host state: validated and stored without a device
    import adtomo_hip as A
    base = A.get_tune("float64")
    try:
        assert base["diag_max_l"] == 1 << 30 and base["adj_wg"] in (256, 512, 1024)
        assert base["fuse_max_n"] == 1 << 21
        t = A.set_tune("float64", diag_max_l=0, adj_wg=256, order=0, fuse_max_n=0)
        assert A.get_tune("float64")["adj_wg"] == 256 and t["diag_max_l"] == 0
        assert t["fuse_max_n"] == 0
        assert A.get_tune("float32")["adj_wg"] == base["adj_wg"]  # per precision
        assert A.get_tune("float32")["fuse_max_n"] == base["fuse_max_n"]
        for bad in ({"diag_max_l": (1 << 30) + 1}, {"diag_max_l": -1}, {"adj_wg": 300},
                    {"order": 2}, {"fuse_max_n": -1}, {"adj_algo": 2}, {"fwd_segs": -1},
                    {"fwd_segs": 9}):
            with pytest.raises(A.AdtError):
                A.set_tune("float64", **bad)
        with pytest.raises(TypeError):
            A.set_tune("float64", tile=4)
    finally:
        A.set_tune("float64", **{k: base[k] for k in ("diag_max_l", "adj_wg", "order",
                                                      "fuse_max_n", "adj_algo", "fwd_segs")})


def test_tune_classes_and_json_round_trip_without_device(tmp_path):
    # class settings and the JSON file are host state too
    import adtomo_hip as A
    saved = str(tmp_path / "saved.json")
    A.save_tune(saved)
    try:
        A.clear_tune_classes("float64")
        A.clear_tune_classes("float32")
        assert A.tune_classes("float64") == []
        A.set_tune_class((128, 128, 64), "float64", fwd_segs=4, adj_algo=1)
        A.set_tune_class((48, 100, 24), "float64", diag_max_l=0)
        A.set_tune_class((128, 130, 64), "float64", fwd_segs=2)  # same class as the first
        A.set_tune_class((40, 40, 40), "float32", fwd_segs=5)
        c64 = A.tune_classes("float64")
        assert [(c["l_blocks"], c["bands"], c["log2n"]) for c in c64] == [(3, 1, 16), (8, 2, 20)]
        assert c64[1]["fwd_segs"] == 2 and c64[0]["diag_max_l"] == 0
        assert [(c["l_blocks"], c["fwd_segs"]) for c in A.tune_classes("float32")] == [(5, 5)]
        with pytest.raises(A.AdtError):
            A.set_tune_class((16, 16, 16), "float64", fwd_segs=9)
        with pytest.raises(A.AdtError):
            A.set_tune_class((1, 16, 16), "float64", fwd_segs=2)
        f = str(tmp_path / "t.json")
        A.save_tune(f)
        json.load(open(f))  # valid JSON
        A.clear_tune_classes("float64")
        A.set_tune("float64", order=1 - A.get_tune("float64")["order"])
        A.load_tune(f)
        assert A.tune_classes("float64") == c64
        assert A.get_tune("float64")["order"] == json.load(open(f))["global"][0]["order"]
        A.set_tune_class((128, 128, 64), "float64", clear=True)
        assert len(A.tune_classes("float64")) == 1
        bad = str(tmp_path / "bad.json")
        open(bad, "w").write('{"global": [{"dtype": 0, "adj_wg": 256}]}')
        with pytest.raises(A.AdtError):
            A.load_tune(bad)
        with pytest.raises(A.AdtError):
            A.load_tune(str(tmp_path / "missing.json"))
    finally:
        A.load_tune(saved)


def _c_struct_fields(name):
    import re
    hdr = open(os.path.join(os.path.dirname(GOLD), "..", "include", "adtomo_hip.h")).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), hdr, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1] if not decl.startswith("const ") else decl.split(None, 2)[2]
        for nm in names.split(","):
            nm = nm.strip().lstrip("*").split("[")[0].strip()
            fields.append(nm)
    return fields


def _julia_struct_fields(name):
    import re
    jl = open(os.path.join(os.path.dirname(GOLD), "..", "adtomo.jl_amd", "julia",
                           "ADTomoHIP.jl")).read()
    body = re.search(r"\nstruct %s\n(.*?)\nend" % name, jl, re.S).group(1)
    return [ln.split("::")[0].strip() for ln in body.splitlines() if "::" in ln]


@pytest.mark.parametrize("c_name,py_name,jl_name", [
    ("adt_info", "AdtInfo", "AdtInfo"), ("adt_event3d", "AdtEvent3d", "AdtEvent3d"),
    ("adt_source3d", "AdtSource3d", "AdtSource3d"), ("adt_misfit3d", "AdtMisfit3d", "AdtMisfit3d"),
    ("adt_tune", "AdtTune", "AdtTune"), ("adt_tune_class", "AdtTuneClass", "AdtTuneClass")])
def test_struct_layouts_agree_across_bindings(c_name, py_name, jl_name):
    # the C header, the ctypes twin and the Julia module declare the same
    # fields in the same order
    from adtomo_hip import _lib as L
    c = _c_struct_fields(c_name)
    py = [f[0] for f in getattr(L, py_name)._fields_]
    assert py == c, (py, c)
    if jl_name:
        assert _julia_struct_fields(jl_name) == c


def test_objective_and_tuning_argument_errors_without_device():
    # every new entry point validates its arguments before touching a device
    from adtomo_hip import _lib as L
    lib = L.lib()
    buf = np.zeros(64)
    p = buf.ctypes.data
    st = (L.AdtSource3d * 1)(L.AdtSource3d(0.5, 2.0, 2.0, 1.0, p))  # x below 1
    assert lib.adt_source_u0_3d_batched_f64(st, 1, p, 4, 4, 4, 1.0, None) == -1
    assert b"outside" in lib.adt_last_error()
    mis = (L.AdtMisfit3d * 1)(L.AdtMisfit3d(4, 4, 4, p, p, p, None, None, p, 0.0))
    coords = np.array([1.0, 2.0, 4.5])  # z beyond l = 4
    ptr = coords.ctypes.data
    assert lib.adt_misfit3d_batched_f64(mis, 1, ptr, ptr, ptr + 16, 1, None) == -1
    assert b"outside" in lib.adt_last_error()
    assert lib.adt_misfit3d_batched_f64(mis, -1, ptr, ptr, ptr, 1, None) == -1
    assert lib.adt_smooth3d_l1_f64(p, 4, 4, 4, 2, 3, 1.0, p, p, None) == -1
    assert lib.adt_smooth3d_l1_f64(p, 4, 4, 4, 5, 3, 1.0, p, p, None) == -1
    assert lib.adt_sum_arrays_f64(None, 0, 8, p, None) == -1
    assert lib.adt_slowness3d_f64(p, -1, 1.7, p, None) == -1
    assert lib.adt_slowness3d_grad_f64(None, 8, 1.7, None, p, p, None) == -1
    assert lib.adt_model3d_f64(None, p, 8, p, p, None) == -1
    assert lib.adt_model3d_grad_f64(p, p, None, None, 8, None, None) == -1
    ev = (L.AdtEvent3d * 1)()
    best = L.AdtTune()
    assert lib.adt_autotune(ev, 0, L.ADT_FORWARD, 0, 1, ctypes.byref(best), None) == -1
    assert lib.adt_autotune(ev, 1, 0, 0, 1, ctypes.byref(best), None) == -1
    assert lib.adt_autotune(ev, 1, L.ADT_FORWARD, 2, 1, ctypes.byref(best), None) == -1
    g = ctypes.c_void_p()
    assert lib.adt_graph_create(ev, 1, 0, 0, ctypes.byref(g), None) == -1
    # the one-launch inversion step: both flags, and items that name their
    # event's box, u and grad_u
    evs = (L.AdtEvent3d * 1)(L.AdtEvent3d(4, 4, 4, 1.0, 1e-3, p, p, p, p, None, p))
    good = L.AdtMisfit3d(4, 4, 4, p, p, p, None, p, p, 0.0)
    other_u = L.AdtMisfit3d(4, 4, 4, p + 8, p, p, None, p, p, 0.0)
    xyz = np.array([2.0, 2.0, 2.0]).ctypes.data
    both = L.ADT_FORWARD | L.ADT_ADJOINT
    assert lib.adt_eikonal3d_misfit_batched_f64(evs, (L.AdtMisfit3d * 1)(good), 1, xyz, xyz, xyz,
                                                1, L.ADT_FORWARD, None, None) == -1
    assert lib.adt_eikonal3d_misfit_batched_f64(evs, (L.AdtMisfit3d * 1)(other_u), 1, xyz, xyz,
                                                xyz, 1, both, None, None) == -1
    assert b"u and grad_u" in lib.adt_last_error()
    far = np.array([9.0]).ctypes.data
    assert lib.adt_graph_create_misfit(evs, (L.AdtMisfit3d * 1)(good), 1, xyz, xyz, far, 1, both,
                                       0, ctypes.byref(g), None) == -1
    assert b"outside" in lib.adt_last_error()
    # two events of one inversion step sharing a grad_u (or a u) buffer
    p2 = p + 8 * 32
    ev2 = (L.AdtEvent3d * 2)(L.AdtEvent3d(4, 4, 4, 1.0, 1e-3, p, p, p, p, None, p),
                             L.AdtEvent3d(4, 4, 4, 1.0, 1e-3, p, p, p, p2, None, p))
    it2 = (L.AdtMisfit3d * 2)(L.AdtMisfit3d(4, 4, 4, p, p, p, None, p, p, 0.0),
                              L.AdtMisfit3d(4, 4, 4, p2, p, p, None, p, p, 0.0))
    assert lib.adt_eikonal3d_misfit_batched_f64(ev2, it2, 2, xyz, xyz, xyz, 1, both, None,
                                                None) == -1
    assert b"share one grad_u" in lib.adt_last_error()
    ev3 = (L.AdtEvent3d * 2)(L.AdtEvent3d(4, 4, 4, 1.0, 1e-3, p, p, p, p, None, p),
                             L.AdtEvent3d(4, 4, 4, 1.0, 1e-3, p, p, p2, p, None, p))
    it3 = (L.AdtMisfit3d * 2)(L.AdtMisfit3d(4, 4, 4, p, p, p, None, p, p, 0.0),
                              L.AdtMisfit3d(4, 4, 4, p, p, p, None, p2, p, 0.0))
    assert lib.adt_graph_create_misfit(ev3, it3, 2, xyz, xyz, xyz, 1, both, 0, ctypes.byref(g),
                                       None) == -1
    assert b"share one u" in lib.adt_last_error()


def test_picks_and_coordinates_rejected_before_device_work():
    # the device cannot check the lengths of the pick arrays: the host
    # validates dtype, length and coordinate lengths first (CPU tensors here)
    import torch
    from adtomo_hip import objective as O
    u = [torch.zeros((4, 4, 4), dtype=torch.float64) for _ in range(2)]
    rx = ry = rz = [2.0, 3.0, 1.5]
    t64 = [torch.zeros(3, dtype=torch.float64) for _ in range(2)]
    with pytest.raises(ValueError, match="equal lengths"):
        O.misfit(u, rx, ry, [2.0, 3.0], t64, t64)
    with pytest.raises(ValueError, match="dtype"):
        O.misfit(u, rx, ry, rz, [t.float() for t in t64], t64)
    with pytest.raises(ValueError, match="values for 3 receivers"):
        O.misfit(u, rx, ry, rz, t64, [torch.zeros(2, dtype=torch.float64)] * 2)
    with pytest.raises(ValueError, match="one tobs and one weights"):
        O.misfit(u, rx, ry, rz, t64[:1], t64)
    for fn in (O.inversion_step, O.InversionStepGraph):
        with pytest.raises(ValueError, match="equal lengths"):
            fn(u, u, [1.0] * 2, [1e-3] * 2, rx, [2.0], rz, t64, t64)
        with pytest.raises(ValueError, match="values for 3 receivers"):
            fn(u, u, [1.0] * 2, [1e-3] * 2, rx, ry, rz, t64, [torch.zeros(4, dtype=torch.float64)] * 2)
        with pytest.raises(ValueError, match="dtype"):
            fn(u, u, [1.0] * 2, [1e-3] * 2, rx, ry, rz, [t.float() for t in t64], t64)
