"""The C ABI library: it loads, exports every entry point include/mq_engine.h
declares, and its host-only entry points behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mqhip
import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(mq_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = mqhip.load()
    names = declared_functions()
    assert {"mq_engine_create", "mq_run_batch", "mq_run_batch_ions", "mq_summarize", "mq_last_error"} <= names
    for n in sorted(names):
        assert hasattr(lib, n), n
    # the Python binding declares a signature for every exported entry point
    assert names <= set(mqhip.SIGNATURES), names - set(mqhip.SIGNATURES)


def test_summarize_matches_numpy():
    rng = np.random.default_rng(3)
    n, nj = 10000, 9
    st = rng.integers(1, 6, n).astype(np.uint8)
    steps = rng.integers(0, 20000, n).astype(np.uint32)
    nc = rng.integers(0, 40, n).astype(np.uint32)
    job = rng.integers(0, nj, n).astype(np.uint32)
    out = mqhip.summarize(st, nc, steps, job, nj, 1e-7)
    for j in range(nj):
        s = job == j
        o = out[j]
        assert o.n == s.sum()
        assert [o.crashed, o.returned, o.transmitted, o.left, o.timeout] == [int((st[s] == k).sum()) for k in range(1, 6)]
        assert o.mean_time == pytest.approx(steps[s].mean() * 1e-7, rel=1e-14)
        assert o.mean_coll == pytest.approx(nc[s].mean(), rel=1e-14)


def test_summarize_rejects_bad_job_index():
    with pytest.raises(mqhip.MqError, match="job index"):
        mqhip.summarize(np.ones(3, np.uint8), np.zeros(3, np.uint32), np.zeros(3, np.uint32),
                        np.array([0, 5, 1], np.uint32), 2, 1e-7)


def test_engine_create_validates_arguments():
    lib = mqhip.load()
    h = C.c_void_p()
    assert lib.mq_engine_create(None, C.byref(h)) == -1
    assert b"NULL" in lib.mq_last_error()
    g = synth.ideal_channel(8, 8, 8, 1e-3)
    g["nx"] = 1
    with pytest.raises(mqhip.MqError, match="at least 2"):
        mqhip.Engine(g)
    g = synth.ideal_channel(8, 8, 8, 1e-3)
    g["h"] = 0.0
    with pytest.raises(mqhip.MqError, match="positive"):
        mqhip.Engine(g)


def test_engine_create_without_gpu_fails_cleanly():
    """On a host without a HIP device the call returns an error code, never a fallback."""
    g = synth.ideal_channel(8, 8, 8, 1e-3)
    try:
        e = mqhip.Engine(g)
    except mqhip.MqError as err:
        assert "HIP" in str(err) or "device" in str(err)
        return
    e.close()  # a GPU is present: creation succeeded


def test_legacy_dims_are_host_side_and_match_the_oracle(oracle_mod):
    o = mqhip.LegacyOpts.default()
    assert (o.scl, o.precision_pct, o.max_iter) == (10.0, np.float32(0.001), 6000)
    for scl in (10.0, 2.0, 2.5, 4.0, 7.3):
        assert mqhip.legacy_dims(mqhip.LegacyOpts.default(scl=scl)) == oracle_mod.legacy_dims(scl)
    with pytest.raises(mqhip.MqError, match="scl"):
        mqhip.legacy_dims(mqhip.LegacyOpts.default(scl=-1.0))


def test_fields_from_theirs_reports_missing_files(tmp_path):
    with pytest.raises(mqhip.MqError, match="eject_RF.f32"):
        mqhip.fields_from_theirs(str(tmp_path))
    # a short file is refused with its expected size
    (tmp_path / "eject_RF.f32").write_bytes(b"\0" * 16)
    with pytest.raises(mqhip.MqError, match="expected 24780800 bytes"):
        mqhip.fields_from_theirs(str(tmp_path))
