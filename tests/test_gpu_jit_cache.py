"""GPU: the persistent code-object cache (BOLT_JIT_CACHE_DIR), the role of
src/jit/disk_cache.rs.  A second process loads the generated kernel from the
directory instead of compiling it and gets the same result; a file whose
header or body does not match is recompiled, never loaded; a directory that
other users can write is not used (disk_cache.rs:70-90)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# One process: a filtered scalar aggregate (a hipRTC-generated kernel), then
# the cache counters, as one JSON line.
CHILD = r"""
import ctypes as C, json, sys
import numpy as np
sys.path.insert(0, %r)
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import Term, execute_aggregate
x = np.arange(1_000_000, dtype=np.int64) %% 977
r = execute_aggregate([DeviceColumn.from_numpy(x)], [(N.SUM, 0), (N.COUNT_STAR, 0)], [Term(0, N.GT, 500)])
c, h, s = C.c_int64(), C.c_int64(), C.c_int64()
N.lib().bolt_jit_cache_stats(C.byref(c), C.byref(h), C.byref(s))
print(json.dumps({"res": r, "compiled": c.value, "hits": h.value, "stores": s.value}))
""" % ROOT


def run_child(cache_dir):
    env = dict(os.environ, BOLT_JIT_CACHE_DIR=str(cache_dir), BOLT_JIT="1")
    p = subprocess.run([sys.executable, "-c", CHILD], env=env, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return json.loads(p.stdout.strip().splitlines()[-1])


def cache_files(d):
    return sorted(f for f in os.listdir(d) if f.startswith("bolt_") and f.endswith(".co"))


def test_second_process_loads_from_the_cache(gpu, tmp_path):
    d = tmp_path / "co"
    d.mkdir(mode=0o700)
    first = run_child(d)
    assert first["compiled"] >= 1 and first["stores"] == first["compiled"] and first["hits"] == 0
    files = cache_files(d)
    assert len(files) == first["stores"]
    assert not [f for f in os.listdir(d) if ".tmp." in f]
    second = run_child(d)
    assert second["res"] == first["res"]
    assert second["compiled"] == 0 and second["hits"] == first["compiled"]
    # a damaged file (body byte flipped past the header, or a header field)
    # is rejected and the kernel recompiled and rewritten
    for off in (-1, 8):
        p = d / files[0]
        b = bytearray(p.read_bytes())
        if off < 0:
            b = b[:-1]  # truncated body
        else:
            b[off] ^= 0xFF  # the key field
        p.write_bytes(bytes(b))
        third = run_child(d)
        assert third["res"] == first["res"]
        assert third["compiled"] == 1 and third["hits"] == first["compiled"] - 1


def test_untrusted_directory_is_not_used(gpu, tmp_path):
    d = tmp_path / "open"
    d.mkdir()
    os.chmod(d, 0o777)
    r = run_child(d)
    assert r["stores"] == 0 and r["hits"] == 0 and r["compiled"] >= 1
    assert cache_files(d) == []
