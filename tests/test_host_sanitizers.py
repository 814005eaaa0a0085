"""The library's host code built with AddressSanitizer and
UndefinedBehaviorSanitizer (host only: each -fsanitize= goes after
-Xarch_host, the device code is built as always) and driven through the C
ABI by tests/host_sanitizer_check.cpp: parameter checks, run-folder names,
failure without a device and the HDF5 snapshot writer with its deflate
threads.  Runs on the CPU; the build takes about a minute."""
import concurrent.futures as cf
import os
import shutil
import subprocess

import pytest

import coolbubble_amd as cb
from coolbubble_amd import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined", "-g"]


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc absent")
def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    hipcc = B.hipcc()
    objs = []

    def compile_one(src):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        cmd = [hipcc, "-x", "hip", "--offload-arch=" + B.ARCH] + B.COMMON + SAN + ["-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return obj

    with cf.ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, B.sources() + [os.path.join(HERE, "host_sanitizer_check.cpp")]))
    exe = str(tmp_path / "host_check")
    cmd = [hipcc, "--offload-arch=" + B.ARCH] + SAN + objs + ["-o", exe, "-lz", "-ldl", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "0 failure(s)" in r.stdout
    if cb.lib().cb_hdf5_available() == 0:
        assert os.path.exists(str(tmp_path / "snap_3.h5"))
