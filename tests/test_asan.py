"""Host code under the address and undefined-behaviour sanitizers: the CPU
oracle's entry points and the drivers' host code (KS generation, delay
embedding, .npy / HDF5 I/O, point loading), each built with
-fsanitize=address,undefined and run on small inputs."""
import os
import subprocess

from kernel_evd_amd import build as b

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _check(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300, env=ENV)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout


def test_oracle_under_asan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True, timeout=300)
    _check([os.path.join(ROOT, "oracle", "_asan", "asan_check")])


def test_driver_host_code_under_asan(tmp_path):
    exe = b.build_asan()
    _check([exe, str(tmp_path)])
