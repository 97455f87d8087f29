"""Host logic of the library that needs no GPU: NTT pass plans and the proof
server's result ring, compiled with g++ from their headers and run here, once
plain and once under AddressSanitizer and UndefinedBehaviorSanitizer; and the
oracle's main paths under the same sanitizers."""

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g"]


@pytest.mark.parametrize("sanitize", [False, True])
def test_host_logic(tmp_path, sanitize):
    exe = tmp_path / "host_logic_test"
    flags = SANITIZE if sanitize else []
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", *flags,
                    os.path.join(HERE, "host", "host_logic_test.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True)
    assert "host logic: ok" in out.stdout
    assert "runtime error" not in out.stderr


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitize_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-Wall", *SANITIZE,
                    os.path.join(HERE, "host", "oracle_sanitize_test.cpp"), os.path.join(ROOT, "oracle", "oracle.cpp"),
                    "-o", str(exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, env=env, timeout=600)
    assert "oracle under sanitizers: ok" in out.stdout
    assert "runtime error" not in out.stderr
