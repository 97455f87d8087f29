"""Host logic of the library that needs no GPU: NTT pass plans and the proof
server's result ring, compiled with g++ from their headers and run here."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_host_logic(tmp_path):
    exe = tmp_path / "host_logic_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", os.path.join(HERE, "host", "host_logic_test.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True)
    assert "host logic: ok" in out.stdout
