"""The C++ facade (include/skyweaver_amd.hpp) on the host: DedispersionPlan
parsing, channel frequencies, DADA headers and MultiFileWriter file layout
(tests/cpp/facade_host_test.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run __graft_entry__.build() (make -C tests/cpp)")
    return path


def test_facade_host(tmp_path):
    r = subprocess.run([_bin("facade_host_test"), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "passed" in r.stdout


def test_dada_reader(tmp_path):
    """read_dada_header on the reference fixture's headers and MultiFileReader
    over ten files built from them (cpp/skyweaver/test/src/MultiFileReaderTester.cu)."""
    r = subprocess.run([_bin("dada_reader_test"), os.path.join(HERE, "golden", "dada"), str(tmp_path)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "passed" in r.stdout
