"""Runs the C++ test programs of the cusat_dsp API (tests/cpp, built by
cusat_amd.build into tests/cpp/_build): host checks on the CPU, the chain
and decoder checks on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build")


def _run(name, timeout):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail("%s is not built (run cusat_amd.build)" % exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


def test_cpp_host_checks():
    _run("qa_host", 120)


@pytest.mark.gpu
def test_cpp_rx_fec_chain_on_gpu():
    _run("qa_rx_fec_chain_gpu", 110)


@pytest.mark.gpu
def test_cpp_rx_fec_pool_on_gpu():
    _run("qa_rx_fec_pool_gpu", 110)


@pytest.mark.gpu
def test_cpp_rx_fec_stream_pool_on_gpu():
    _run("qa_rx_fec_stream_pool_gpu", 110)


@pytest.mark.gpu
def test_cpp_dd_snr_on_gpu():
    _run("qa_dd_snr_gpu", 110)
