"""bench.py command line (no GPU): the one-GPU build refuses --gpus N > 1
with a message before touching the device, and every documented option
parses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench('--gpus', '2')
    assert r.returncode == 2
    assert 'one GPU' in r.stderr and r.stdout == ''


def test_options_parse():
    r = _bench('--help')
    assert r.returncode == 0
    for opt in ('--steps', '--warmup', '--batch', '--precision', '--workload', '--dmrg-sites',
                '--chi-init', '--two-site-sweeps', '--pipelines', '--serving-requests',
                '--serving-rate', '--serving-max-active', '--roofline-only', '--no-cpu'):
        assert opt in r.stdout, opt
