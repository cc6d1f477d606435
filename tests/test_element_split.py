"""The plane-split HEX8 element equals the fused element bitwise (host compile of the device header, CPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, 'tests', 'cpp', 'element_split_check.cpp')
INC = os.path.join(ROOT, 'partial-green-jacobi_amd', 'csrc')


@pytest.mark.skipif(shutil.which('g++') is None, reason='g++ not available')
def test_plane_split_element_is_bitwise_equal(tmp_path):
    exe = tmp_path / 'check'
    subprocess.run(['g++', '-O2', '-std=c++17', '-I', INC, SRC, '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert 'max |diff| = 0' in out.stdout
