"""The reference's driver flow (main.cpp:24-107 with the field block that wrote
the golden files) as a C++ program over bem.hpp: its output file must match
sphere171Hz_CHIEF and the analytic pulsating sphere (sphere.m:11)."""
import os
import subprocess

import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C

pytestmark = pytest.mark.gpu

APP = os.path.join(bem.PKG_DIR, "apps", "pulsating_sphere")


def test_reference_driver_reproduces_chief_golden(tmp_path):
    assert os.path.exists(APP), "driver not built"
    out = tmp_path / "p.txt"
    r = subprocess.run([APP, os.path.join(C.GOLDEN, "sphere.obj"), str(out), "1", "1"], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    got = C.read_golden(str(out))
    gold = C.read_golden("sphere171Hz_CHIEF")
    k = float(np.float32(2 * np.pi * 171.5 / 343.21))
    ana = C.pulsating_sphere(np.arange(1.0, 3.01, 0.2), 1.0, k, 1.2041, 343.21)
    assert got.shape == (11,)
    assert np.linalg.norm(got - gold) / np.linalg.norm(gold) < 0.015
    assert np.linalg.norm(got - ana) / np.linalg.norm(ana) < 0.025
