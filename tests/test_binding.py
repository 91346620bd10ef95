"""The reference's own main program with NuOsc::step_rk4 on the library.

oracle/_ref/main_binding_<presets> is src/main.cpp of the reference,
unmodified, compiled with the binding of include/cosenu_nuosc_binding.hpp
applied to class NuOsc (oracle/ref_binding.hpp, built by `make -C oracle
binding`); oracle/_ref/main_<presets> is the same program as the reference
builds it.  Both run the same job.config in separate directories.  In STRICT
mode every file the program writes (G0.bin, state0.bin, rho_<t>.dat,
conserved_quantities.dat, survival_probability.dat) and its console output
must be byte for byte the reference's; only time.txt and the "elapsed time"
line (wall time) differ.  In FAST mode the dumped states agree normwise.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

CONFIG = """scheme : {scheme}
nz : 64
nvz : 16
z0 : -32.0
z1 : 32.0
v0 : -1.0
v1 : 1.0
CFL : 0.4
gz : {gz}
N_ITER : 41
ANAL_EVERY : 5
pmo : 1.0
omega : 0.3
theta : 37.0
mu : 1.0
perturbation_size : 0.001
n_dump_rho : 4
"""

# presets -> (scheme, gz) of the job
CASES = {"fv_vc": ("FV", 4), "fd_vc": ("FD", 2), "fv_open_vc": ("FV", 4), "fd_vmc": ("FD", 2)}


def _exe(name):
    path = os.path.join(REF, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (make -C oracle ref binding, with the reference checkout)")
    return path


def _run(exe, cwd, env=None, timeout=300):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([exe, "--conf", "job.config"], cwd=cwd, capture_output=True, text=True,
                          timeout=timeout, env=e)


def _outputs(d):
    return {n: (d / n).read_bytes() for n in sorted(os.listdir(d))
            if n not in ("job.config", "time.txt")}


def _console(text):
    return [l for l in text.splitlines() if not l.startswith("elapsed time:")]


def test_binding_fails_loudly_without_a_device(tmp_path):
    """No CPU fallback: with no GPU the first step reports the library's error
    in the reference's "[ FAIL ]" style and exits 1."""
    exe = _exe("main_binding_fv_vc")
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    (tmp_path / "job.config").write_text(CONFIG.format(scheme="FV", gz=4))
    r = _run(exe, tmp_path)
    assert r.returncode == 1, r.stdout[-2000:]
    assert "[ FAIL ]...cosenu:" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("presets", sorted(CASES))
def test_reference_main_with_the_binding_is_the_reference(gpu, tmp_path, presets):
    scheme, gz = CASES[presets]
    ref, hip = _exe(f"main_{presets}"), _exe(f"main_binding_{presets}")
    a, b = tmp_path / "ref", tmp_path / "hip"
    a.mkdir()
    b.mkdir()
    for d in (a, b):
        (d / "job.config").write_text(CONFIG.format(scheme=scheme, gz=gz))
    ra = _run(ref, a)
    rb = _run(hip, b, {"COSENU_BINDING_MODE": "strict"})
    print(f"{presets}: reference rc {ra.returncode}, binding rc {rb.returncode}", flush=True)
    assert ra.returncode == 0, ra.stdout[-2000:]
    assert rb.returncode == 0, rb.stdout[-2000:] + rb.stderr[-2000:]
    oa, ob = _outputs(a), _outputs(b)
    assert sorted(oa) == sorted(ob)
    assert "rho_40.dat" in oa and "conserved_quantities.dat" in oa
    for name in oa:
        assert oa[name] == ob[name], name
    assert _console(ra.stdout) == _console(rb.stdout)


def _rho(path):
    return np.loadtxt(path)


@pytest.mark.gpu
@pytest.mark.parametrize("presets", ["fv_vc", "fd_vc"])
def test_reference_main_with_the_binding_fast_mode(gpu, tmp_path, presets):
    scheme, gz = CASES[presets]
    ref, hip = _exe(f"main_{presets}"), _exe(f"main_binding_{presets}")
    a, b = tmp_path / "ref", tmp_path / "hip"
    a.mkdir()
    b.mkdir()
    for d in (a, b):
        (d / "job.config").write_text(CONFIG.format(scheme=scheme, gz=gz))
    assert _run(ref, a).returncode == 0
    rb = _run(hip, b, {"COSENU_BINDING_MODE": "fast"})
    assert rb.returncode == 0, rb.stdout[-2000:]
    for t in (10, 20, 30, 40):
        ra_, rb_ = _rho(a / f"rho_{t}.dat"), _rho(b / f"rho_{t}.dat")
        assert np.array_equal(ra_[:, :2], rb_[:, :2])  # z, v columns
        rel = np.max(np.abs(rb_[:, 2:] - ra_[:, 2:])) / np.max(np.abs(ra_[:, 2:]))
        assert rel <= 1e-10, (t, rel)
