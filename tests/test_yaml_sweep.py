"""The YAML sweep front-end (cosenu_amd/yaml_sweep.py): job expansion as the
reference's manage.py configure() derives it (manage.py:213-256), and on the
GPU the batched run against the C++ driver running each job alone."""
import os
import subprocess

import numpy as np
import pytest
import yaml

from cosenu_amd import yaml_sweep as Y

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
YML = os.path.join(ROOT, "tests", "golden", "sweep_small.yaml")
COSENU = os.path.join(ROOT, "cosenu_amd", "bin", "cosenu")


def test_expansion_follows_manage_py(tmp_path):
    assert Y.main([YML, "--scheme", "fv", "--out", str(tmp_path)]) == 0
    cfg = yaml.safe_load(open(YML))
    jobs = Y.expand(cfg, "fv")
    assert [j for j, _ in jobs] == ["64_16_0.4", "64_16_0.5", "64_24_0.4", "64_24_0.5",
                                    "128_16_0.4", "128_16_0.5", "128_24_0.4", "128_24_0.5"]
    for jid, job in jobs:
        nz, nvz, cfl = jid.split("_")
        dz = 80.0 / int(nz)
        dt = abs(float(cfl) * dz / 1.0)
        assert job["dz"] == dz and job["dt"] == dt and job["gz"] == 4
        assert job["N_ITER"] == int(8 / dt)
        assert job["ANAL_EVERY"] == max(1, int(job["N_ITER"] / 4))
        text = (tmp_path / jid / "job.config").read_text().splitlines()
        assert text[0].startswith("scheme") and text[0].endswith(": FV ")
        assert any(l.startswith("boundary : periodic") for l in text)
    assert len((tmp_path / "jobs_list.txt").read_text().splitlines()) == 8
    assert Y.expand(cfg, "fd")[0][1]["gz"] == 2


def test_reference_factors_reproduce_reference_initial_state():
    from cosenu_amd.initial import reference_ic
    import cosenu_amd as C
    p = C.RunParams(nz=64, nvz=16, z0=-40, z1=40)
    a, b, G, bG = Y.reference_factors(p, 1e-3)
    G2, bG2, s = reference_ic(p, 1e-3)
    st = a[:, :, None] * b[:, None, :]
    assert np.array_equal(st, s) and np.array_equal(np.signbit(st), np.signbit(s))
    assert np.array_equal(G, G2[:, 0]) and np.array_equal(bG, bG2[:, 0])


@pytest.mark.gpu
def test_batched_sweep_matches_driver(gpu, tmp_path):
    out = tmp_path / "sweep"
    assert Y.main([YML, "--scheme", "fv", "--out", str(out), "--run"]) == 0
    for jid in ("64_16_0.4", "128_24_0.5"):
        solo = tmp_path / ("solo_" + jid)
        solo.mkdir()
        cfg = (out / jid / "job.config").read_text() + "mode : fast\n"
        (solo / "job.config").write_text(cfg)
        r = subprocess.run([COSENU, "--conf", "job.config"], cwd=solo, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-1500:]
        for name in ("conserved_quantities.dat", "survival_probability.dat"):
            assert (out / jid / name).read_text() == (solo / name).read_text(), (jid, name)
        dumps = sorted(f for f in os.listdir(solo) if f.startswith("rho_"))
        assert dumps and dumps == sorted(f for f in os.listdir(out / jid) if f.startswith("rho_"))
        for f in dumps:
            assert (out / jid / f).read_bytes() == (solo / f).read_bytes(), (jid, f)
