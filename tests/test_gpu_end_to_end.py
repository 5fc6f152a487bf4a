"""runMPMYDW end to end on the device (driver.main): a JSON config, the
reference's HDF5 domain format (tests/golden), the reference's defaults
(geostatic F, volumetric locking correction), snapshots and a checkpoint, then
a restart from that checkpoint that resumes at the next big step."""
import gzip
import glob
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "coupling_column_same.h5.gz")


def test_config_run_checkpoint_and_restart(tmp_path):
    from mpmwarpydw_amd import driver, fs5PlotUtils
    dom = tmp_path / "domain.h5"
    dom.write_bytes(gzip.decompress(open(GOLDEN, "rb").read()))
    out = tmp_path / "out"
    cfg = {"domainFile": str(dom), "dt": 1e-4, "dtxpbd": 1e-3, "bigStepDuration": 2.5e-3, "bigSteps": 1,
           "xpbdOnlyDuration": 2.5e-3, "saveFlag": 1, "outputFolder": str(out), "render_interval": 1e-3}
    (tmp_path / "c.json").write_text(json.dumps(cfg))
    print("[end to end] first run", flush=True)
    rec = driver.main(["--config", str(tmp_path / "c.json")], verbose=False)
    assert [r["bigStep"] for r in rec] == [0] and rec[0]["phase1_steps"] == 25 and rec[0]["phase2_steps"] == 2
    vtps = sorted(glob.glob(str(out / "*_particles.vtp")))
    assert len(glob.glob(str(out / "*_grid.vti"))) == len(vtps) >= 3
    ck = fs5PlotUtils.load_restart_from_vtp(vtps[-1])  # the big-step checkpoint is written last
    assert ck["bigStep"] == 0 and abs(ck["sim_time"] - rec[0]["t"]) == 0.0
    assert np.isfinite(ck["particle_F"]).all() and not np.allclose(ck["particle_F"], np.eye(3))  # geostatic
    cfg.update({"bigSteps": 2, "restart": vtps[-1], "outputFolder": str(tmp_path / "out2")})
    (tmp_path / "c2.json").write_text(json.dumps(cfg))
    print("[end to end] restart", flush=True)
    rec2 = driver.main(["--config", str(tmp_path / "c2.json")], verbose=False)
    assert [r["bigStep"] for r in rec2] == [1]
    assert rec2[0]["t"] > rec[0]["t"]
