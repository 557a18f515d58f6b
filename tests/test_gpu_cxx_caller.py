"""The C++ caller of INTEGRATION.md (rockable_amd/examples/rk_run.cpp,
compiled against include/rockable_amd.h alone) runs Rockable's
512-sphere regression deck to its end on the GPU; its last conf file equals
the reference's golden line for line (compareConf, src/Apps/run.cpp:38-92)."""
import os
import subprocess

import pytest

from tests.conftest import DECKS, ROOT

pytestmark = pytest.mark.gpu


def test_cxx_caller_reproduces_the_512_golden(tmp_path):
    from oracle.rk_oracle import compare_conf
    exe = os.path.join(ROOT, "rockable_amd", "rk_run")
    conf = os.path.join(DECKS, "redirection_test_512_particles_bruteforce_LawDefault_noPeriodicity.conf")
    r = subprocess.run([exe, conf, str(tmp_path), "-1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "steps" in r.stdout
    ok, k, a, b = compare_conf(str(tmp_path / "conf101"),
                               os.path.join(DECKS, "regression", "regression_512_particles_LawDefault_noPeriodicity.conf"))
    assert ok, "first difference at line %s:\nREF %s\nGPU %s" % (k, a, b)
