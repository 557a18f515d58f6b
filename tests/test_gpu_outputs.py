"""integrate()'s records on the GPU: kineticEnergy.txt, staticBalance.txt and
the dnStat / TrackDamage extractor files the engine writes equal the
oracle's line for line (perf.txt holds wall-clock figures: its times only),
with conf outputs, on Rockable's 512-sphere deck, on hull polyhedra with the
device displacement check (batched guard reads, the extractors breaking the
batches) and with captured graphs."""
import os

import pytest

from tests.conftest import DECKS

pytestmark = pytest.mark.gpu


def _both(path, tmp_path, ext_lines, engine_opts=None, steps=-1, driving=None):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    res = {}
    for who in ("gpu", "cpu"):
        d = tmp_path / who
        d.mkdir()
        ext = d / "dataExtractors.txt"
        ext.write_text("\n".join(ext_lines) + "\n")
        out = d / "out"
        out.mkdir()
        if who == "gpu":
            e = Engine(path, driving=driving)
            for k, v in (engine_opts or {}).items():
                e.set_option(k, v)
            e.load_data_extractors(str(ext))
            e.update_nl()
            n = e.integrate(str(out), steps)
            e.close()
        else:
            o = Oracle(path, driving=driving)
            o.load_data_extractors(str(ext))
            o.update_nl()
            n = o.integrate(str(out), steps)
            o.close()
        res[who] = (d, n)
    return res


def _same_files(res, names):
    (dg, ng), (dc, nc) = res["gpu"], res["cpu"]
    assert ng == nc
    for name in names:
        if os.path.basename(name).startswith("conf"):
            from oracle.rk_oracle import compare_conf
            ok, k, x, y = compare_conf(os.path.join(dg, name), os.path.join(dc, name))
            assert ok, "%s differs at line %s:\nCPU %s\nGPU %s" % (name, k, x, y)
            continue
        a = open(os.path.join(dg, name)).read().splitlines()
        b = open(os.path.join(dc, name)).read().splitlines()
        assert len(a) == len(b) and len(b) > 0, name
        if name.endswith("perf.txt"):
            a = [ln.split()[0] for ln in a]
            b = [ln.split()[0] for ln in b]
        assert a == b, "%s: first difference %s" % (name, next((x, y) for x, y in zip(a, b) if x != y))


def test_outputs_512_deck(tmp_path):
    path = os.path.join(DECKS, "redirection_test_512_particles_bruteforce_LawDefault_noPeriodicity.conf")
    res = _both(path, tmp_path, ["dnStat dn.txt 1", "TrackDamage damage.txt 3"])
    _same_files(res, ["dn.txt", "damage.txt", "out/kineticEnergy.txt", "out/staticBalance.txt", "out/perf.txt"])


@pytest.mark.parametrize("opts", [{"record_geometry": 0, "guard_batch": 16},
                                  {"record_geometry": 0, "graph_steps": 8, "guard_batch": 64}])
def test_outputs_hulls_with_device_guard(tmp_path, opts):
    from rockable_amd import decks as D
    from tests.test_gpu_parity import _dynamic_hull_deck
    path = _dynamic_hull_deck(tmp_path)
    path = D.conf_with(path, str(tmp_path / "deck2.conf"), interConf=50 * 2e-7, tmax=600 * 2e-7)
    res = _both(path, tmp_path, ["dnStat dn.txt 7", "# a comment line", "dnStat dn2.txt 50"], opts)
    _same_files(res, ["dn.txt", "dn2.txt", "out/kineticEnergy.txt", "out/staticBalance.txt", "out/perf.txt",
                      "out/conf12"])
