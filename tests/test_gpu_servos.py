"""The servo decks of tests/test_oracle_servos.py on the GPU: tritriCustom,
tritriLodeAngle, the three shakers and the ramp, and the re-read of the
driving file at a conf save.  The engine is bit-identical to the oracle on
every one (the shakers' and the ramp's control values, functions of the time
only, are evaluated on the host with the C library, as the oracle does)."""
import pytest

from rockable_amd import decks as D
from tests import kat_decks as K
from tests.test_gpu_parity import assert_bodies_equal
from tests.test_oracle_pins import servo_box_deck

pytestmark = pytest.mark.gpu


def _servo_case(name, d):
    if name == "custom":
        return servo_box_deck(d, "tritriCustom 0 0.1 1 40.0 0 0.0 0 0.05 1 30.0 0 0.0")
    if name == "lode":
        return servo_box_deck(d, "tritriLodeAngle 40.0 45.0 200.0")
    if name == "isostatic":
        return servo_box_deck(d, "tritriIsostaticCompression 25.0")
    lines = {"shaker": "Servo shaker 0 1 2 2 0.02 25.0",
             "triangle": "Servo triangle_shaker 0 0 0 1 0.01 20.0",
             "sawtooth": "Servo sawtooth_shaker 0 0 1 0 0.01 20.0 0.013",
             "ramp": "Servo ramp _y_Vel_ 0 0.1 -0.3 0.01 0.03"}[name]
    path, drv = K.driven_plate(d, [lines, "Control _zrot_Vel_ 0 0.5"], dt=1e-5)
    return path, drv, 4000


@pytest.mark.parametrize("name", ["custom", "lode", "isostatic", "shaker", "triangle", "sawtooth", "ramp"])
def test_servo_bit_exact(tmp_path, name):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    path, drv, steps = _servo_case(name, str(tmp_path))
    e = Engine(path, driving=drv)
    o = Oracle(path, driving=drv)
    e.update_nl()
    o.update_nl()
    for _ in range(4):
        e.step(steps // 4)
        o.run_steps(steps // 4)
        assert_bodies_equal(e.bodies(), o.bodies())


@pytest.mark.parametrize("graph_steps", [0, 8])
def test_driving_file_reread_at_conf_save(tmp_path, graph_steps):
    """A control changed in the driving file during the run takes effect
    after the next conf save, on the GPU as in the oracle, whether the
    steps go out one by one or as captured graphs."""
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    dt = 1e-5
    path, drv = K.driven_plate(str(tmp_path), ["Control _x_Vel_ 0 0.1", "Control _z_For_ 0 40.0"], dt=dt)
    path = D.conf_with(path, str(tmp_path / "deck2.conf"), interConf=100 * dt)
    e = Engine(path, driving=drv)
    e.set_option("graph_steps", graph_steps)
    o = Oracle(path, driving=drv)
    e.update_nl()
    o.update_nl()
    e.step(50)
    o.run_steps(50)
    with open(drv, "w") as fh:
        fh.write("Control _x_Vel_ 0 -0.2\nControl _z_For_ 0 -25.0\nControl _yrot_Vel_ 0 0.3\n")
    for _ in range(3):
        e.step(50)
        o.run_steps(50)
        assert_bodies_equal(e.bodies(), o.bodies())
    assert e.bodies()["vel"][0, 0] == -0.2
