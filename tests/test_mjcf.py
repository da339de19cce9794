"""MJCF-subset loader: XML text compiles to the same System as the builders.

Two systems compile to the same physics when their float64 oracle
trajectories agree to the last bit; that checks every array the step reads
(tree, frames, inertias, joint parameters, contact pairs, actuators, options)
without naming them one by one.
"""
import os

import numpy as np
import pytest

from brax_amd import systems as S
from oracle.cstep import CStep

_TRIPLE = """
<mujoco model="triple pendulum">
  <option timestep="0.001"/>
  <worldbody>
    <body name="body1" pos="0 0 0">
      <joint name="hinge1" type="hinge" axis="1 0 0"/>
      <geom name="sphere1" type="sphere" size="0.15" pos="0 0.5 0" mass="1"
            contype="0" conaffinity="0"/>
      <body name="body2" pos="0 0.5 0">
        <joint name="hinge2" type="hinge" axis="1 0 0"/>
        <geom name="sphere2" type="sphere" size="0.15" pos="0 0.5 0" mass="1"
              contype="0" conaffinity="0"/>
        <body name="body3" pos="0 0.5 0">
          <joint name="hinge3" type="hinge" axis="1 0 0"/>
          <geom name="sphere3" type="sphere" size="0.15" pos="0 0.5 0" mass="1"
                contype="0" conaffinity="0"/>
        </body>
      </body>
    </body>
  </worldbody>
  <actuator>
    <motor joint="hinge1" gear="150" ctrlrange="-100 100"/>
  </actuator>
</mujoco>
"""

# brax/envs/assets/ant.xml, read where the reference tree is present.
_ANT_XML = "/root/reference/brax/envs/assets/ant.xml"


def _same_physics(a, b, steps=5, n=8, seed=0):
  rng = np.random.default_rng(seed)
  q0 = b.init_q + rng.uniform(-0.1, 0.1, (n, b.q_size()))
  qd0 = 0.1 * rng.standard_normal((n, b.qd_size()))
  acts = rng.uniform(-1, 1, (steps, n, b.act_size()))
  qa, qda = CStep(a).rollout(q0, qd0, acts, steps, dtype=np.float64)
  qb, qdb = CStep(b).rollout(q0, qd0, acts, steps, dtype=np.float64)
  np.testing.assert_array_equal(qa, qb)
  np.testing.assert_array_equal(qda, qdb)


def test_triple_pendulum_xml_matches_builder():
  a = S.load_mjcf(_TRIPLE)
  b = S.pendulum(3, motor_gear=150.0)
  assert a.link_types == b.link_types and a.link_parents == b.link_parents
  _same_physics(a, b)


@pytest.mark.skipif(not os.path.exists(_ANT_XML), reason="reference assets absent")
def test_ant_xml_matches_builder():
  with open(_ANT_XML) as f:
    a = S.load_mjcf(f.read())
  b = S.ant()
  assert a.link_types == b.link_types and a.link_parents == b.link_parents
  np.testing.assert_allclose(a.init_q, b.init_q)
  _same_physics(a, b, steps=10, n=16)


def test_radian_angles_and_joint_defaults():
  """compiler angle="radian" and a default joint class reach the hinge."""
  deg = _TRIPLE.replace('<option timestep="0.001"/>',
                        '<option timestep="0.001"/><default><joint damping="0.5" '
                        'range="-30 30" limited="true"/></default>')
  rad = deg.replace("<option", '<compiler angle="radian"/><option').replace(
      'range="-30 30"', f'range="{-np.pi / 6!r} {np.pi / 6!r}"')
  a, b = S.load_mjcf(deg), S.load_mjcf(rad)
  _same_physics(a, b)
  # damping and limits changed the physics against the plain pendulum
  with pytest.raises(AssertionError):
    _same_physics(a, S.pendulum(3, motor_gear=150.0))


@pytest.mark.parametrize("edit", [
    ('type="hinge"', 'type="slide"'),
    ('type="sphere"', 'type="box"'),
    ('<joint name="hinge3"', '<inertial mass="1" pos="0 0 0"/><joint name="hinge3"'),
    ('<motor joint', '<general joint'),
])
def test_outside_subset_raises(edit):
  with pytest.raises(NotImplementedError):
    S.load_mjcf(_TRIPLE.replace(*edit, 1))


def test_not_mjcf_raises():
  with pytest.raises(ValueError):
    S.load_mjcf("<robot/>")


@pytest.mark.parametrize("edit", [
    ('<option timestep="0.001"/>', '<option timestep="0.001" integrator="RK4"/>'),
    ('<body name="body2" pos="0 0.5 0">', '<body name="body2" pos="0 0.5 0" euler="0 0 10">'),
    ('contype="0" conaffinity="0"/>', 'contype="0" conaffinity="0" condim="1"/>'),
    ('<joint name="hinge2"', '<joint name="hinge2" margin="0.1"'),
    ('<mujoco model="triple pendulum">', '<mujoco><compiler coordinate="global"/>'),
    ('<actuator>', '<equality><joint joint1="hinge1" joint2="hinge2"/></equality><actuator>'),
    ('<actuator>', '<include file="more.xml"/><actuator>'),
    ('<option timestep="0.001"/>', '<option timestep="0.001"><flag contact="disable"/></option>'),
])
def test_unread_physics_attributes_raise(edit):
  with pytest.raises(NotImplementedError):
    S.load_mjcf(_TRIPLE.replace(*edit, 1))


_REF = "/root/reference/brax"
_needs_ref = pytest.mark.skipif(not os.path.isdir(_REF), reason="reference assets absent")


def _load(rel):
  with open(os.path.join(_REF, rel)) as f:
    return S.load_mjcf(f.read())


@_needs_ref
def test_reference_triple_pendulum_fixture():
  """Known answers of brax/io/mjcf_test.py:27-50 on test_data/triple_pendulum.xml."""
  s = _load("test_data/triple_pendulum.xml")
  np.testing.assert_allclose(s.gravity, [0, 0, -9.81])
  assert list(s.link_names) == ["body1", "body2", "body3"]
  np.testing.assert_allclose(s.link.transform.pos, [[0, 0, 0], [0, 0.5, 0], [0, 0.5, 0]])
  np.testing.assert_allclose(s.link.transform.rot, np.tile([1.0, 0, 0, 0], (3, 1)))
  np.testing.assert_allclose(s.link.inertia.i, np.tile(np.eye(3), (3, 1, 1)) * 0.009,
                             atol=1e-12)
  np.testing.assert_allclose(s.link.inertia.transform.pos, np.tile([0, 0.5, 0], (3, 1)))
  np.testing.assert_allclose(s.link.inertia.mass, [1, 1, 1])
  assert s.link_types == "111" and tuple(s.link_parents) == (-1, 0, 1)


@_needs_ref
def test_reference_ant_names_and_init_q():
  """brax/io/mjcf_test.py:52-61 and brax/kinematics_test.py:69-75."""
  s = _load("envs/assets/ant.xml")
  assert list(s.link_names) == ["torso", "aux_1", "", "aux_2", "", "aux_3", "", "aux_4", ""]
  assert s.link_types == "f11111111"
  assert tuple(s.link_parents) == (-1, 0, 1, 0, 3, 0, 5, 0, 7)
  np.testing.assert_allclose(
      s.init_q, [0, 0, 0.55, 1, 0, 0, 0, 0, 1, 0, -1, 0, -1, 0, 1], atol=1e-7)


@_needs_ref
def test_reference_humanoid_names():
  """brax/io/mjcf_test.py:63-82 on envs/assets/humanoid.xml, plus its sizes
  and options (humanoid.xml:2-24): 17 motors, timestep 0.003, 8 iterations,
  20 Newton-Schulz iterations, 15 line-search steps."""
  s = _load("envs/assets/humanoid.xml")
  assert list(s.link_names) == [
      "torso", "lwaist", "pelvis", "right_thigh", "right_shin", "left_thigh",
      "left_shin", "right_upper_arm", "right_lower_arm", "left_upper_arm",
      "left_lower_arm"]
  assert s.link_types == "f2131312121"
  assert s.q_size() == 24 and s.qd_size() == 23 and s.act_size() == 17
  assert s.opt.timestep == pytest.approx(0.003) and s.solver_iterations == 8
  assert s.matrix_inv_iterations == 20 and s.solver_maxls == 15


@_needs_ref
def test_reference_actuator_fixtures():
  """Position / velocity / motor actuators of brax/test_data (as
  actuator_test.py:50-123 loads them) compile to the builder's servos:
  position kp -> gain kp, bias_q -kp; velocity kv -> gain kv, bias_qd -kv
  (brax/io/mjcf.py:394-423 reading MuJoCo's gainprm / biasprm)."""
  a = _load("test_data/single_pendulum_position_frclimit.xml")
  b = S.pendulum(1, position_servo=(10.0, 10.0, (-2.5, 3.1)), timestep=0.001)
  _same_physics(a, b)
  a = _load("test_data/single_pendulum_motor.xml")
  b = S.pendulum(1, motor_gear=150.0, ctrlrange=(-100.0, 100.0), timestep=0.001)
  _same_physics(a, b)
  v = _load("test_data/single_pendulum_velocity.xml").actuator
  np.testing.assert_allclose([v.gain[0], v.bias_q[0], v.bias_qd[0]], [10.0, 0.0, -10.0])
  assert np.isinf(v.ctrl_range).all() and np.isinf(v.force_range).all()
  p = _load("test_data/single_pendulum_position.xml").actuator
  np.testing.assert_allclose([p.gain[0], p.gear[0], p.bias_q[0]], [10.0, 1.0, -10.0])
  t = _load("test_data/triple_pendulum_motor.xml")
  assert t.act_size() == 3 and t.link_types == "111"
  np.testing.assert_allclose(t.actuator.ctrl_range, np.tile([-100.0, 100.0], (3, 1)))


def test_humanoid_xml_builder_equals_loaded_file():
  """systems.humanoid_xml() (the transcription the GPU tests step, where the
  reference tree is absent) compiles to exactly the System that load_mjcf
  makes of brax/envs/assets/humanoid.xml: every array bitwise equal."""
  import dataclasses

  built = S.humanoid_xml()
  loaded = _load("envs/assets/humanoid.xml")

  def diff(x, y, path=""):
    if dataclasses.is_dataclass(x):
      return sum((diff(getattr(x, f.name), getattr(y, f.name), f"{path}.{f.name}")
                  for f in dataclasses.fields(x)), [])
    if isinstance(x, (np.ndarray, list, tuple)):
      xa, ya = np.asarray(x), np.asarray(y)
      same = xa.shape == ya.shape and np.array_equal(xa, ya, equal_nan=xa.dtype.kind == "f")
      return [] if same else [path]
    return [] if (x == y or path == ".name") else [path]

  assert diff(built, loaded) == []
  assert built.link_types == "f2131312121" and built.qd_size() == 23
  assert built.num_contacts() == 2 and built.num_limit_rows() == 17
