"""System arrays for the generalized-coordinates step, and model builders.

``System`` holds exactly the fields the reference's generalized pipeline reads
(brax/base.py:415-540; filled by brax/io/mjcf.py:317-489), with the same names
and nesting (``sys.link.inertia.i``, ``sys.dof.motion.ang``,
``sys.actuator.gear`` ...), as float64 numpy arrays on the host.  The device
upload (``brax_amd.generalized.pipeline``) converts them to float32 tables.

The builders replace the MuJoCo compiler for the models this project runs.
``ModelBuilder`` is a small compiler for the MJCF subset those models use:
bodies with a free joint or up to three hinges, sphere/capsule/plane geoms,
motors.  It derives the compiler quantities the step consumes from documented
formulas:

* geom mass and inertia from geometry and density (sphere, capsule as
  cylinder plus two hemispheres), combined per body by the parallel-axis
  theorem, principal axes from the eigen-decomposition;
* ``body_invweight0`` (link.invweight) = trace of the 3x3 translational block
  of J_com M^-1 J_com^T / 3 at qpos0, floored at 1e-15;
* ``dof_invweight0`` = diag(M^-1) at qpos0, averaged over the translational
  and the rotational triples of a free joint;
* collision pairs and their mixed parameters with MJX's rules (contype /
  conaffinity filter, no same-body or non-world parent/child pairs, friction =
  elementwise max, solref/solimp = solmix-weighted mean, plane first).
"""

from __future__ import annotations

import dataclasses
import itertools
from typing import List, Optional, Tuple

import numpy as np

# brax/base.py:33-34
Q_WIDTHS = {"f": 7, "1": 1, "2": 2, "3": 3}
QD_WIDTHS = {"f": 6, "1": 1, "2": 2, "3": 3}

# MuJoCo geom type enumeration (mjtGeom)
GEOM_PLANE = 0
GEOM_SPHERE = 2
GEOM_CAPSULE = 3

# pair kinds handled by the step (MJX narrow phase), keyed by (type1, type2)
# with type1 <= type2, and their contact counts
PAIR_PLANE_SPHERE = 0
PAIR_PLANE_CAPSULE = 1
PAIR_SPHERE_SPHERE = 2
PAIR_SPHERE_CAPSULE = 3
PAIR_CAPSULE_CAPSULE = 4
_PAIR_KIND = {(GEOM_PLANE, GEOM_SPHERE): PAIR_PLANE_SPHERE,
              (GEOM_PLANE, GEOM_CAPSULE): PAIR_PLANE_CAPSULE,
              (GEOM_SPHERE, GEOM_SPHERE): PAIR_SPHERE_SPHERE,
              (GEOM_SPHERE, GEOM_CAPSULE): PAIR_SPHERE_CAPSULE,
              (GEOM_CAPSULE, GEOM_CAPSULE): PAIR_CAPSULE_CAPSULE}
_PAIR_NCON = {PAIR_PLANE_SPHERE: 1, PAIR_PLANE_CAPSULE: 2, PAIR_SPHERE_SPHERE: 1,
              PAIR_SPHERE_CAPSULE: 1, PAIR_CAPSULE_CAPSULE: 1}

_MJ_MINVAL = 1e-15
_DEFAULT_SOLREF = (0.02, 1.0)
_DEFAULT_SOLIMP = (0.9, 0.95, 0.001, 0.5, 2.0)


# ---------------------------------------------------------------------------
# containers (brax/base.py:170-540)


@dataclasses.dataclass
class Transform:
  pos: np.ndarray  # (..., 3)
  rot: np.ndarray  # (..., 4) quaternion (w, x, y, z)


@dataclasses.dataclass
class Motion:
  ang: np.ndarray  # (..., 3)
  vel: np.ndarray  # (..., 3)


@dataclasses.dataclass
class Inertia:
  transform: Transform
  i: np.ndarray  # (L, 3, 3)
  mass: np.ndarray  # (L,)


@dataclasses.dataclass
class Link:
  transform: Transform  # link frame relative to the parent link frame
  joint: Transform  # joint frame relative to the link frame
  inertia: Inertia
  invweight: np.ndarray  # (L,)


@dataclasses.dataclass
class DoF:
  motion: Motion  # (nv, 3) each
  armature: np.ndarray
  stiffness: np.ndarray
  damping: np.ndarray
  limit: Optional[Tuple[np.ndarray, np.ndarray]]
  invweight: np.ndarray
  solver_params: np.ndarray  # (nv, 7) = solref(2) + solimp(5)


@dataclasses.dataclass
class Actuator:
  q_id: np.ndarray
  qd_id: np.ndarray
  ctrl_range: np.ndarray  # (nu, 2)
  force_range: np.ndarray  # (nu, 2)
  gain: np.ndarray
  gear: np.ndarray
  bias_q: np.ndarray
  bias_qd: np.ndarray


@dataclasses.dataclass
class Opt:
  timestep: float
  iterations: int


@dataclasses.dataclass
class ContactPair:
  """One collision pair after MJX's filtering and parameter mixing."""

  kind: int  # PAIR_*
  geom1: int  # the geom of the lower type (the plane of a plane pair)
  geom2: int
  ncon: int
  friction: np.ndarray  # (5,)
  solref: np.ndarray  # (2,)
  solimp: np.ndarray  # (5,)


@dataclasses.dataclass
class System:
  """Model arrays read by the generalized step (brax/base.py:415-540)."""

  name: str
  gravity: np.ndarray
  link: Link
  dof: DoF
  actuator: Actuator
  init_q: np.ndarray
  opt: Opt
  link_names: List[str]
  link_types: str
  link_parents: Tuple[int, ...]
  matrix_inv_iterations: int
  solver_iterations: int
  solver_maxls: int
  # geoms (MuJoCo model arrays; body id 0 is the world, link i is body i+1)
  geom_bodyid: np.ndarray
  geom_type: np.ndarray
  geom_pos: np.ndarray
  geom_quat: np.ndarray
  geom_size: np.ndarray
  geom_friction: np.ndarray
  geom_solref: np.ndarray
  geom_solimp: np.ndarray
  geom_solmix: np.ndarray
  geom_contype: np.ndarray
  geom_conaffinity: np.ndarray
  elasticity: np.ndarray
  geom_names: List[str]
  enable_fluid: bool = False

  # -- index helpers (brax/base.py:482-540) --------------------------------

  def num_links(self) -> int:
    return len(self.link_types)

  def q_size(self) -> int:
    return sum(Q_WIDTHS[t] for t in self.link_types)

  def qd_size(self) -> int:
    return sum(QD_WIDTHS[t] for t in self.link_types)

  def act_size(self) -> int:
    return int(self.actuator.q_id.shape[0])

  def dof_link(self, depth: bool = False) -> np.ndarray:
    """Link index of each dof; with depth=True the index within its level."""
    idx = []
    for i, t in enumerate(self.link_types):
      idx.extend([i] * QD_WIDTHS[t])
    if depth:
      dpt = self.link_depths()
      count, within = {}, []
      for i in range(self.num_links()):
        within.append(count.get(dpt[i], 0))
        count[dpt[i]] = within[-1] + 1
      idx = [within[i] for i in idx]
    return np.array(idx, dtype=np.int64)

  def dof_ranges(self) -> List[List[int]]:
    beg, out = 0, []
    for t in self.link_types:
      out.append(list(range(beg, beg + QD_WIDTHS[t])))
      beg += QD_WIDTHS[t]
    return out

  def q_ranges(self) -> List[List[int]]:
    beg, out = 0, []
    for t in self.link_types:
      out.append(list(range(beg, beg + Q_WIDTHS[t])))
      beg += Q_WIDTHS[t]
    return out

  def q_idx(self, link_type: str) -> np.ndarray:
    idx, out = 0, []
    for t in self.link_types:
      if t in link_type:
        out.extend(range(idx, idx + Q_WIDTHS[t]))
      idx += Q_WIDTHS[t]
    return np.array(out, dtype=np.int64)

  def qd_idx(self, link_type: str) -> np.ndarray:
    idx, out = 0, []
    for t in self.link_types:
      if t in link_type:
        out.extend(range(idx, idx + QD_WIDTHS[t]))
      idx += QD_WIDTHS[t]
    return np.array(out, dtype=np.int64)

  def link_depths(self) -> List[int]:
    d = []
    for i, p in enumerate(self.link_parents):
      d.append(0 if p < 0 else d[p] + 1)
    return d

  def contact_pairs(self) -> List[ContactPair]:
    """Collision pairs with MJX filtering and mixing: geom1 has the lower
    geom type (the lower geom id between equal types), so the plane of a
    plane pair is geom1 and contact normals point from geom1 to geom2."""
    body_parent = [-1] + [p + 1 for p in self.link_parents]  # MuJoCo body ids
    pairs = []
    ng = len(self.geom_type)
    for g1, g2 in itertools.combinations(range(ng), 2):
      b1, b2 = int(self.geom_bodyid[g1]), int(self.geom_bodyid[g2])
      if b1 == b2:
        continue
      if ((self.geom_contype[g1] & self.geom_conaffinity[g2]) |
          (self.geom_contype[g2] & self.geom_conaffinity[g1])) == 0:
        continue
      # parent/child filter, except when the parent is the world body
      if (body_parent[b2] == b1 and b1 != 0) or (body_parent[b1] == b2 and b2 != 0):
        continue
      t1, t2 = int(self.geom_type[g1]), int(self.geom_type[g2])
      if t1 > t2:
        g1, g2, t1, t2 = g2, g1, t2, t1
      kind = _PAIR_KIND.get((t1, t2))
      if kind is None:
        raise NotImplementedError(
            f"collision pair types ({t1}, {t2}) are not supported by this step")
      friction = np.maximum(self.geom_friction[g1], self.geom_friction[g2])
      friction5 = np.array([friction[0], friction[0], friction[1],
                            friction[2], friction[2]])
      s1, s2 = float(self.geom_solmix[g1]), float(self.geom_solmix[g2])
      mix = 0.5 if (s1 + s2) <= 0 else s1 / (s1 + s2)
      solref = mix * self.geom_solref[g1] + (1 - mix) * self.geom_solref[g2]
      solimp = mix * self.geom_solimp[g1] + (1 - mix) * self.geom_solimp[g2]
      pairs.append(ContactPair(kind, g1, g2, _PAIR_NCON[kind], friction5,
                               solref, solimp))
    # MJX runs one collision function per pair kind, in kind order
    pairs.sort(key=lambda p: p.kind)
    return pairs

  def num_contacts(self) -> int:
    return sum(p.ncon for p in self.contact_pairs())

  def num_limit_rows(self) -> int:
    if self.dof.limit is None:
      return 0
    return int(len(self.qd_idx("123")))

  def replace(self, **kw) -> "System":
    return dataclasses.replace(self, **kw)


# ---------------------------------------------------------------------------
# small quaternion helpers (host, float64)


def quat_mul(u, v):
  return np.array([
      u[0] * v[0] - u[1] * v[1] - u[2] * v[2] - u[3] * v[3],
      u[0] * v[1] + u[1] * v[0] + u[2] * v[3] - u[3] * v[2],
      u[0] * v[2] - u[1] * v[3] + u[2] * v[0] + u[3] * v[1],
      u[0] * v[3] + u[1] * v[2] - u[2] * v[1] + u[3] * v[0],
  ])


def quat_to_mat(q):
  q = np.asarray(q, dtype=np.float64)
  q = q / np.linalg.norm(q)
  w, x, y, z = q
  return np.array([
      [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
  ])


def mat_to_quat(m):
  """Rotation matrix -> unit quaternion (w >= 0)."""
  t = np.trace(m)
  if t > 0:
    s = np.sqrt(t + 1.0) * 2
    q = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s,
         (m[1, 0] - m[0, 1]) / s]
  elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
    s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
    q = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s,
         (m[0, 2] + m[2, 0]) / s]
  elif m[1, 1] > m[2, 2]:
    s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
    q = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s,
         (m[1, 2] + m[2, 1]) / s]
  else:
    s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
    q = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s,
         (m[1, 2] + m[2, 1]) / s, 0.25 * s]
  q = np.array(q)
  q = q / np.linalg.norm(q)
  return q if q[0] >= 0 else -q


def axis_angle_quat(axis, angle):
  axis = np.asarray(axis, dtype=np.float64)
  axis = axis / np.linalg.norm(axis)
  return np.concatenate([[np.cos(angle / 2)], axis * np.sin(angle / 2)])


def z_to_vec_quat(vec):
  """Quaternion rotating +z onto the direction of ``vec`` (capsule fromto)."""
  v = np.asarray(vec, dtype=np.float64)
  v = v / np.linalg.norm(v)
  axis = np.cross([0.0, 0.0, 1.0], v)
  s = np.linalg.norm(axis)
  if s < 1e-10:
    return np.array([1.0, 0, 0, 0]) if v[2] > 0 else np.array([0.0, 1, 0, 0])
  ang = np.arctan2(s, v[2])
  return np.concatenate([[np.cos(ang / 2)], axis / s * np.sin(ang / 2)])


# ---------------------------------------------------------------------------
# geom mass properties


def _geom_volume(gtype, size):
  if gtype == GEOM_SPHERE:
    return 4.0 / 3.0 * np.pi * size[0] ** 3
  if gtype == GEOM_CAPSULE:
    r, h = size[0], size[1]
    return np.pi * r * r * 2 * h + 4.0 / 3.0 * np.pi * r ** 3
  return 0.0


def _geom_inertia(gtype, size, mass):
  """Principal inertia (geom frame) of a solid geom of the given mass."""
  if mass <= 0:
    return np.zeros(3)
  if gtype == GEOM_SPHERE:
    i = 0.4 * mass * size[0] ** 2
    return np.array([i, i, i])
  if gtype == GEOM_CAPSULE:
    r, h = size[0], size[1]
    height = 2 * h
    vol_cyl = np.pi * r * r * height
    vol_sph = 4.0 / 3.0 * np.pi * r ** 3
    rho = mass / (vol_cyl + vol_sph)
    mc, ms = rho * vol_cyl, rho * vol_sph
    ixx = (mc * (3 * r * r + height * height) / 12 +
           ms * (0.4 * r * r + height * height / 4 + 3 * r * height / 8))
    izz = mc * r * r / 2 + ms * 0.4 * r * r
    return np.array([ixx, ixx, izz])
  return np.zeros(3)


# ---------------------------------------------------------------------------
# model builder (subset compiler)


@dataclasses.dataclass
class _Joint:
  typ: str  # 'free' or 'hinge'
  axis: np.ndarray
  pos: np.ndarray
  limited: bool
  range: Tuple[float, float]
  armature: float
  damping: float
  stiffness: float
  solref: Tuple[float, float]
  solimp: Tuple[float, ...]
  name: str


@dataclasses.dataclass
class _Geom:
  body: int  # MuJoCo body id (0 = world)
  typ: int
  size: np.ndarray
  pos: np.ndarray
  quat: np.ndarray
  mass: float
  contype: int
  conaffinity: int
  friction: np.ndarray
  solref: np.ndarray
  solimp: np.ndarray
  solmix: float
  name: str
  elasticity: float = 0.0


@dataclasses.dataclass
class _Body:
  name: str
  parent: int  # MuJoCo body id of the parent (0 = world)
  pos: np.ndarray
  quat: np.ndarray
  joints: List[_Joint]


class ModelBuilder:
  """Compiles a body tree into a ``System`` (MJCF-subset compiler)."""

  def __init__(self, name: str, timestep: float, iterations: int,
               solver_maxls: int = 20, matrix_inv_iterations: int = 10,
               gravity=(0.0, 0.0, -9.81)):
    self.name = name
    self.timestep = float(timestep)
    self.iterations = int(iterations)
    self.solver_maxls = int(solver_maxls)
    self.matrix_inv_iterations = int(matrix_inv_iterations)
    self.gravity = np.array(gravity, dtype=np.float64)
    self.bodies: List[_Body] = []
    self.geoms: List[_Geom] = []
    self.motors = []  # (joint name, gear, ctrlrange or None)
    self.init_q_override = None

  # body ids: 0 is the world; add_body returns the new body id
  def add_body(self, name: str, parent: int, pos=(0, 0, 0),
               quat=(1, 0, 0, 0)) -> int:
    self.bodies.append(_Body(name, parent, np.array(pos, np.float64),
                             np.array(quat, np.float64) /
                             np.linalg.norm(quat), []))
    return len(self.bodies)

  def add_free_joint(self, body: int, name: str = "root"):
    self.bodies[body - 1].joints.append(
        _Joint("free", np.zeros(3), np.zeros(3), False, (0.0, 0.0), 0.0, 0.0,
               0.0, _DEFAULT_SOLREF, _DEFAULT_SOLIMP, name))

  def add_hinge(self, body: int, name: str, axis, pos=(0, 0, 0),
                range_rad: Optional[Tuple[float, float]] = None,
                armature: float = 0.0, damping: float = 0.0,
                stiffness: float = 0.0, solref=_DEFAULT_SOLREF,
                solimp=_DEFAULT_SOLIMP):
    axis = np.array(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    limited = range_rad is not None
    rng = tuple(range_rad) if limited else (0.0, 0.0)
    self.bodies[body - 1].joints.append(
        _Joint("hinge", axis, np.array(pos, np.float64), limited, rng,
               armature, damping, stiffness, tuple(solref), tuple(solimp),
               name))

  def add_geom(self, body: int, typ: int, size, pos=(0, 0, 0),
               quat=(1, 0, 0, 0), fromto=None, density: float = 1000.0,
               mass: Optional[float] = None, contype: int = 1,
               conaffinity: int = 1, friction=(1.0, 0.005, 0.0001),
               solref=_DEFAULT_SOLREF, solimp=_DEFAULT_SOLIMP,
               solmix: float = 1.0, name: str = "", elasticity: float = 0.0):
    size = np.array(list(size) + [0.0] * (3 - len(size)), np.float64)
    pos = np.array(pos, np.float64)
    quat = np.array(quat, np.float64) / np.linalg.norm(quat)
    if fromto is not None:
      a, b = np.array(fromto[:3], np.float64), np.array(fromto[3:], np.float64)
      pos = 0.5 * (a + b)
      quat = z_to_vec_quat(b - a)
      size[1] = 0.5 * np.linalg.norm(b - a)
    if mass is None:
      mass = density * _geom_volume(typ, size)
    self.geoms.append(_Geom(body, typ, size, pos, quat, float(mass), contype,
                            conaffinity, np.array(friction, np.float64),
                            np.array(solref, np.float64),
                            np.array(solimp, np.float64), solmix, name,
                            float(elasticity)))

  def add_motor(self, joint: str, gear: float, ctrlrange=None):
    self.add_actuator(joint, gear, ctrlrange=ctrlrange)

  def add_actuator(self, joint: str, gear: float, gain: float = 1.0,
                   bias_q: float = 0.0, bias_qd: float = 0.0, ctrlrange=None,
                   forcerange=None):
    """Joint actuator: force = clip(gain*ctrl + gear*(bias_q*q +
    bias_qd*qd), forcerange) * gear (MuJoCo motor: gain 1, no bias; position
    servo with kp: gain kp, bias_q -kp)."""
    self.motors.append((joint, float(gear), ctrlrange, float(gain),
                        float(bias_q), float(bias_qd), forcerange))

  # -- compilation ----------------------------------------------------------

  def _body_inertia(self, b: int):
    """Combined mass properties of body b (MuJoCo id) in the body frame."""
    gs = [g for g in self.geoms if g.body == b and g.mass > 0]
    if not gs:
      return 0.0, np.zeros(3), np.array([1.0, 0, 0, 0]), np.zeros(3)
    m = sum(g.mass for g in gs)
    com = sum(g.mass * g.pos for g in gs) / m
    if len(gs) == 1:
      g = gs[0]
      return m, com, g.quat.copy(), _geom_inertia(g.typ, g.size, g.mass)
    inert = np.zeros((3, 3))
    for g in gs:
      r = quat_to_mat(g.quat)
      ig = r @ np.diag(_geom_inertia(g.typ, g.size, g.mass)) @ r.T
      d = g.pos - com
      inert += ig + g.mass * (np.dot(d, d) * np.eye(3) - np.outer(d, d))
    evals, evecs = np.linalg.eigh(inert)
    if np.linalg.det(evecs) < 0:
      evecs[:, 0] = -evecs[:, 0]
    return m, com, mat_to_quat(evecs), evals

  def build(self) -> System:
    nb = len(self.bodies)
    link_types, link_parents, names = "", [], []
    for bi, b in enumerate(self.bodies):
      kinds = [j.typ for j in b.joints]
      if kinds == ["free"]:
        link_types += "f"
      elif kinds and all(k == "hinge" for k in kinds) and len(kinds) <= 3:
        if len({tuple(j.pos) for j in b.joints}) != 1:
          raise RuntimeError("invalid joint stack: one joint position allowed")
        link_types += str(len(kinds))
      else:
        raise NotImplementedError(
            f"body {b.name}: joints {kinds} are outside the supported subset "
            "(bodies without joints must be fused into their parent)")
      if b.parent - 1 >= bi:
        raise ValueError("bodies must be added parent-first")
      link_parents.append(b.parent - 1)
      names.append(b.name)

    # links (brax/io/mjcf.py:321-345)
    tpos = np.array([b.pos for b in self.bodies])
    trot = np.array([b.quat for b in self.bodies])
    jpos = np.array([b.joints[0].pos for b in self.bodies])
    for i, t in enumerate(link_types):
      if t == "f":  # free q is in world frame (brax/io/mjcf.py:466-470)
        tpos[i] = 0.0
        trot[i] = (1.0, 0.0, 0.0, 0.0)
    mass, ipos, irot, idiag = [], [], [], []
    for bi in range(nb):
      m, c, qi, d = self._body_inertia(bi + 1)
      mass.append(m), ipos.append(c), irot.append(qi), idiag.append(d)
    inertia_i = np.array([np.diag(d) for d in idiag])

    # dofs (brax/io/mjcf.py:347-392)
    ang, vel, arm, stiff, damp, lo, hi, sp = [], [], [], [], [], [], [], []
    joint_dof = {}
    any_limited = False
    for b in self.bodies:
      for j in b.joints:
        if j.typ == "free":
          joint_dof[j.name] = len(ang)
          ang += [np.zeros(3)] * 3 + list(np.eye(3))
          vel += list(np.eye(3)) + [np.zeros(3)] * 3
          arm += [j.armature] * 6
          damp += [j.damping] * 6
          stiff += [0.0] * 6
          lo += [-np.inf] * 6
          hi += [np.inf] * 6
          sp += [list(j.solref) + list(j.solimp)] * 6
        else:
          joint_dof[j.name] = len(ang)
          ang.append(j.axis.copy())
          vel.append(np.zeros(3))
          arm.append(j.armature)
          damp.append(j.damping)
          stiff.append(j.stiffness)
          if j.limited:
            any_limited = True
            lo.append(j.range[0])
            hi.append(j.range[1])
          else:
            lo.append(-np.inf)
            hi.append(np.inf)
          sp.append(list(j.solref) + list(j.solimp))
    nv = len(ang)
    limit = (np.array(lo), np.array(hi)) if any_limited else None

    # actuators (brax/io/mjcf.py:394-423); motors only: gain 1, no bias
    q_start = {}
    qi = 0
    for b in self.bodies:
      for j in b.joints:
        q_start[j.name] = qi
        qi += 7 if j.typ == "free" else 1
    q_id, qd_id, ctrl, gear, gain, bq, bqd, frc = [], [], [], [], [], [], [], []
    for jn, g, cr, gn, b1, b2, fr in self.motors:
      q_id.append(q_start[jn])
      qd_id.append(joint_dof[jn])
      gear.append(g)
      ctrl.append(cr if cr is not None else (-np.inf, np.inf))
      gain.append(gn)
      bq.append(b1)
      bqd.append(b2)
      frc.append(fr if fr is not None else (-np.inf, np.inf))
    nu = len(q_id)
    actuator = Actuator(
        q_id=np.array(q_id, np.int64), qd_id=np.array(qd_id, np.int64),
        ctrl_range=np.array(ctrl, np.float64).reshape(nu, 2),
        force_range=np.array(frc, np.float64).reshape(nu, 2),
        gain=np.array(gain, np.float64), gear=np.array(gear, np.float64),
        bias_q=np.array(bq, np.float64), bias_qd=np.array(bqd, np.float64))

    # qpos0: free joints at the body pose, hinges at zero
    init_q = []
    for b in self.bodies:
      if b.joints[0].typ == "free":
        init_q += list(b.pos) + list(b.quat)
      else:
        init_q += [0.0] * len(b.joints)
    qpos0 = np.array(init_q, np.float64)
    init_q = qpos0.copy()
    if self.init_q_override is not None:
      init_q = np.array(self.init_q_override, np.float64)

    # MuJoCo numbers geoms body by body (the world's first), in declaration
    # order within a body; pair order and roles follow these ids
    g = sorted(self.geoms, key=lambda x: x.body)
    sys = System(
        name=self.name,
        gravity=self.gravity.copy(),
        link=Link(
            transform=Transform(tpos, trot),
            joint=Transform(jpos, np.tile([1.0, 0, 0, 0], (nb, 1))),
            inertia=Inertia(Transform(np.array(ipos), np.array(irot)),
                            inertia_i, np.array(mass)),
            invweight=np.zeros(nb)),
        dof=DoF(motion=Motion(np.array(ang).reshape(nv, 3),
                              np.array(vel).reshape(nv, 3)),
                armature=np.array(arm), stiffness=np.array(stiff),
                damping=np.array(damp), limit=limit,
                invweight=np.zeros(nv),
                solver_params=np.array(sp, np.float64).reshape(nv, 7)),
        actuator=actuator,
        init_q=init_q,
        opt=Opt(self.timestep, self.iterations),
        link_names=names,
        link_types=link_types,
        link_parents=tuple(link_parents),
        matrix_inv_iterations=self.matrix_inv_iterations,
        solver_iterations=self.iterations,
        solver_maxls=self.solver_maxls,
        geom_bodyid=np.array([x.body for x in g], np.int64),
        geom_type=np.array([x.typ for x in g], np.int64),
        geom_pos=np.array([x.pos for x in g]).reshape(-1, 3),
        geom_quat=np.array([x.quat for x in g]).reshape(-1, 4),
        geom_size=np.array([x.size for x in g]).reshape(-1, 3),
        geom_friction=np.array([x.friction for x in g]).reshape(-1, 3),
        geom_solref=np.array([x.solref for x in g]).reshape(-1, 2),
        geom_solimp=np.array([x.solimp for x in g]).reshape(-1, 5),
        geom_solmix=np.array([x.solmix for x in g], np.float64),
        geom_contype=np.array([x.contype for x in g], np.int64),
        geom_conaffinity=np.array([x.conaffinity for x in g], np.int64),
        elasticity=np.array([x.elasticity for x in g], np.float64),
        geom_names=[x.name for x in g],
    )
    _set_invweights(sys, qpos0)
    return sys


# ---------------------------------------------------------------------------
# independent rigid-body kinematics (float64), used for invweight0 and as a
# brute-force mass-matrix check


def body_frames(sys: System, q: np.ndarray):
  """World pose of every link and the world axis/anchor of every dof.

  Joints compose in the link's local frame: free joints place the link at
  q[0:3], q[3:7]; hinge k of a stacked joint rotates about its axis after the
  rotations of hinges 0..k-1 (the MuJoCo convention).
  """
  L = sys.num_links()
  pos = np.zeros((L, 3))
  rot = np.zeros((L, 4))
  dof_axis, dof_anchor, dof_lin = [], [], []
  qr = sys.q_ranges()
  for i, t in enumerate(sys.link_types):
    p = sys.link_parents[i]
    ppos = pos[p] if p >= 0 else np.zeros(3)
    prot = rot[p] if p >= 0 else np.array([1.0, 0, 0, 0])
    if t == "f":
      qq = q[qr[i]]
      pos[i] = qq[0:3]
      rot[i] = qq[3:7] / np.linalg.norm(qq[3:7])
      rm = quat_to_mat(rot[i])
      for k in range(3):
        dof_axis.append(np.eye(3)[k])
        dof_anchor.append(pos[i].copy())
        dof_lin.append(True)
      for k in range(3):
        dof_axis.append(rm[:, k])
        dof_anchor.append(pos[i].copy())
        dof_lin.append(False)
      continue
    bpos = ppos + quat_to_mat(prot) @ sys.link.transform.pos[i]
    brot = quat_mul(prot, sys.link.transform.rot[i])
    anchor_local = sys.link.joint.pos[i]
    anchor = bpos + quat_to_mat(brot) @ anchor_local
    cur = brot
    dr = sys.dof_ranges()[i]
    for k, d in enumerate(dr):
      axis_world = quat_to_mat(cur) @ sys.dof.motion.ang[d]
      dof_axis.append(axis_world)
      dof_anchor.append(anchor.copy())
      dof_lin.append(False)
      cur = quat_mul(cur, axis_angle_quat(sys.dof.motion.ang[d], q[qr[i][k]]))
    rot[i] = cur / np.linalg.norm(cur)
    # the joint rotates about the anchor: frame origin moves with it
    pos[i] = anchor + quat_to_mat(rot[i]) @ (-anchor_local)
  return pos, rot, np.array(dof_axis), np.array(dof_anchor), np.array(dof_lin)


def _ancestors(sys: System, i: int):
  out = []
  while i >= 0:
    out.append(i)
    i = sys.link_parents[i]
  return out


def point_jacobians(sys: System, q: np.ndarray, link: int, point: np.ndarray):
  """Linear (3, nv) and angular (3, nv) Jacobians of a world point on a link."""
  _, _, axis, anchor, lin = body_frames(sys, q)
  nv = sys.qd_size()
  jl, ja = np.zeros((3, nv)), np.zeros((3, nv))
  dl = sys.dof_link()
  anc = set(_ancestors(sys, link))
  for d in range(nv):
    if dl[d] not in anc:
      continue
    if lin[d]:
      jl[:, d] = axis[d]
    else:
      ja[:, d] = axis[d]
      jl[:, d] = np.cross(axis[d], point - anchor[d])
  return jl, ja


def mass_matrix_bruteforce(sys: System, q: np.ndarray) -> np.ndarray:
  """M = sum_links m J_v^T J_v + J_w^T I_world J_w + diag(armature)."""
  pos, rot, _, _, _ = body_frames(sys, q)
  m = np.diag(sys.dof.armature.astype(np.float64))
  for i in range(sys.num_links()):
    mass = sys.link.inertia.mass[i]
    if mass <= 0:
      continue
    r = quat_to_mat(rot[i])
    com = pos[i] + r @ sys.link.inertia.transform.pos[i]
    ri = r @ quat_to_mat(sys.link.inertia.transform.rot[i])
    iw = ri @ sys.link.inertia.i[i] @ ri.T
    jl, ja = point_jacobians(sys, q, i, com)
    m += mass * jl.T @ jl + ja.T @ iw @ ja
  return m


def _set_invweights(sys: System, q0=None):
  """body_invweight0 and dof_invweight0 at qpos0 (MuJoCo set0 semantics)."""
  q = sys.init_q if q0 is None else q0
  if not np.any(sys.link.inertia.mass > 0):
    return
  m = mass_matrix_bruteforce(sys, q)
  minv = np.linalg.inv(m)
  pos, rot, _, _, _ = body_frames(sys, q)
  for i in range(sys.num_links()):
    com = pos[i] + quat_to_mat(rot[i]) @ sys.link.inertia.transform.pos[i]
    jl, _ = point_jacobians(sys, q, i, com)
    a = jl @ minv @ jl.T
    sys.link.invweight[i] = max(_MJ_MINVAL, np.trace(a) / 3.0)
  diag = np.diag(minv)
  dv = np.zeros_like(diag)
  for i, t in enumerate(sys.link_types):
    dr = sys.dof_ranges()[i]
    if t == "f":
      dv[dr[0:3]] = diag[dr[0:3]].mean()
      dv[dr[3:6]] = diag[dr[3:6]].mean()
    else:
      dv[dr] = diag[dr]
  sys.dof.invweight[:] = dv


# ---------------------------------------------------------------------------
# builders


def _quadruped_builder(seed: int, knee_pitch_deg: float):
  """Synthetic quadruped: free torso sphere + 4 legs of 2 hinged capsules."""
  rng = np.random.default_rng(seed)
  lengths = rng.uniform(0.2, 0.4, size=(4, 2))
  limits = np.deg2rad(rng.uniform(30.0, 70.0, size=(4, 2)))
  beta = np.deg2rad(knee_pitch_deg)
  cap_r = 0.08
  ground_z = float(np.max(lengths[:, 1] * np.sin(beta)) + cap_r - 0.005)

  b = ModelBuilder("quadruped", timestep=0.01, iterations=4, solver_maxls=15,
                   matrix_inv_iterations=10)
  b.add_geom(0, GEOM_PLANE, (40, 40, 40), contype=0, conaffinity=1,
             friction=(1.0, 0.005, 0.0001), name="floor")
  torso = b.add_body("torso", 0, pos=(0, 0, ground_z))
  b.add_free_joint(torso)
  b.add_geom(torso, GEOM_SPHERE, (0.25,), mass=1.0, contype=1, conaffinity=0,
             friction=(1.0, 0.5, 0.5), name="torso_geom")
  for k in range(4):
    phi = np.deg2rad(45.0 + 90.0 * k)
    d = np.array([np.cos(phi), np.sin(phi), 0.0])
    a_knee = np.cross([0.0, 0.0, 1.0], d)
    up = b.add_body(f"upper_{k}", torso, pos=0.25 * d)
    b.add_hinge(up, f"hip_{k}", (0, 0, 1),
                range_rad=(-limits[k, 0], limits[k, 0]), armature=1.0,
                damping=1.0)
    b.add_geom(up, GEOM_CAPSULE, (cap_r,),
               fromto=np.concatenate([np.zeros(3), lengths[k, 0] * d]),
               density=1000.0, contype=1, conaffinity=0,
               friction=(1.0, 0.5, 0.5), name=f"upper_geom_{k}")
    lo = b.add_body(f"lower_{k}", up, pos=lengths[k, 0] * d)
    b.add_hinge(lo, f"knee_{k}", a_knee,
                range_rad=(-limits[k, 1], limits[k, 1]), armature=1.0,
                damping=1.0)
    tip = lengths[k, 1] * (np.cos(beta) * d - np.sin(beta) * np.array([0, 0, 1.0]))
    b.add_geom(lo, GEOM_CAPSULE, (cap_r,),
               fromto=np.concatenate([np.zeros(3), tip]), density=1000.0,
               contype=1, conaffinity=0, friction=(1.0, 0.5, 0.5),
               name=f"lower_geom_{k}")
  for k in range(4):
    b.add_motor(f"hip_{k}", 150.0, (-1.0, 1.0))
    b.add_motor(f"knee_{k}", 150.0, (-1.0, 1.0))
  return b


def quadruped(seed: int = 0) -> System:
  """The synthetic quadruped of the 1024..65536-env workloads.

  Free-joint torso sphere (r=0.25 m, 1 kg); 4 legs at 45+90k degrees of 2
  hinge-jointed capsules each (length U(0.2, 0.4) m, radius 0.08 m, density
  1000); hips about +z, knees about z x leg; hinge limits +-U(30, 70) deg;
  motors with gear 150 on ctrl [-1, 1]; ground plane; only limb/plane pairs
  collide.  Options follow the ant task (brax/envs/assets/ant.xml:3,16,18):
  timestep 0.01, 4 solver iterations, 15 line-search steps, 10 Newton-Schulz
  iterations, hinge armature 1 and damping 1.  qd size 14.
  """
  return _quadruped_builder(seed, knee_pitch_deg=60.0).build()


def humanoid(seed: int = 0) -> System:
  """Synthetic humanoid-like tree (16384-env workload).

  13 capsule bodies: pelvis (free joint), torso (1-DoF waist), head, two
  thighs (3-DoF hips as hinge triples x, z, y), two shins (1-DoF knees), two
  feet (two heel-to-toe capsules each), two upper arms (3-DoF shoulders as
  hinge triples x, y, z) and two forearms (1-DoF elbows) each carrying a hand
  capsule.  The head and feet have no joint and are fused into their parents
  (brax/io/mjcf.py:76-98), leaving 10 links; qd size 23, q size 24.  Body
  masses U(1, 8) kg (split over a body's capsules by volume).  Contacts: the
  4 foot and 2 hand capsules and a torso sphere against the plane (13 MJX
  contacts); every hinge is limited (17 limit rows).  Options and joint
  parameters follow brax/envs/assets/humanoid.xml:2-9,23-24,46-108:
  timestep 0.003, 8 solver iterations, 15 line-search steps, 20 Newton-Schulz
  iterations, motors with ctrl range [-0.4, 0.4] and gears 100/300/200/25.
  """
  rng = np.random.default_rng(seed)
  masses = iter(rng.uniform(1.0, 8.0, size=13))
  b = ModelBuilder("humanoid", timestep=0.003, iterations=8, solver_maxls=15,
                   matrix_inv_iterations=20)
  b.add_geom(0, GEOM_PLANE, (20, 20, 0.125), contype=0, conaffinity=1,
             friction=(1.0, 0.1, 0.1), name="floor")
  touch = dict(contype=1, conaffinity=0)
  none = dict(contype=0, conaffinity=0)

  def capsules(body, specs, **kw):
    """Adds capsules (fromto, radius, contact kw) sharing one body mass."""
    m = next(masses)
    vols = []
    for ft, r, _ in specs:
      half = 0.5 * np.linalg.norm(np.subtract(ft[3:], ft[:3]))
      vols.append(_geom_volume(GEOM_CAPSULE, np.array([r, half, 0.0])))
    for (ft, r, ckw), v in zip(specs, vols):
      b.add_geom(body, GEOM_CAPSULE, (r,), fromto=ft, mass=m * v / sum(vols),
                 **ckw, **kw)

  def hinge(body, name, axis, deg, armature, damping, stiffness):
    b.add_hinge(body, name, axis, range_rad=np.deg2rad(deg),
                armature=armature, damping=damping, stiffness=stiffness)

  pelvis = b.add_body("pelvis", 0, pos=(0, 0, 0.885))
  b.add_free_joint(pelvis)
  capsules(pelvis, [((0, -0.07, 0, 0, 0.07, 0), 0.09, none)])
  torso = b.add_body("torso", pelvis, pos=(0, 0, 0.1))
  hinge(torso, "abdomen_y", (0, 1, 0), (-45, 30), 0.02, 5.0, 10.0)
  capsules(torso, [((0, 0, 0, 0, 0, 0.3), 0.07, none)])
  b.add_geom(torso, GEOM_SPHERE, (0.11,), pos=(0, 0, 0.3), mass=0.0,
             name="chest", **touch)
  head = b.add_body("head", torso, pos=(0, 0, 0.45))
  capsules(head, [((0, 0, 0, 0, 0, 0.1), 0.08, none)])
  for side, s in (("right", -1.0), ("left", 1.0)):
    thigh = b.add_body(f"{side}_thigh", pelvis, pos=(0, s * 0.1, -0.04))
    hinge(thigh, f"{side}_hip_x", (s, 0, 0), (-25, 5), 0.01, 5.0, 10.0)
    hinge(thigh, f"{side}_hip_z", (0, 0, s), (-60, 35), 0.01, 5.0, 10.0)
    hinge(thigh, f"{side}_hip_y", (0, 1, 0), (-110, 20), 0.008, 5.0, 20.0)
    capsules(thigh, [((0, 0, 0, 0, 0, -0.34), 0.06, none)])
    shin = b.add_body(f"{side}_shin", thigh, pos=(0, 0, -0.4))
    hinge(shin, f"{side}_knee", (0, -1, 0), (-150, 0), 0.006, 1.0, 0.0)
    capsules(shin, [((0, 0, 0, 0, 0, -0.3), 0.05, none)])
    foot = b.add_body(f"{side}_foot", shin, pos=(0, 0, -0.4))
    capsules(foot, [((-0.05, -0.03, 0, 0.12, -0.03, 0), 0.03, touch),
                    ((-0.05, 0.03, 0, 0.12, 0.03, 0), 0.03, touch)])
  for side, s in (("right", -1.0), ("left", 1.0)):
    uarm = b.add_body(f"{side}_upper_arm", torso, pos=(0, s * 0.17, 0.25))
    hinge(uarm, f"{side}_shoulder_x", (s, 0, 0), (-85, 60), 0.0068, 1.0, 1.0)
    hinge(uarm, f"{side}_shoulder_y", (0, 1, 0), (-85, 60), 0.0051, 1.0, 1.0)
    hinge(uarm, f"{side}_shoulder_z", (0, 0, s), (-60, 60), 0.005, 1.0, 1.0)
    capsules(uarm, [((0, 0, 0, 0, 0, -0.26), 0.04, none)])
    farm = b.add_body(f"{side}_forearm", uarm, pos=(0, 0, -0.28))
    hinge(farm, f"{side}_elbow", (0, 1, 0), (-90, 0), 0.0028, 1.0, 0.0)
    capsules(farm, [((0, 0, 0, 0, 0, -0.22), 0.035, none),
                    ((0, 0, -0.25, 0, 0, -0.33), 0.03, touch)])
  b.bodies = _fuse_jointless(b)
  gears = {"abdomen_y": 100.0}
  for side in ("right", "left"):
    gears.update({f"{side}_hip_x": 100.0, f"{side}_hip_z": 100.0,
                  f"{side}_hip_y": 300.0, f"{side}_knee": 200.0,
                  f"{side}_shoulder_x": 25.0, f"{side}_shoulder_y": 25.0,
                  f"{side}_shoulder_z": 25.0, f"{side}_elbow": 25.0})
  for j, g in gears.items():
    b.add_motor(j, g, (-0.4, 0.4))
  return b.build()


def _fuse_jointless(b: "ModelBuilder") -> List[_Body]:
  """Fuses bodies without joints into their parents (brax/io/mjcf.py:76-98).

  Geoms of a fused body move to the parent with the body's offset applied
  (body quaternions of fused bodies must be identity); children are re-parented
  and body ids are renumbered.  Returns the new body list and rewrites geom
  body ids in place.
  """
  keep, newid = [], {0: 0}
  offset = {0: np.zeros(3)}
  target = {0: 0}
  for i, body in enumerate(b.bodies, start=1):
    p = body.parent
    if body.joints:
      keep.append(body)
      newid[i] = len(keep)
      target[i] = i
      offset[i] = np.zeros(3)
    else:
      if not np.allclose(body.quat, (1, 0, 0, 0)):
        raise NotImplementedError("fused bodies must have identity rotation")
      target[i] = target[p]
      offset[i] = offset[p] + body.pos
  for body in keep:
    body.pos = body.pos + offset[body.parent]
    body.parent = newid[target[body.parent]]
  for g in b.geoms:
    if g.body:
      g.pos = g.pos + offset[g.body]
      g.body = newid[target[g.body]]
  return keep


def ant() -> System:
  """The ant of brax/envs/assets/ant.xml after body fusing (mjcf.py:76-98).

  Numbers transcribed from ant.xml:30-75 (density 5, friction 1 0.5 0.5,
  hinge armature 1 / damping 1, 8 motors with gear 150).  Used for the
  reference's scan and contact known-answer tests.
  """
  b = ModelBuilder("ant", timestep=0.01, iterations=4, solver_maxls=15,
                   matrix_inv_iterations=10)
  b.add_geom(0, GEOM_PLANE, (40, 40, 40), contype=0, conaffinity=1,
             name="floor")
  kw = dict(density=5.0, contype=0, conaffinity=0, friction=(1.0, 0.5, 0.5))
  torso = b.add_body("torso", 0, pos=(0, 0, 0.75))
  b.add_free_joint(torso)
  b.add_geom(torso, GEOM_SPHERE, (0.25,), name="torso_geom", **kw)
  legs = [  # (hip offset, ankle axis, ankle range deg, name)
      ((0.2, 0.2), (-1, 1, 0), (30, 70), "1"),
      ((-0.2, 0.2), (1, 1, 0), (-70, -30), "2"),
      ((-0.2, -0.2), (-1, 1, 0), (-70, -30), "3"),
      ((0.2, -0.2), (1, 1, 0), (30, 70), "4"),
  ]
  # fused aux geoms first (they belong to the torso after fusing)
  for (hx, hy), _, _, n in legs:
    b.add_geom(torso, GEOM_CAPSULE, (0.08,),
               fromto=(0, 0, 0, hx, hy, 0), name=f"aux_{n}_geom", **kw)
  for (hx, hy), axis, rng, n in legs:
    aux = b.add_body(f"aux_{n}", torso, pos=(hx, hy, 0))
    b.add_hinge(aux, f"hip_{n}", (0, 0, 1),
                range_rad=np.deg2rad((-30.0, 30.0)), armature=1.0,
                damping=1.0)
    b.add_geom(aux, GEOM_CAPSULE, (0.08,), fromto=(0, 0, 0, hx, hy, 0),
               name=f"leg_geom_{n}", **kw)
    low = b.add_body("", aux, pos=(hx, hy, 0))
    b.add_hinge(low, f"ankle_{n}", axis, range_rad=np.deg2rad(rng),
                armature=1.0, damping=1.0)
    b.add_geom(low, GEOM_CAPSULE, (0.08,),
               fromto=(0, 0, 0, 2 * hx, 2 * hy, 0), name=f"ankle_geom_{n}",
               **kw)
    b.add_geom(low, GEOM_SPHERE, (0.08,), pos=(2 * hx, 2 * hy, 0), mass=0.0,
               contype=1, conaffinity=0, friction=(1.0, 0.5, 0.5),
               name=f"foot_geom_{n}")
  for n in ["4", "1", "2", "3"]:
    b.add_motor(f"hip_{n}", 150.0, (-1.0, 1.0))
    b.add_motor(f"ankle_{n}", 150.0, (-1.0, 1.0))
  # ant.xml:6 custom init_qpos
  b.init_q_override = [0.0, 0.0, 0.55, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0,
                       -1.0, 0.0, -1.0, 0.0, 1.0]
  return b.build()


def humanoid_xml() -> System:
  """The reference's humanoid, brax/envs/assets/humanoid.xml, transcribed.

  The same compiler calls load_mjcf makes for that file (bodies, joints,
  geoms and motors in document order, its defaults: joint armature 1,
  damping 1, limited; geoms contype 0, conaffinity 0; motor ctrlrange
  +-0.4; options timestep 0.003, 8 iterations, matrix_inv_iterations 20,
  solver_maxls 15), so that the GPU tests, which cannot read the reference
  tree, step the reference's own model; tests/test_mjcf.py checks every
  array against load_mjcf of the file.  After fusing the jointless feet: 11
  links, 23 dofs, 17 motors, 17 limit rows and 2 foot-sphere contacts.
  """
  deg = np.pi / 180.0
  b = ModelBuilder("humanoid_xml", timestep=0.003, iterations=8, solver_maxls=15,
                   matrix_inv_iterations=20)
  g = dict(contype=0, conaffinity=0)
  b.add_geom(0, GEOM_PLANE, (20, 20, 0.125), contype=0, conaffinity=1, friction=(1.0, 0.1, 0.1),
             name="floor")

  def hinge(body, name, axis, rng, armature, damping=1.0, stiffness=0.0, pos=(0, 0, 0)):
    b.add_hinge(body, name, axis, pos=pos, range_rad=(rng[0] * deg, rng[1] * deg),
                armature=armature, damping=damping, stiffness=stiffness)

  torso = b.add_body("torso", 0, pos=(0, 0, 1.4))
  b.add_free_joint(torso, "root")
  b.add_geom(torso, GEOM_CAPSULE, (0.07,), fromto=(0, -.07, 0, 0, .07, 0), name="torso1", **g)
  b.add_geom(torso, GEOM_SPHERE, (.09,), pos=(0, 0, .19), name="head", **g)
  b.add_geom(torso, GEOM_CAPSULE, (0.06,), fromto=(-.01, -.06, -.12, -.01, .06, -.12),
             name="uwaist", **g)
  lwaist = b.add_body("lwaist", torso, pos=(-.01, 0, -0.260), quat=(1.000, 0, -0.002, 0))
  b.add_geom(lwaist, GEOM_CAPSULE, (0.06,), fromto=(0, -.06, 0, 0, .06, 0), name="lwaist", **g)
  hinge(lwaist, "abdomen_z", (0, 0, 1), (-45, 45), 0.02, 5, 20, pos=(0, 0, 0.065))
  hinge(lwaist, "abdomen_y", (0, 1, 0), (-75, 30), 0.02, 5, 10, pos=(0, 0, 0.065))
  pelvis = b.add_body("pelvis", lwaist, pos=(0, 0, -0.165), quat=(1.000, 0, -0.002, 0))
  hinge(pelvis, "abdomen_x", (1, 0, 0), (-35, 35), 0.02, 5, 10, pos=(0, 0, 0.1))
  b.add_geom(pelvis, GEOM_CAPSULE, (0.09,), fromto=(-.02, -.07, 0, -.02, .07, 0), name="butt",
             **g)
  for side, sy, hx, hz, hy_arm, knee_stiff in (("right", -1, (1, 0, 0), (0, 0, 1), 0.0080, 0.0),
                                               ("left", 1, (-1, 0, 0), (0, 0, -1), 0.01, 1.0)):
    thigh = b.add_body(f"{side}_thigh", pelvis, pos=(0, 0.1 * sy, -0.04))
    hinge(thigh, f"{side}_hip_x", hx, (-25, 5), 0.01, 5, 10)
    hinge(thigh, f"{side}_hip_z", hz, (-60, 35), 0.01, 5, 10)
    hinge(thigh, f"{side}_hip_y", (0, 1, 0), (-110, 20), hy_arm, 5, 20)
    b.add_geom(thigh, GEOM_CAPSULE, (0.06,), fromto=(0, 0, 0, 0, -0.01 * sy, -.34),
               name=f"{side}_thigh1", **g)
    shin = b.add_body(f"{side}_shin", thigh, pos=(0, -0.01 * sy, -0.403))
    hinge(shin, f"{side}_knee", (0, -1, 0), (-160, -2), 0.0060, stiffness=knee_stiff,
          pos=(0, 0, .02))
    b.add_geom(shin, GEOM_CAPSULE, (0.049,), fromto=(0, 0, 0, 0, 0, -.3), name=f"{side}_shin1",
               **g)
    foot = b.add_body(f"{side}_foot", shin, pos=(0, 0, -0.45))
    b.add_geom(foot, GEOM_SPHERE, (0.075,), pos=(0, 0, 0.1), contype=1, conaffinity=0,
               name=f"{side}_foot")
  for side, sy, s1, s2, rng, el in (("right", -1, (2, 1, 1), (0, -1, 1), (-85, 60), (0, -1, 1)),
                                    ("left", 1, (2, -1, 1), (0, 1, 1), (-60, 85), (0, -1, -1))):
    upper = b.add_body(f"{side}_upper_arm", torso, pos=(0, 0.17 * sy, 0.06))
    hinge(upper, f"{side}_shoulder1", s1, rng, 0.0068, stiffness=1)
    hinge(upper, f"{side}_shoulder2", s2, rng, 0.0051, stiffness=1)
    b.add_geom(upper, GEOM_CAPSULE, (0.04, 0.16), fromto=(0, 0, 0, .16, .16 * sy, -.16),
               name=f"{side}_uarm1", **g)
    lower = b.add_body(f"{side}_lower_arm", upper, pos=(.18, .18 * sy, -.18))
    hinge(lower, f"{side}_elbow", el, (-90, 50), 0.0028, stiffness=0)
    b.add_geom(lower, GEOM_CAPSULE, (0.031,), fromto=(0.01, -0.01 * sy, 0.01, .17, -.17 * sy, .17),
               name=f"{side}_larm", **g)
    b.add_geom(lower, GEOM_SPHERE, (0.04,), pos=(.18, -.18 * sy, .18), name=f"{side}_hand", **g)
  for joint, gear in (("abdomen_y", 100), ("abdomen_z", 100), ("abdomen_x", 100),
                      ("right_hip_x", 100), ("right_hip_z", 100), ("right_hip_y", 300),
                      ("right_knee", 200), ("left_hip_x", 100), ("left_hip_z", 100),
                      ("left_hip_y", 300), ("left_knee", 200), ("right_shoulder1", 25),
                      ("right_shoulder2", 25), ("right_elbow", 25), ("left_shoulder1", 25),
                      ("left_shoulder2", 25), ("left_elbow", 25)):
    b.add_motor(joint, float(gear), (-.4, .4))
  b.bodies = _fuse_jointless(b)
  return b.build()


def pendulum(n_links: int = 1, motor_gear: Optional[float] = None,
             ctrlrange=(-100.0, 100.0), timestep: float = 0.001,
             position_servo: Optional[Tuple[float, float, Tuple[float, float]]] = None
             ) -> System:
  """Chain of n hinged unit-mass spheres (brax/test_data/*pendulum*.xml).

  Each body hangs 0.5 m along +y from its hinge (axis +x); spheres r=0.15,
  mass 1 (inertia 0.009).  With n_links=3 and no motor this is
  triple_pendulum.xml; with n_links=1 and gear 150 it is
  single_pendulum_motor.xml.
  """
  b = ModelBuilder(f"pendulum{n_links}", timestep=timestep, iterations=100)
  parent = 0
  for k in range(n_links):
    body = b.add_body(f"body{k + 1}", parent,
                      pos=(0, 0.5, 0) if k > 0 else (0, 0, 0))
    b.add_hinge(body, f"hinge{k + 1}", (1, 0, 0))
    b.add_geom(body, GEOM_SPHERE, (0.15,), pos=(0, 0.5, 0), mass=1.0,
               contype=0, conaffinity=0, name=f"sphere{k + 1}")
    parent = body
  if motor_gear is not None:
    b.add_motor("hinge1", motor_gear, ctrlrange)
  if position_servo is not None:
    kp, gear, frc = position_servo
    b.add_actuator("hinge1", gear, gain=kp, bias_q=-kp, forcerange=frc)
  return b.build()


def chain(n_links: int, seed: int = 0) -> System:
  """Serial hinge chain of capsule links hanging from the world.

  The heterogeneous-batch workload's chains: link k is a capsule of length
  U(0.1, 0.5) m (radius 0.04) hanging along -z from its hinge; its mass is
  log-uniform on [0.1, 10] kg; hinge axes alternate y, x (so the chain moves
  in 3D); the first hinge is fixed to the world at z = 2; every link has a
  motor (gear 10, ctrl [-1, 1]); armature 0.1 and damping 0.5 per hinge;
  no contacts and no limits.  Options as the quadruped (timestep 0.01,
  4 solver iterations).  q size = qd size = n_links.
  """
  rng = np.random.default_rng(seed)
  lengths = rng.uniform(0.1, 0.5, size=n_links)
  masses = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=n_links))
  b = ModelBuilder(f"chain{n_links}", timestep=0.01, iterations=4, solver_maxls=15,
                   matrix_inv_iterations=10)
  parent = 0
  for k in range(n_links):
    pos = (0, 0, 2.0) if k == 0 else (0, 0, -lengths[k - 1])
    body = b.add_body(f"link{k}", parent, pos=pos)
    b.add_hinge(body, f"hinge{k}", (0, 1, 0) if k % 2 == 0 else (1, 0, 0),
                armature=0.1, damping=0.5)
    b.add_geom(body, GEOM_CAPSULE, (0.04,), fromto=(0, 0, 0, 0, 0, -lengths[k]),
               mass=masses[k], contype=0, conaffinity=0, name=f"capsule{k}")
    parent = body
  for k in range(n_links):
    b.add_motor(f"hinge{k}", 10.0, (-1.0, 1.0))
  return b.build()


def two_spheres() -> System:
  """Two free bodies whose spheres overlap by 0.01 at the initial pose: body1
  at the origin with a sphere of radius 0.2 (elasticity 0.1), body2 at the
  origin with a sphere of radius 0.11 at (0, 0.3, 0); geom elasticity
  default 0.2; MuJoCo defaults otherwise (density 1000, timestep 0.002,
  100 solver iterations).  The system of the reference's sphere-sphere
  contact test (brax/contact_test.py:45-75)."""
  b = ModelBuilder("two_spheres", timestep=0.002, iterations=100)
  b1 = b.add_body("body1", 0)
  b.add_free_joint(b1, "free1")
  b.add_geom(b1, GEOM_SPHERE, (0.2,), name="sphere1", elasticity=0.1)
  b2 = b.add_body("body2", 0)
  b.add_free_joint(b2, "free2")
  b.add_geom(b2, GEOM_SPHERE, (0.11,), pos=(0.0, 0.3, 0.0), name="sphere2", elasticity=0.2)
  return b.build()


PILE_BODIES = ("ball1", "ball2", "rod1", "rod2")


def pile() -> System:
  """Four free bodies over a floor, beside a capsule fixed to the world: two
  spheres (radii 0.1, 0.08) and two capsules (radius 0.05, half length 0.2
  along x; radius 0.04, half length 0.15 along y), density 500.  Every geom
  collides with every other (the world's two with the bodies'), so the
  contacts cover each pair kind the step handles: plane-sphere,
  plane-capsule, sphere-sphere, sphere-capsule and capsule-capsule, with a
  world geom on either side of a pair.  Friction 0.8 (tangential), timestep
  0.005, 6 solver iterations; the bodies start apart and settle onto each
  other and the floor."""
  b = ModelBuilder("pile", timestep=0.005, iterations=6)
  kw = dict(density=500.0, friction=(0.8, 0.005, 0.0001))
  b.add_geom(0, GEOM_PLANE, (10, 10, 0.1), name="floor", **kw)
  b.add_geom(0, GEOM_CAPSULE, (0.06,), fromto=(-0.6, -0.25, 0.06, 0.6, -0.25, 0.06),
             name="rail", **kw)
  specs = [("ball1", (0.0, -0.06, 0.13), GEOM_SPHERE, (0.1,), None),
           ("ball2", (0.22, 0.05, 0.1), GEOM_SPHERE, (0.08,), None),
           ("rod1", (0.05, 0.02, 0.36), GEOM_CAPSULE, (0.05,), (-0.2, 0, 0, 0.2, 0, 0)),
           ("rod2", (0.0, -0.05, 0.55), GEOM_CAPSULE, (0.04,), (0, -0.15, 0, 0, 0.15, 0))]
  for name, pos, typ, size, fromto in specs:
    body = b.add_body(name, 0, pos=pos)
    b.add_free_joint(body, name)
    b.add_geom(body, typ, size, fromto=fromto, name=name, **kw)
  return b.build()


def free_body(gravity=(0.0, 0.0, -9.81), radius: float = 0.1,
              mass: float = 1.0, timestep: float = 0.01,
              pos=(0.0, 0.0, 1.0)) -> System:
  """One free sphere with no contacts: closed-form integrator checks."""
  b = ModelBuilder("free_body", timestep=timestep, iterations=4,
                   gravity=gravity)
  body = b.add_body("ball", 0, pos=pos)
  b.add_free_joint(body)
  b.add_geom(body, GEOM_SPHERE, (radius,), mass=mass, contype=0,
             conaffinity=0, name="ball")
  return b.build()


def from_arrays_check(sys: System) -> None:
  """Validates the parts of a System the step supports (mjcf.py:236-314)."""
  if sys.enable_fluid:
    raise NotImplementedError("fluid forces are not part of this step")
  if not np.allclose(sys.link.joint.rot, [1.0, 0, 0, 0]):
    raise NotImplementedError("link joint frames must be unrotated")
  for t in sys.link_types:
    if t not in Q_WIDTHS:
      raise NotImplementedError(f"link type {t!r}")
  _ = sys.contact_pairs()


# ---------------------------------------------------------------------------
# MJCF subset from XML (brax.io.mjcf.loads / load_model, brax/io/mjcf.py:317-512)

_MJCF_GEOMS = {"plane": GEOM_PLANE, "sphere": GEOM_SPHERE, "capsule": GEOM_CAPSULE}
_MJCF_IGNORED = {"camera", "light", "site"}
# top-level elements read (compiler ... actuator) or without effect on the
# step (assets, sizes, visuals, statistics, sensors, keyframes)
_MJCF_TOP = {"compiler", "option", "custom", "default", "worldbody", "actuator", "asset",
             "size", "visual", "statistic", "sensor", "keyframe"}
# attributes with no effect on the generalized step: names, visuals, and
# fields the step does not read (a free joint's margin, geom user data)
_MJCF_VISUAL = {"name", "rgba", "material", "user", "group"}
_MJCF_ATTRS = {
    "geom": {"type", "size", "pos", "quat", "fromto", "density", "mass", "contype",
             "conaffinity", "friction", "condim", "solref", "solimp"},
    "joint": {"type", "axis", "pos", "range", "limited", "armature", "damping",
              "stiffness", "solreflimit", "solimplimit", "margin"},
    "body": {"pos", "quat", "childclass"},
    "motor": {"joint", "gear", "ctrlrange", "ctrllimited", "forcerange", "forcelimited"},
}
_MJCF_ATTRS["position"] = _MJCF_ATTRS["motor"] | {"kp", "kv"}
_MJCF_ATTRS["velocity"] = _MJCF_ATTRS["motor"] | {"kv"}


def _mjcf_check(tag: str, a: dict) -> None:
  extra = set(a) - _MJCF_ATTRS[tag] - _MJCF_VISUAL
  if extra:
    raise NotImplementedError(f"<{tag}> attributes {sorted(extra)}")
  if a.get("condim", "3") != "3":
    raise NotImplementedError("condim other than 3")
  if "childclass" in a:
    raise NotImplementedError("default classes")
  if tag == "joint" and "margin" in a and a.get("type", "hinge") != "free":
    raise NotImplementedError("joint limit margin")


def _floats(s) -> List[float]:
  return [float(v) for v in str(s).split()]


def load_mjcf(xml: str, name: Optional[str] = None) -> System:
  """Compiles MJCF text into a System: the subset the step supports.

  Read: <compiler angle>, <option timestep iterations gravity>, brax's
  <custom> numerics init_qpos, solver_maxls, matrix_inv_iterations
  (defaults 20 and 10 as mjcf.py:154-172; the other numerics belong to the
  positional and spring pipelines), the unnamed top-level <default> of
  joint / geom / motor, <worldbody> bodies (pos, quat) with free or hinge
  joints (axis, pos, range, limited, armature, damping, stiffness,
  solreflimit, solimplimit), plane / sphere / capsule geoms (size, pos,
  quat, fromto, density, mass, contype, conaffinity, friction, solref,
  solimp) and <actuator> motor / position / velocity actuators (joint,
  gear, ctrlrange, ctrllimited, forcerange, forcelimited, kp, kv).
  Jointless bodies are fused into their parents (mjcf.py:76-98).  The
  reference compiles the XML with MuJoCo first; this reads the same subset
  directly.  Any other element or attribute that changes the physics raises
  NotImplementedError; names, visuals, cameras, lights, sites and assets
  are ignored.
  """
  import xml.etree.ElementTree as ET
  root = ET.fromstring(xml)
  if root.tag != "mujoco":
    raise ValueError("not an MJCF document")
  for c in root:
    if c.tag not in _MJCF_TOP:
      raise NotImplementedError(f"top-level element <{c.tag}>")
  comp = root.find("compiler")
  comp = comp.attrib if comp is not None else {}
  degrees = comp.get("angle", "degree") == "degree"
  if comp.get("coordinate", "local") != "local":
    raise NotImplementedError("global coordinates")
  if comp.get("inertiafromgeom", "true") not in ("true", "auto"):
    raise NotImplementedError("inertia other than from geoms")
  autolimits = comp.get("autolimits", "true") == "true"

  def limited(a: dict, flag: str, rng) -> bool:
    v = a.get(flag, "auto")
    return (rng is not None and autolimits) if v == "auto" else v == "true"
  opt = root.find("option")
  if opt is not None and len(opt):
    raise NotImplementedError("<option> children (flags, overrides)")
  opt = opt.attrib if opt is not None else {}
  if set(opt) - {"timestep", "iterations", "gravity"}:
    raise NotImplementedError(f"<option> attributes {sorted(set(opt))}")
  custom = {}
  for c in root.findall("custom"):
    for num in c.findall("numeric"):
      custom[num.get("name")] = _floats(num.get("data"))
  defaults = {"joint": {}, "geom": {}, "motor": {}, "position": {}, "velocity": {}}
  for d in root.findall("default"):
    if d.get("class") or d.findall("default"):
      raise NotImplementedError("named or nested default classes")
    for k in defaults:
      e = d.find(k)
      if e is not None:
        defaults[k].update(e.attrib)
  b = ModelBuilder(
      name or root.get("model", "mjcf"), timestep=float(opt.get("timestep", 0.002)),
      iterations=int(opt.get("iterations", 100)),
      solver_maxls=int(custom.get("solver_maxls", [20])[0]),
      matrix_inv_iterations=int(custom.get("matrix_inv_iterations", [10])[0]),
      gravity=_floats(opt.get("gravity", "0 0 -9.81")))
  scale = np.pi / 180.0 if degrees else 1.0

  def geom(body: int, e):
    a = dict(defaults["geom"], **e.attrib)
    _mjcf_check("geom", a)
    typ = a.get("type", "sphere")
    if typ not in _MJCF_GEOMS:
      raise NotImplementedError(f"geom type {typ!r}")
    size = _floats(a.get("size", "0"))
    kw = dict(contype=int(a.get("contype", 1)), conaffinity=int(a.get("conaffinity", 1)),
              density=float(a.get("density", 1000.0)), name=a.get("name", ""))
    fr = _floats(a.get("friction", "1 0.005 0.0001"))
    kw["friction"] = tuple(fr + [0.005, 0.0001][len(fr) - 1:]) if len(fr) < 3 else tuple(fr)
    if "mass" in a:
      kw["mass"] = float(a["mass"])
    for k in ("solref", "solimp"):
      if k in a:
        kw[k] = tuple(_floats(a[k]))
    if "fromto" in a:
      kw["fromto"] = _floats(a["fromto"])
    else:
      kw["pos"] = _floats(a.get("pos", "0 0 0"))
      kw["quat"] = _floats(a.get("quat", "1 0 0 0"))
    b.add_geom(body, _MJCF_GEOMS[typ], size, **kw)

  def body(e, parent: int):
    _mjcf_check("body", e.attrib)
    bid = b.add_body(e.get("name", ""), parent, pos=_floats(e.get("pos", "0 0 0")),
                     quat=_floats(e.get("quat", "1 0 0 0")))
    for c in e:
      if c.tag == "joint":
        a = dict(defaults["joint"], **c.attrib)
        _mjcf_check("joint", a)
        typ = a.get("type", "hinge")
        if typ == "free":
          b.add_free_joint(bid, a.get("name", "root"))
        elif typ == "hinge":
          rng = _floats(a["range"]) if "range" in a else None
          lim = limited(a, "limited", rng)
          b.add_hinge(bid, a.get("name", ""), _floats(a.get("axis", "0 0 1")),
                      pos=_floats(a.get("pos", "0 0 0")),
                      range_rad=(rng[0] * scale, rng[1] * scale) if lim and rng else None,
                      armature=float(a.get("armature", 0.0)),
                      damping=float(a.get("damping", 0.0)),
                      stiffness=float(a.get("stiffness", 0.0)),
                      **{k: tuple(_floats(a[m])) for k, m in
                         (("solref", "solreflimit"), ("solimp", "solimplimit")) if m in a})
        else:
          raise NotImplementedError(f"joint type {typ!r}")
      elif c.tag == "geom":
        geom(bid, c)
      elif c.tag == "body":
        body(c, bid)
      elif c.tag == "inertial":
        raise NotImplementedError("explicit inertials (inertia comes from the geoms)")
      elif c.tag not in _MJCF_IGNORED:
        raise NotImplementedError(f"body element <{c.tag}>")

  for wb in root.findall("worldbody"):
    for c in wb:
      if c.tag == "geom":
        geom(0, c)
      elif c.tag == "body":
        body(c, 0)
      elif c.tag not in _MJCF_IGNORED:
        raise NotImplementedError(f"worldbody element <{c.tag}>")
  # actuators (mjcf.py:394-423): motor gain 1; position kp: gain kp, bias
  # -kp q (and -kv qd); velocity kv: gain kv, bias -kv qd
  for act in root.findall("actuator"):
    for m in act:
      if m.tag not in ("motor", "position", "velocity"):
        raise NotImplementedError(f"actuator <{m.tag}>")
      a = dict(defaults[m.tag], **m.attrib)
      _mjcf_check(m.tag, a)
      crng = _floats(a["ctrlrange"]) if "ctrlrange" in a else None
      frng = _floats(a["forcerange"]) if "forcerange" in a else None
      gain, bq, bqd = 1.0, 0.0, 0.0
      if m.tag == "position":
        gain = float(a.get("kp", 1.0))
        bq, bqd = -gain, -float(a.get("kv", 0.0))
      elif m.tag == "velocity":
        gain = float(a.get("kv", 1.0))
        bqd = -gain
      b.add_actuator(a["joint"], float(_floats(a.get("gear", "1"))[0]), gain=gain,
                     bias_q=bq, bias_qd=bqd,
                     ctrlrange=tuple(crng) if limited(a, "ctrllimited", crng) and crng else None,
                     forcerange=tuple(frng) if limited(a, "forcelimited", frng) and frng else None)
  # per-geom elasticity (mjcf.py:154-233): the custom numeric sets every
  # geom's value (one value, or one per geom in geom-id order), then custom
  # tuple elements override it per named geom
  if "elasticity" in custom:
    v = custom["elasticity"]
    order = sorted(b.geoms, key=lambda x: x.body)
    if len(v) not in (1, len(order)):
      raise ValueError(f"custom elasticity needs 1 or {len(order)} values, got {len(v)}")
    for i, g in enumerate(order):
      g.elasticity = float(v[0] if len(v) == 1 else v[i])
  for c in root.findall("custom"):
    for tup in c.findall("tuple"):
      if tup.get("name") != "elasticity":
        continue
      by_name = {g.name: g for g in b.geoms if g.name}
      for el in tup.findall("element"):
        if el.get("objtype") != "geom":
          raise NotImplementedError("custom elasticity tuples hold geoms")
        by_name[el.get("objname")].elasticity = float(el.get("prm"))
  b.bodies = _fuse_jointless(b)
  if "init_qpos" in custom:
    b.init_q_override = custom["init_qpos"]
  return b.build()
