"""Topology headers and model tables for the HIP step.

A System splits into two parts, the same split the reference makes between
its static (non-pytree) fields and its arrays (brax/base.py:471-480):

* the *topology*: link types and parents, q/qd addresses, actuator and limit
  dof maps, and the static list of contact candidates (link + geom kind).
  It is compiled into the kernel as a ``constexpr`` struct, so every
  per-link and per-dof array of the step lives in registers with static
  indices.  ``write_topology_header`` emits these structs for the registered
  systems; ``build()`` regenerates them before compiling.
* the *model table*: every number the step reads (link frames, inertias,
  dof axes, limits, solver parameters, actuator gains, geom sizes and
  contact parameters), packed by ``pack_model`` into one float32 array with
  the fixed record layout below, uploaded once per model.
"""

from __future__ import annotations

import hashlib
import os
from typing import Optional, Dict, List, Tuple

import numpy as np

from brax_amd import systems as S

# record layouts of the float32 model table (mirrored in csrc/bgx_layout.h)
HEADER = 9  # gravity[3], dt, solver_iterations, solver_maxls, tol, ls_eps, 0
LINK_STRIDE = 24  # tpos[3] trot[4] jpos[3] ipos[3] irot[4] I[6] mass
DOF_STRIDE = 24  # ang[3] vel[3] armature damping stiffness lo hi invweight imp[9]
ACT_STRIDE = 8  # ctrl_lo ctrl_hi force_lo force_hi gain gear bias_q bias_qd
CAND_STRIDE = 36  # gpos[3] gquat[4] radius halflen sign mu imp[9] tw ppos[3] pn[3] t1[3] t2[3]
IMP_SIZE = 9  # dmin dmax 1/width mid power 1/mid^(p-1) 1/(1-mid)^(p-1) b k

KIND_SPHERE = 0  # plane - sphere
KIND_CAPSULE = 1  # plane - one capsule end
KIND_SS = 2  # sphere - sphere
KIND_SC = 3  # sphere - capsule
KIND_CC = 4  # capsule - capsule
_PAIR_CAND_KIND = {S.PAIR_SPHERE_SPHERE: KIND_SS, S.PAIR_SPHERE_CAPSULE: KIND_SC,
                   S.PAIR_CAPSULE_CAPSULE: KIND_CC}

# fast-path contact-slot capacity per topology (other candidates take the
# scratch path); sized from the active-contact distribution of the workload
DEFAULT_SLOTS = 6
# fast contact slots per system name (default DEFAULT_SLOTS): the humanoid
# stands on 4-8 foot contacts and its batches leave most of each CU's LDS
# free, so every candidate gets a slot and the overflow path is not needed
SLOTS = {"humanoid": 13}
# other slot counts (LDS tiles) the step kernels are compiled for, per system
# name, for the autotuner to time: the quadruped trades LDS per env (6 L +
# 16 KC floats: 4 workgroups per CU up to KC = 6, 3 at 8) against envs
# taking the slot-per-candidate path (more than KC active contacts)
SLOT_TILES = {"quadruped": (4, 8)}

# tree systems whose topology gets the persistent batch kernel (k_batch): the
# heterogeneous workload mixes chains with quadrupeds
BATCH_SYSTEMS = {"quadruped"}

# tree systems whose step also compiles the multi-lane tree kernel
# (bgx_mtree.h: a lane per dof, the default for their homogeneous
# launches): large trees whose batches leave the chip short of wavefronts
# at one thread per env
MLANE_SYSTEMS = {"humanoid", "humanoid_xml"}

JAXOPT_TOL = 1e-3  # jaxopt ProximalGradient default tol
LS_EPS = float(np.finfo(np.float32).eps)


class Topology:
  """Static structure of a System, as consumed by the kernel templates."""

  def __init__(self, sys: S.System, slots: Optional[int] = None, per_env: bool = False):
    S.from_arrays_check(sys)
    # per_env: every env has its own model table (same structure, its own
    # numbers); otherwise one table is shared by the whole batch
    self.per_env = bool(per_env)
    if slots is None:
      slots = SLOTS.get(sys.name, DEFAULT_SLOTS)
    self.name = _ident(sys.name)
    self.link_types = sys.link_types
    self.parents = list(sys.link_parents)
    self.L = sys.num_links()
    self.nq, self.nv, self.nu = sys.q_size(), sys.qd_size(), sys.act_size()
    self.ltype = [0 if t == "f" else int(t) for t in sys.link_types]
    for l, t in enumerate(sys.link_types):
      if t == "f":  # the kernel hard-codes the free-joint motion (mjcf.py:354-356)
        dr = sys.dof_ranges()[l]
        if not (np.array_equal(sys.dof.motion.ang[dr], np.eye(6, 3, -3)) and
                np.array_equal(sys.dof.motion.vel[dr], np.eye(6, 3))):
          raise NotImplementedError("free joint with non-standard dof motion")
    self.qadr = [r[0] for r in sys.q_ranges()]
    self.dadr = [r[0] for r in sys.dof_ranges()]
    roots = []
    for i in range(self.L):
      r = i
      while self.parents[r] >= 0:
        r = self.parents[r]
      roots.append(r)
    self.root = roots
    self.dof_link = [int(v) for v in sys.dof_link()]
    # previous dof on the path to the root (-1 at a root link's first dof)
    dp = []
    for l, dr in enumerate(sys.dof_ranges()):
      for k, d in enumerate(dr):
        if k > 0:
          dp.append(d - 1)
        else:
          p = self.parents[l]
          dp.append(sys.dof_ranges()[p][-1] if p >= 0 else -1)
    self.dof_parent = dp
    if self.nv > 32:
      raise NotImplementedError("at most 32 dofs per tree")
    anc = []
    for d in range(self.nv):
      m, j = 0, dp[d]
      while j >= 0:
        m |= 1 << j
        j = dp[j]
      anc.append(m)
    self.dof_anc = anc  # bit j set: dof j is a strict ancestor of dof d
    self.act_q = [int(v) for v in sys.actuator.q_id]
    self.act_d = [int(v) for v in sys.actuator.qd_id]
    if sys.dof.limit is None:
      self.lim_d, self.lim_q = [], []
    else:
      self.lim_d = [int(v) for v in sys.qd_idx("123")]
      self.lim_q = [int(v) for v in sys.q_idx("123")]
    # contact candidates: a plane candidate is one contact of a plane pair
    # (the plane on the world, the sphere or capsule on link cand_link); a
    # pair candidate is a sphere / capsule pair whose rows act on the
    # velocity of cand_link minus that of cand_linka (-1: the world).  A
    # pair whose geom 2 is a world geom puts geom 1's link in cand_link and
    # sets cand_flip (the rows' directions change sign).
    self.cand_link, self.cand_linka, self.cand_flip = [], [], []
    self.cand_kind, self.cand_src = [], []
    for p in sys.contact_pairs():
      la = int(sys.geom_bodyid[p.geom1]) - 1
      lb = int(sys.geom_bodyid[p.geom2]) - 1
      if p.kind in _PAIR_CAND_KIND:
        link, linka, flip = (la, -1, 1) if lb < 0 else (lb, la, 0)
        cands = [(_PAIR_CAND_KIND[p.kind], 0.0)]
      else:
        if la != -1 or lb < 0:
          raise NotImplementedError("planes must belong to the world body")
        link, linka, flip = lb, -1, 0
        cands = ([(KIND_SPHERE, 0.0)] if p.kind == S.PAIR_PLANE_SPHERE
                 else [(KIND_CAPSULE, 1.0), (KIND_CAPSULE, -1.0)])
      for kind, sign in cands:
        self.cand_link.append(link)
        self.cand_linka.append(linka)
        self.cand_flip.append(flip)
        self.cand_kind.append(kind)
        self.cand_src.append((p, sign))
    self.ncand = len(self.cand_link)
    self.pairs = any(k >= KIND_SS for k in self.cand_kind)
    # every contact plane is the z-up floor and there are no pair
    # candidates: frames and distances specialize
    self.plane_z = (not self.pairs) and (all(
        np.allclose(S.quat_to_mat(sys.geom_quat[p.geom1])[:, 2], [0, 0, 1], atol=0)
        for p, _ in self.cand_src) if self.cand_src else True)
    # per dof: 1 for a pure rotation axis (motion.vel == 0)
    self.dof_hinge = [int(not np.any(sys.dof.motion.vel[d])) for d in range(self.nv)]
    self.slots = max(1, min(slots, self.ncand)) if self.ncand else 1
    alt = [k for k in SLOT_TILES.get(sys.name, ()) if 1 <= k <= self.ncand and k != self.slots]
    self.tiles = (alt + [0, 0])[:2]
    # a hinge chain (links in series from the world, one motor per hinge,
    # no contacts, no limits) with per-env tables runs the chain kernels
    # (bgx_chain.h: several lanes per env) instead of the tree kernels
    self.chain = (self.per_env and 1 <= self.L <= 32 and
                  all(t == "1" for t in sys.link_types) and
                  self.parents == list(range(-1, self.L - 1)) and
                  self.nq == self.nv == self.nu == self.L and
                  self.act_q == list(range(self.L)) and self.act_d == list(range(self.L)) and
                  not self.lim_d and self.ncand == 0 and all(self.dof_hinge))
    # the persistent heterogeneous-batch kernel is compiled for this tree
    # topology (its envs can share a launch with chains)
    self.batch = (not self.chain) and sys.name in BATCH_SYSTEMS
    self.mlane = sys.name in MLANE_SYSTEMS and self._mlane_family(sys)
    # LTL factor and limit-row solver vectors in LDS instead of registers:
    # supported, but off for every registered topology (the 16384-humanoid
    # batch ran at 5.2e5 env-steps/s with it against 5.8e5 without: at one
    # wave per CU, LDS latency in the triangular solves is not hidden)
    self.lm_lds = False
    self.off_link = HEADER
    self.off_dof = self.off_link + LINK_STRIDE * self.L
    self.off_act = self.off_dof + DOF_STRIDE * self.nv
    self.off_cand = self.off_act + ACT_STRIDE * max(self.nu, 1)
    self.size = self.off_cand + CAND_STRIDE * max(self.ncand, 1)

  def _mlane_family(self, sys) -> bool:
    """The multi-lane tree kernel's family (bgx_mtree.h mtree_ok): one tree
    rooted at link 0, a free joint only there, no sphere / capsule pairs,
    at most 32 dofs and candidates, each actuator on its own dof's
    hinge coordinate, one actuator per dof."""
    if self.per_env or self.pairs or self.nv > 32 or self.ncand > 32:
      return False
    if self.parents[0] >= 0 or any(p < 0 for p in self.parents[1:]):
      return False
    if any(t == 0 for t in self.ltype[1:]) or (self.ltype[0] == 0 and self.qadr[0] != 0):
      return False
    if len(set(self.act_d)) != len(self.act_d):
      return False
    for q_, d_ in zip(self.act_q, self.act_d):
      l = self.dof_link[d_]
      if self.ltype[l] == 0 or q_ != self.qadr[l] + (d_ - self.dadr[l]):
        return False
    return True

  def signature(self) -> str:
    key = repr((self.link_types, self.parents, self.act_q, self.act_d,
                self.lim_d, self.cand_link, self.cand_kind, self.slots,
                self.plane_z, self.dof_hinge) + ((True,) if self.per_env else ()) +
               ((tuple(self.tiles),) if any(self.tiles) else ()) +
               ((tuple(self.cand_linka), tuple(self.cand_flip)) if self.pairs else ()) +
               (("mlane",) if self.mlane else ()))
    return hashlib.sha1(key.encode()).hexdigest()[:16]

  def cpp_struct(self) -> str:
    def arr(name, vals):
      vals = list(vals) if vals else [0]
      return (f"  static constexpr int {name}[{len(vals)}] = "
              f"{{{', '.join(str(int(v)) for v in vals)}}};\n")
    s = f"struct Topo_{self.name} {{\n"
    s += f'  static constexpr const char* NAME = "{self.name}";\n'
    s += (f"  static constexpr int L = {self.L}, NQ = {self.nq}, NV = {self.nv}, "
          f"NU = {self.nu}, NLIM = {len(self.lim_d)}, NCAND = {self.ncand}, "
          f"KC = {self.slots}, PLANE_Z = {int(self.plane_z)}, "
          f"PER_ENV = {int(self.per_env)}, LM_LDS = {int(self.lm_lds)}, "
          f"CHAIN = {int(self.chain)}, BATCH = {int(self.batch)}, "
          f"TILE1 = {self.tiles[0]}, TILE2 = {self.tiles[1]}, PAIRS = {int(self.pairs)}, "
          f"MLANE = {int(self.mlane)};\n")
    s += (f"  static constexpr int OFF_LINK = {self.off_link}, OFF_DOF = "
          f"{self.off_dof}, OFF_ACT = {self.off_act}, OFF_CAND = "
          f"{self.off_cand}, SIZE = {self.size};\n")
    s += arr("LTYPE", self.ltype) + arr("PARENT", self.parents)
    s += arr("QADR", self.qadr) + arr("DADR", self.dadr) + arr("ROOT", self.root)
    s += arr("DOF_LINK", self.dof_link) + arr("DOF_PARENT", self.dof_parent)
    s += (f"  static constexpr unsigned DANC[{self.nv}] = "
          f"{{{', '.join(hex(v) + 'u' for v in self.dof_anc)}}};\n")
    s += arr("ACT_Q", self.act_q) + arr("ACT_D", self.act_d)
    s += arr("LIM_D", self.lim_d) + arr("LIM_Q", self.lim_q)
    s += arr("CAND_LINK", self.cand_link) + arr("CAND_KIND", self.cand_kind)
    s += (arr("CAND_LINKA", self.cand_linka) if self.cand_linka else
          "  static constexpr int CAND_LINKA[1] = {-1};\n")
    s += arr("CAND_FLIP", self.cand_flip)
    s += arr("DOF_HINGE", self.dof_hinge)
    s += f'  static constexpr const char* SIG = "{self.signature()}";\n'
    s += "};\n"
    return s


def impedance_record(solref, solimp) -> np.ndarray:
  """Constants of constraint.py:29-65 for solver params solref(2)+solimp(5).

  imp(pos) = clip(dmin + y (dmax - dmin)) with x = |pos| / width,
  y = x^p / mid^(p-1) below mid, 1 - (1-x)^p / (1-mid)^(p-1) above;
  aref = -b vel - k imp pos with b = 2 / (dmax timeconst) (or -dampratio /
  dmax if dampratio <= 0) and k = 1 / (dmax^2 timeconst^2 dampratio^2) (or
  -timeconst / dmax^2 if timeconst <= 0)."""
  tc, dr = float(solref[0]), float(solref[1])
  dmin, dmax, width, mid, power = [float(v) for v in solimp]
  b = -dr / dmax if dr <= 0 else 2.0 / (dmax * tc)
  k = -tc / (dmax * dmax) if tc <= 0 else 1.0 / (dmax * dmax * tc * tc * dr * dr)
  return np.array([dmin, dmax, 1.0 / width, mid, power,
                   1.0 / mid ** (power - 1), 1.0 / (1.0 - mid) ** (power - 1),
                   b, k])


def _ident(name: str) -> str:
  out = "".join(c if c.isalnum() else "_" for c in name)
  return out if out and not out[0].isdigit() else "t_" + out


def pack_model(sys: S.System, topo: Topology = None) -> np.ndarray:
  """float32 model table of ``sys`` in the layout of its topology."""
  topo = topo or Topology(sys)
  m = np.zeros(topo.size, np.float64)
  m[0:3] = sys.gravity
  m[3] = sys.opt.timestep
  m[4] = sys.solver_iterations
  m[5] = sys.solver_maxls
  m[6] = JAXOPT_TOL
  m[7] = LS_EPS
  for l in range(topo.L):
    o = topo.off_link + l * LINK_STRIDE
    m[o:o + 3] = sys.link.transform.pos[l]
    m[o + 3:o + 7] = sys.link.transform.rot[l]
    m[o + 7:o + 10] = sys.link.joint.pos[l]
    m[o + 10:o + 13] = sys.link.inertia.transform.pos[l]
    m[o + 13:o + 17] = sys.link.inertia.transform.rot[l]
    i = sys.link.inertia.i[l]
    m[o + 17:o + 23] = [i[0, 0], i[1, 1], i[2, 2], i[0, 1], i[0, 2], i[1, 2]]
    m[o + 23] = sys.link.inertia.mass[l]
  lo = hi = None
  if sys.dof.limit is not None:
    lo, hi = sys.dof.limit
  for d in range(topo.nv):
    o = topo.off_dof + d * DOF_STRIDE
    m[o:o + 3] = sys.dof.motion.ang[d]
    m[o + 3:o + 6] = sys.dof.motion.vel[d]
    m[o + 6] = sys.dof.armature[d]
    m[o + 7] = sys.dof.damping[d]
    m[o + 8] = sys.dof.stiffness[d]
    m[o + 9] = -np.inf if lo is None else lo[d]
    m[o + 10] = np.inf if hi is None else hi[d]
    m[o + 11] = sys.dof.invweight[d]
    sp = sys.dof.solver_params[d]
    m[o + 12:o + 12 + IMP_SIZE] = impedance_record(sp[:2], sp[2:])
  a = sys.actuator
  for u in range(topo.nu):
    o = topo.off_act + u * ACT_STRIDE
    m[o:o + 8] = [a.ctrl_range[u, 0], a.ctrl_range[u, 1], a.force_range[u, 0],
                  a.force_range[u, 1], a.gain[u], a.gear[u], a.bias_q[u],
                  a.bias_qd[u]]
  for c in range(topo.ncand):
    p, sign = topo.cand_src[c]
    o = topo.off_cand + c * CAND_STRIDE
    g2, g1 = p.geom2, p.geom1
    if topo.cand_kind[c] >= KIND_SS:
      # geom 2, then geom 1 in the plane's fields: local pose (the world
      # pose for a world geom), radius, half length
      m[o:o + 3] = sys.geom_pos[g2]
      m[o + 3:o + 7] = sys.geom_quat[g2]
      m[o + 7] = sys.geom_size[g2][0]
      m[o + 8] = sys.geom_size[g2][1]
      m[o + 10] = p.friction[0]
      m[o + 11:o + 11 + IMP_SIZE] = impedance_record(p.solref, p.solimp)
      la = int(sys.geom_bodyid[g1]) - 1
      lb = int(sys.geom_bodyid[g2]) - 1
      # constraint.py:168: invweight[link_a] (link_a > -1) + invweight[link_b];
      # a world geom 2 (link_b = -1) indexes the last link, as jax does
      m[o + 20] = sys.link.invweight[la] * (la > -1) + sys.link.invweight[lb]
      m[o + 21:o + 24] = sys.geom_pos[g1]
      m[o + 24:o + 28] = sys.geom_quat[g1]
      m[o + 28] = sys.geom_size[g1][0]
      m[o + 29] = sys.geom_size[g1][1]
      continue
    m[o:o + 3] = sys.geom_pos[g2]
    m[o + 3:o + 7] = sys.geom_quat[g2]
    m[o + 7] = sys.geom_size[g2][0]
    m[o + 8] = sys.geom_size[g2][1]
    m[o + 9] = sign
    m[o + 10] = p.friction[0]
    m[o + 11:o + 11 + IMP_SIZE] = impedance_record(p.solref, p.solimp)
    lb = int(sys.geom_bodyid[g2]) - 1
    m[o + 20] = sys.link.invweight[lb]  # world side contributes 0
    # plane on the world body: pose = geom pose (contact.py:48-53)
    pr = S.quat_to_mat(sys.geom_quat[g1])
    n = pr[:, 2]
    m[o + 21:o + 24] = sys.geom_pos[g1]
    m[o + 24:o + 27] = n
    # sphere frame: [n, b, n x b] with b = orthogonals(n) (math.py:247-254)
    nn = n / np.linalg.norm(n)
    b = np.array([0.0, 1.0, 0.0]) if -0.5 < nn[1] < 0.5 else np.array([0.0, 0.0, 1.0])
    b = b - nn * nn.dot(b)
    b = b / np.linalg.norm(b)
    m[o + 27:o + 30] = b
    m[o + 30:o + 33] = np.cross(nn, b)
  return m.astype(np.float32)


# ---------------------------------------------------------------------------
# registry of topologies compiled into the library


CHAIN_LINKS = range(2, 33)  # hinge chains of the heterogeneous-batch workload


def registered_systems() -> List[Tuple[S.System, bool]]:
  """(system, per-env tables) whose topologies are compiled into
  libbrax_gen.so."""
  out = [
      (S.quadruped(0), False),
      (S.ant(), False),
      (S.pendulum(1, motor_gear=150.0), False),
      (S.pendulum(3), False),
      (S.humanoid(0), False),
      (S.humanoid_xml(), False),
      (S.two_spheres(), False),
      (S.pile(), False),
  ]
  out += [(S.chain(n, 0), True) for n in CHAIN_LINKS]
  return out



def registered_topologies() -> List[Topology]:
  seen: Dict[str, Topology] = {}
  for sys, per_env in registered_systems():
    t = Topology(sys, per_env=per_env)
    seen.setdefault(t.signature(), t)
  return list(seen.values())


def write_topology_header(path: str) -> Tuple[str, List[Topology]]:
  topos = registered_topologies()
  names = set()
  out = ["// Generated by brax_amd/codegen.py from the registered systems.",
         "// Static structure of each topology compiled into the library.",
         "#pragma once", ""]
  for t in topos:
    if t.name in names:
      t.name = f"{t.name}_{t.signature()[:6]}"
    names.add(t.name)
    out.append(t.cpp_struct())
  for macro, keep in (("BGX_FOR_EACH_TOPOLOGY", lambda t: True),
                      ("BGX_FOR_EACH_TREE_TOPOLOGY", lambda t: not t.chain),
                      ("BGX_FOR_EACH_CHAIN_TOPOLOGY", lambda t: t.chain)):
    out.append(f"#define {macro}(X) \\")
    for i, t in enumerate(topos):
      if keep(t):
        out.append(f"  X({i}, Topo_{t.name}) \\")
    out.append("")
  out.append(f"#define BGX_NUM_TOPOLOGIES {len(topos)}")
  text = "\n".join(out) + "\n"
  old = open(path).read() if os.path.exists(path) else None
  if old != text:
    with open(path, "w") as f:
      f.write(text)
  return text, topos
