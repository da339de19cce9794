"""Oracle: the reference's generalized pipeline restated in torch on the CPU.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Every function follows the
reference file:line it cites, batched over a leading env axis N, in float64
(default) or float32.  Two pieces of the path live in third-party code that is
not under the reference tree; they are restated from their published source:

* ``pg_solve``: jaxopt.ProjectedGradient (FISTA with backtracking line search,
  ``jaxopt`` unpinned in pyproject.toml:41; call site constraint.py:229-238).
* ``collide``: MJX narrow phase for plane-sphere, plane-capsule,
  sphere-sphere, sphere-capsule and capsule-capsule pairs (``mujoco-mjx``,
  pyproject.toml:44-45; call site contact.py:57).

Parity pins (tests that would fail on a plausible mistake):
  scan schedules ........ scan_test.py KATs (tests/test_oracle_kat.py)
  spatial algebra ....... closed forms (rotation orthogonality, q*q^-1)
  kinematics ............ hand-derived pendulum FK; FD of positions vs xd
  contact geometry ...... contact_test.py:29-44 ant KAT and :67-75 sphere-sphere
                          KAT; closed forms; segment closest points against
                          brute force
  mass matrix ........... brute-force sum_links J^T I J (independent code)
  RNEA bias ............. FD of potential energy at qd=0; FD of momentum
  actuator .............. actuator_test.py:57-60 and :114-118 closed forms
  Newton-Schulz ......... converges to numpy.linalg.inv (math_test.py:38-46)
  constraint rows ....... J = FD of contact-point / joint position along qd;
                          _imp_aref closed form
  PG solve .............. converged limit = scipy.optimize.nnls; x >= 0;
                          objective non-increasing (jaxopt semantics)
  integrator ............ free fall / constant-omega closed forms
The exact 4-iteration iterate of jaxopt's line search (the form of its
acceptance test) and the MJX capsule tangent frame cannot be pinned by any
fixture available here: parity unpinned for those two details.
"""

from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional

import numpy as np
import torch

EPS_F32 = float(np.finfo(np.float32).eps)


def _t(x, dtype):
  return torch.as_tensor(np.asarray(x), dtype=dtype)


# ---------------------------------------------------------------------------
# math (brax/math.py)


def cross(a, b):
  a, b = torch.broadcast_tensors(a, b)
  return torch.stack([
      a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
      a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
      a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def dot(a, b):
  return (a * b).sum(-1)


def rotate(vec, quat):
  """math.py:25-40."""
  s, u = quat[..., 0:1], quat[..., 1:]
  r = 2 * (dot(u, vec)[..., None] * u) + (s * s - dot(u, u)[..., None]) * vec
  return r + 2 * s * cross(u, vec)


def quat_mul(u, v):
  """math.py:86-101."""
  u, v = torch.broadcast_tensors(u, v)
  return torch.stack([
      u[..., 0] * v[..., 0] - u[..., 1] * v[..., 1] - u[..., 2] * v[..., 2] - u[..., 3] * v[..., 3],
      u[..., 0] * v[..., 1] + u[..., 1] * v[..., 0] + u[..., 2] * v[..., 3] - u[..., 3] * v[..., 2],
      u[..., 0] * v[..., 2] - u[..., 1] * v[..., 3] + u[..., 2] * v[..., 0] + u[..., 3] * v[..., 1],
      u[..., 0] * v[..., 3] + u[..., 1] * v[..., 2] - u[..., 2] * v[..., 1] + u[..., 3] * v[..., 0]], -1)


def quat_inv(q):
  """math.py:122-131."""
  return q * torch.tensor([1.0, -1.0, -1.0, -1.0], dtype=q.dtype)


def quat_rot_axis(axis, angle):
  """math.py:134-148."""
  s = torch.sin(angle / 2)[..., None]
  return torch.cat([torch.cos(angle / 2)[..., None], axis * s], -1)


def quat_to_3x3(q):
  """math.py:151-165 (rows of the rotation matrix)."""
  d = dot(q, q)
  w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
  s = 2 / d
  xs, ys, zs = x * s, y * s, z * s
  wx, wy, wz = w * xs, w * ys, w * zs
  xx, xy, xz = x * xs, x * ys, x * zs
  yy, yz, zz = y * ys, y * zs, z * zs
  return torch.stack([
      torch.stack([1 - (yy + zz), xy - wz, xz + wy], -1),
      torch.stack([xy + wz, 1 - (xx + zz), yz - wx], -1),
      torch.stack([xz - wy, yz + wx, 1 - (xx + yy)], -1)], -2)


def safe_norm(x, ndims=1):
  """math.py:308-328: norm over the trailing ``ndims`` dims, 0 if all ~0."""
  flat = x.reshape(x.shape[:x.dim() - ndims] + (-1,))
  # jp.allclose(x, 0): |x| <= atol + rtol*0 with atol=1e-8
  is_zero = (flat.abs() <= 1e-8).all(-1).to(x.dtype)
  f = flat + is_zero[..., None]
  return torch.linalg.norm(f, dim=-1) * (1.0 - is_zero)


def normalize(x, ndims=1):
  """math.py:331-345."""
  n = safe_norm(x, ndims)
  shape = n.shape + (1,) * ndims
  return x / (n + 1e-6 * (n == 0.0).to(x.dtype)).reshape(shape), n


def orthogonals(a):
  """math.py:247-254."""
  y = torch.tensor([0.0, 1.0, 0.0], dtype=a.dtype).expand_as(a)
  z = torch.tensor([0.0, 0.0, 1.0], dtype=a.dtype).expand_as(a)
  c = ((-0.5 < a[..., 1]) & (a[..., 1] < 0.5))[..., None]
  b = torch.where(c, y, z)
  b = b - a * dot(a, b)[..., None]
  b = normalize(b)[0] * (a != 0).any(-1, keepdim=True).to(a.dtype)
  return b, cross(a, b)


# -- MJX narrow-phase primitives for sphere and capsule pairs ----------------
# mujoco-mjx (pyproject.toml:44-45, mujoco>=3.8.1; not under the reference
# tree): mjx/_src/collision_primitive.py (_sphere_sphere, sphere_capsule,
# capsule_capsule) and mjx/_src/math.py (make_frame,
# closest_segment_point[_and_dist], closest_segment_to_segment_points),
# restated from the published source; called through contact.py:57.


def make_frame(a):
  """A right-handed frame with rows a / |a|, then orthogonals of it."""
  a = normalize(a)[0]
  b, c = orthogonals(a)
  return torch.stack([a, b, c], -2)


def sphere_sphere(pos1, r1, pos2, r2):
  """One contact of two spheres: normal from sphere 1 to sphere 2 (the x
  axis for coincident centres), dist = |p2 - p1| - r1 - r2, pos midway
  through the overlap."""
  n, d = normalize(pos2 - pos1)
  ex = torch.tensor([1.0, 0.0, 0.0], dtype=pos1.dtype).expand_as(n)
  n = torch.where((d == 0.0)[..., None], ex, n)
  dist = d - (r1 + r2)
  pos = pos1 + n * (r1 + dist * 0.5)[..., None]
  return dist, pos, make_frame(n)


def closest_segment_point(a, b, pt):
  """The point of segment a-b closest to pt (a 1e-6 regularized projection)."""
  ab = b - a
  t = dot(pt - a, ab) / (dot(ab, ab) + 1e-6)
  return a + torch.clamp(t, 0.0, 1.0)[..., None] * ab


def closest_segment_point_and_dist(a, b, pt):
  c = closest_segment_point(a, b, pt)
  return c, dot(pt - c, pt - c)


def closest_segment_to_segment_points(a0, a1, b0, b1):
  """Closest points of segments a0-a1 and b0-b1: the closest points of the
  two lines clipped to the segments, then, if the other segment's point
  projected back is nearer, that pair."""
  dir_a, len_a = normalize(a1 - a0)
  dir_b, len_b = normalize(b1 - b0)
  half_a, half_b = len_a * 0.5, len_b * 0.5
  a_mid = a0 + dir_a * half_a[..., None]
  b_mid = b0 + dir_b * half_b[..., None]
  trans = a_mid - b_mid
  dd = dot(dir_a, dir_b)
  dat = dot(dir_a, trans)
  dbt = dot(dir_b, trans)
  denom = 1 - dd * dd
  ta0 = (-dat + dd * dbt) / (denom + 1e-6)
  tb0 = dbt + ta0 * dd
  ta = torch.minimum(torch.maximum(ta0, -half_a), half_a)
  tb = torch.minimum(torch.maximum(tb0, -half_b), half_b)
  best_a = a_mid + dir_a * ta[..., None]
  best_b = b_mid + dir_b * tb[..., None]
  new_a, d1 = closest_segment_point_and_dist(a0, a1, best_b)
  new_b, d2 = closest_segment_point_and_dist(b0, b1, best_a)
  c = (d1 < d2)[..., None]
  return torch.where(c, new_a, best_a), torch.where(c, best_b, new_b)


def inv_approximate(a, a_inv, num_iter):
  """math.py:278-305: warm-started Newton-Schulz inverse, batched."""
  n = a.shape[-1]
  eye = torch.eye(n, dtype=a.dtype)
  r0 = eye - a @ a_inv
  reset = safe_norm(r0, 2) > 1
  tr = torch.diagonal(a @ a.transpose(-1, -2), dim1=-2, dim2=-1).sum(-1)
  a_inv = torch.where(reset[..., None, None],
                      0.5 * a.transpose(-1, -2) / tr[..., None, None], a_inv)
  r = r0
  err = torch.ones(a.shape[:-2], dtype=a.dtype)
  for _ in range(num_iter):
    nxt = a_inv @ (eye + r)
    r_nxt = eye - a @ nxt
    err_nxt = safe_norm(r_nxt, 2)
    a_inv = torch.where((err_nxt < err)[..., None, None], nxt, a_inv)
    r, err = r_nxt, err_nxt
  return a_inv


# ---------------------------------------------------------------------------
# spatial algebra (brax/base.py:543-614)


def tf_do_tf(pos, rot, pos2, rot2):
  """base.py:559-563: Transform.do(Transform)."""
  return pos + rotate(pos2, rot), quat_mul(rot, rot2)


def tf_do_motion(pos, rot, ang, vel):
  """base.py:566-571: Transform.do(Motion)."""
  rot_t = quat_inv(rot)
  return rotate(ang, rot_t), rotate(vel - cross(pos, ang), rot_t)


def tf_inv_do_motion(pos, rot, ang, vel):
  """base.py:574-579: Transform.inv_do(Motion)."""
  a = rotate(ang, rot)
  return a, rotate(vel, rot) + cross(pos, a)


def motion_cross_motion(s_ang, s_vel, m_ang, m_vel):
  """base.py:604-608: self.cross(m) for a Motion m."""
  return cross(s_ang, m_ang), cross(s_ang, m_vel) + cross(s_vel, m_ang)


def motion_cross_force(s_ang, s_vel, f_ang, f_vel):
  """base.py:611-614: self.cross(f) for a Force f."""
  return cross(s_ang, f_ang) + cross(s_vel, f_vel), cross(s_ang, f_vel)


def inertia_mul(c, i, mass, ang, vel):
  """base.py:297-302: Inertia.mul(Motion); c = transform.pos (mass-weighted)."""
  fa = (i @ ang[..., None])[..., 0] + cross(c, vel)
  fv = mass[..., None] * vel - cross(c, ang)
  return fa, fv


def tf_do_inertia(pos, rot, i, mass):
  """base.py:589-594: Transform.do(Inertia)."""
  eye = torch.eye(3, dtype=pos.dtype)
  h = cross(pos[..., None, :], -eye)  # rows: cross(pos, -e_k)
  r = quat_to_3x3(rot)
  inew = r @ i @ r.transpose(-1, -2) + (h @ h.transpose(-1, -2)) * mass[..., None, None]
  return pos * mass[..., None], rot, inew, mass


# ---------------------------------------------------------------------------
# scan schedules (brax/scan.py)


def depth_levels(sys) -> List[Dict[str, List[int]]]:
  """scan.py:88-99: link/q/qd index sets of each depth level."""
  from brax_amd.systems import Q_WIDTHS, QD_WIDTHS
  depth = sys.link_depths()
  levels = []
  qi = di = 0
  for i, t in enumerate(sys.link_types):
    while depth[i] >= len(levels):
      levels.append({"l": [], "q": [], "d": []})
    levels[depth[i]]["l"].append(i)
    levels[depth[i]]["q"].extend(range(qi, qi + Q_WIDTHS[t]))
    levels[depth[i]]["d"].extend(range(di, di + QD_WIDTHS[t]))
    qi += Q_WIDTHS[t]
    di += QD_WIDTHS[t]
  return levels


def type_groups(sys) -> List[Dict[str, object]]:
  """scan.py:160-171: link/q/qd index sets per link type, first-seen order."""
  from brax_amd.systems import Q_WIDTHS, QD_WIDTHS
  order = sorted(set(sys.link_types), key=sys.link_types.find)
  groups = [{"typ": t, "l": [], "q": [], "d": []} for t in order]
  qi = di = 0
  for i, t in enumerate(sys.link_types):
    g = groups[order.index(t)]
    g["l"].append(i)
    g["q"].extend(range(qi, qi + Q_WIDTHS[t]))
    g["d"].extend(range(di, di + QD_WIDTHS[t]))
    qi += Q_WIDTHS[t]
    di += QD_WIDTHS[t]
  return groups


def tree_forward(sys, own, fn):
  """scan.py:113-125 (root to leaves): y[l] = fn(y[parent] or None, l)."""
  del own
  out = {}
  for lvl in depth_levels(sys):
    for l in lvl["l"]:
      p = sys.link_parents[l]
      out[l] = fn(out[p] if p >= 0 else None, l)
  return [out[l] for l in range(sys.num_links())]


def tree_reverse_sum(sys, own: List[tuple]):
  """scan.py:101-112 with an additive carry: y[l] = own[l] + sum(children)."""
  L = sys.num_links()
  out = [None] * L
  levels = depth_levels(sys)
  for d in range(len(levels) - 1, -1, -1):
    for l in levels[d]["l"]:
      kids = [c for c in range(L) if sys.link_parents[c] == l]
      y = own[l]
      if kids:
        acc = tuple(torch.zeros_like(v) for v in own[l])
        for c in kids:
          acc = tuple(a + v for a, v in zip(acc, out[c]))
        y = tuple(a + b for a, b in zip(y, acc))
      out[l] = y
  return out


# ---------------------------------------------------------------------------
# the oracle


@dataclasses.dataclass
class OState:
  """generalized State (brax/generalized/base.py:25-60), batched over envs."""

  q: torch.Tensor
  qd: torch.Tensor
  x_pos: torch.Tensor
  x_rot: torch.Tensor
  xd_ang: torch.Tensor
  xd_vel: torch.Tensor
  root_com: Optional[torch.Tensor] = None
  cinr: Optional[tuple] = None  # (c, rot, i, mass) per link
  cd: Optional[tuple] = None  # (ang, vel) (N, L, 3)
  cdof: Optional[tuple] = None  # (ang, vel) (N, nv, 3)
  cdofd: Optional[tuple] = None
  mass_mx: Optional[torch.Tensor] = None
  mass_mx_inv: Optional[torch.Tensor] = None
  con_jac: Optional[torch.Tensor] = None
  con_diag: Optional[torch.Tensor] = None
  con_aref: Optional[torch.Tensor] = None
  qf_smooth: Optional[torch.Tensor] = None
  qf_constraint: Optional[torch.Tensor] = None
  qdd: Optional[torch.Tensor] = None
  contact: Optional[dict] = None
  pg_info: Optional[dict] = None


class Oracle:
  """Restatement of brax.generalized.pipeline.init/step for one System.

  ``inverse='ns'`` keeps the reference's warm-started Newton-Schulz inverse
  (mass.py:101-102); ``inverse='exact'`` solves M exactly every step (what
  the reference does at init, mass.py:103-104).
  """

  def __init__(self, sys, dtype=torch.float64, inverse: str = "ns",
               ls_eps: float = EPS_F32):
    self.sys = sys
    self.dtype = dtype
    self.inverse = inverse
    self.ls_eps = ls_eps
    T = lambda x: _t(x, dtype)  # noqa: E731
    s = sys
    self.L = s.num_links()
    self.nv = s.qd_size()
    self.nq = s.q_size()
    self.gravity = T(s.gravity)
    self.dt = float(s.opt.timestep)
    self.lt_pos, self.lt_rot = T(s.link.transform.pos), T(s.link.transform.rot)
    self.lj_pos, self.lj_rot = T(s.link.joint.pos), T(s.link.joint.rot)
    self.li_pos = T(s.link.inertia.transform.pos)
    self.li_rot = T(s.link.inertia.transform.rot)
    self.li_i, self.li_mass = T(s.link.inertia.i), T(s.link.inertia.mass)
    self.l_invweight = T(s.link.invweight)
    self.m_ang, self.m_vel = T(s.dof.motion.ang), T(s.dof.motion.vel)
    self.armature, self.stiffness = T(s.dof.armature), T(s.dof.stiffness)
    self.damping = T(s.dof.damping)
    self.d_invweight = T(s.dof.invweight)
    self.solver_params = T(s.dof.solver_params)
    self.limit = None if s.dof.limit is None else (T(s.dof.limit[0]), T(s.dof.limit[1]))
    self.dof_link = torch.as_tensor(s.dof_link())
    self.dof_ranges = s.dof_ranges()
    self.q_ranges = s.q_ranges()
    self.pairs = s.contact_pairs()
    self.ncon = sum(p.ncon for p in self.pairs)

  # -- kinematics.forward (kinematics.py:31-108) ------------------------------

  def _jcalc(self, l, q, qd):
    """kinematics.py:47-83 for one link; q (N, qw), qd (N, dw)."""
    typ = self.sys.link_types[l]
    if typ == "f":
      return q[:, 0:3], q[:, 3:7], qd[:, 3:6], qd[:, 0:3]
    jp_, jr_, ja_, jv_ = None, None, None, None
    for k, d in enumerate(self.dof_ranges[l]):
      ang, vel = self.m_ang[d], self.m_vel[d]
      qi = q[:, k]
      rot_k = normalize(quat_rot_axis(ang.expand(q.shape[0], 3), qi))[0]
      pos_k = vel * qi[:, None]
      jda, jdv = ang * qd[:, k:k + 1], vel * qd[:, k:k + 1]
      if k == 0:
        jp_, jr_, ja_, jv_ = pos_k, rot_k, jda, jdv
      else:
        jp_, jr_ = tf_do_tf(jp_, jr_, pos_k, rot_k)
        ja_ = ja_ + rotate(jda, rot_k)
        jv_ = jv_ + rotate(jdv + cross(pos_k, jda), rot_k)
    return jp_, jr_, ja_, jv_

  def kinematics(self, q, qd):
    N, L = q.shape[0], self.L
    j_pos, j_rot, jd_ang, jd_vel = [], [], [], []
    for l in range(L):
      jp_, jr_, ja_, jv_ = self._jcalc(l, q[:, self.q_ranges[l]],
                                       qd[:, self.dof_ranges[l]])
      # kinematics.py:85-87: joint anchor offset, then link transform
      anchor_pos = rotate(self.lj_pos[l].expand(N, 3), jr_)
      jp_ = jp_ + self.lj_pos[l] - anchor_pos
      jp_, jr_ = tf_do_tf(self.lt_pos[l].expand(N, 3),
                          self.lt_rot[l].expand(N, 4), jp_, jr_)
      j_pos.append(jp_), j_rot.append(jr_), jd_ang.append(ja_), jd_vel.append(jv_)

    def world(parent, l):  # kinematics.py:89-101
      if parent is None:
        return j_pos[l], j_rot[l], rotate(jd_ang[l], j_rot[l]), jd_vel[l]
      xp, xr, xa, xv = parent
      pos, rot = tf_do_tf(xp, xr, j_pos[l], j_rot[l])
      vel = xv + cross(xa, pos - xp)
      vel = vel + rotate(jd_vel[l], xr)
      ang = xa + rotate(jd_ang[l], rot)
      return pos, rot, ang, vel

    ys = tree_forward(self.sys, None, world)
    x_pos = torch.stack([y[0] for y in ys], 1)
    x_rot = normalize(torch.stack([y[1] for y in ys], 1))[0]
    xd_ang = torch.stack([y[2] for y in ys], 1)
    xd_vel = torch.stack([y[3] for y in ys], 1)
    return x_pos, x_rot, xd_ang, xd_vel

  # -- dynamics.transform_com (dynamics.py:27-134) ---------------------------

  def transform_com(self, st: OState) -> OState:
    s, N, L = self.sys, st.q.shape[0], self.L
    xi_pos, xi_rot = tf_do_tf(st.x_pos, st.x_rot, self.li_pos, self.li_rot)
    root = []
    for i in range(L):
      r = i
      while s.link_parents[r] >= 0:
        r = s.link_parents[r]
      root.append(r)
    root_com = torch.zeros(N, L, 3, dtype=self.dtype)
    for r in set(root):
      members = [i for i in range(L) if root[i] == r]
      mxs = sum(self.li_mass[i] * xi_pos[:, i] for i in members)
      ms = sum(self.li_mass[i] for i in members)
      for i in members:
        root_com[:, i] = mxs / ms
    cinr = tf_do_inertia(xi_pos - root_com, xi_rot, self.li_i, self.li_mass)

    # dof motions in the COM frame
    parent_idx = [i if t == "f" else p for i, (t, p) in
                  enumerate(zip(s.link_types, s.link_parents))]
    px_pos = torch.cat([st.x_pos, torch.zeros(N, 1, 3, dtype=self.dtype)], 1)
    px_rot = torch.cat([st.x_rot, torch.tensor([1.0, 0, 0, 0], dtype=self.dtype).expand(N, 1, 4)], 1)
    pidx = torch.tensor([p if p >= 0 else L for p in parent_idx])
    j_pos, j_rot = tf_do_tf(px_pos[:, pidx], px_rot[:, pidx], self.lt_pos, self.lt_rot)
    j_pos, j_rot = tf_do_tf(j_pos, j_rot, self.lj_pos, self.lj_rot)

    cdof_ang = torch.zeros(N, self.nv, 3, dtype=self.dtype)
    cdof_vel = torch.zeros(N, self.nv, 3, dtype=self.dtype)
    for l in range(L):
      dr = self.dof_ranges[l]
      if s.link_types[l] == "f":  # dynamics.py:56-57
        cdof_ang[:, dr] = self.m_ang[dr]
        cdof_vel[:, dr] = self.m_vel[dr]
        continue
      # dynamics.py:59-83: stacked joints
      acc_p = torch.zeros(N, 3, dtype=self.dtype)
      acc_r = torch.tensor([1.0, 0, 0, 0], dtype=self.dtype).expand(N, 4)
      qv = st.q[:, self.q_ranges[l]]
      for k, d in enumerate(dr):
        ang, vel = self.m_ang[d].expand(N, 3), self.m_vel[d].expand(N, 3)
        a, v = tf_inv_do_motion(acc_p, acc_r, ang, vel)
        cdof_ang[:, d], cdof_vel[:, d] = a, v
        rk = normalize(quat_rot_axis(ang, qv[:, k]))[0]
        pk = vel * qv[:, k:k + 1]
        acc_p, acc_r = tf_do_tf(acc_p, acc_r, pk, rk)
    dl = self.dof_link
    cdof_ang = rotate(cdof_ang, j_rot[:, dl])  # dynamics.py:86-87
    off = root_com - j_pos  # dynamics.py:88-89
    one = torch.tensor([1.0, 0, 0, 0], dtype=self.dtype)
    cdof_ang, cdof_vel = tf_do_motion(off[:, dl], one, cdof_ang, cdof_vel)
    cq_ang = cdof_ang * st.qd[..., None]
    cq_vel = cdof_vel * st.qd[..., None]

    def cd_fn(parent, l):  # dynamics.py:93-103
      if parent is None:
        pa = torch.zeros(N, 3, dtype=self.dtype)
        pv = torch.zeros(N, 3, dtype=self.dtype)
      else:
        pa, pv = parent
      for d in self.dof_ranges[l]:
        pa = pa + cq_ang[:, d]
        pv = pv + cq_vel[:, d]
      return pa, pv

    cds = tree_forward(s, None, cd_fn)
    cd_ang = torch.stack([c[0] for c in cds], 1)
    cd_vel = torch.stack([c[1] for c in cds], 1)

    # dynamics.py:105-130: cdofd
    cdofd_ang = torch.zeros_like(cdof_ang)
    cdofd_vel = torch.zeros_like(cdof_vel)
    zeros3 = torch.zeros(N, 3, dtype=self.dtype)
    for l in range(L):
      dr = self.dof_ranges[l]
      if s.link_types[l] == "f":
        ca = cq_ang[:, dr[0:3]].sum(1)
        cv = cq_vel[:, dr[0:3]].sum(1)
        for k, d in enumerate(dr):
          if k < 3:
            continue
          a, v = motion_cross_motion(ca, cv, cdof_ang[:, d], cdof_vel[:, d])
          cdofd_ang[:, d], cdofd_vel[:, d] = a, v
        continue
      p = parent_idx[l]
      ca = cd_ang[:, p] if p >= 0 else zeros3
      cv = cd_vel[:, p] if p >= 0 else zeros3
      for k, d in enumerate(dr):
        if k > 0:
          ca = ca + cq_ang[:, dr[k - 1]]
          cv = cv + cq_vel[:, dr[k - 1]]
        a, v = motion_cross_motion(ca, cv, cdof_ang[:, d], cdof_vel[:, d])
        cdofd_ang[:, d], cdofd_vel[:, d] = a, v

    return dataclasses.replace(
        st, root_com=root_com, cinr=cinr, cd=(cd_ang, cd_vel),
        cdof=(cdof_ang, cdof_vel), cdofd=(cdofd_ang, cdofd_vel))

  # -- dynamics.inverse / _passive / forward (dynamics.py:137-236) -----------

  def inverse_dynamics(self, st: OState):
    s, N = self.sys, st.q.shape[0]
    cdofd_ang, cdofd_vel = st.cdofd

    def cdd_fn(parent, l):  # dynamics.py:153-161
      if parent is None:
        pa = torch.zeros(N, 3, dtype=self.dtype)
        pv = (-self.gravity).expand(N, 3)
      else:
        pa, pv = parent
      for d in self.dof_ranges[l]:
        pa = pa + cdofd_ang[:, d] * st.qd[:, d:d + 1]
        pv = pv + cdofd_vel[:, d] * st.qd[:, d:d + 1]
      return pa, pv

    cdds = tree_forward(s, None, cdd_fn)
    cdd_ang = torch.stack([c[0] for c in cdds], 1)
    cdd_vel = torch.stack([c[1] for c in cdds], 1)
    c, _, i, mass = st.cinr
    cd_ang, cd_vel = st.cd
    fa1, fv1 = inertia_mul(c, i, mass, cdd_ang, cdd_vel)  # dynamics.py:168-171
    fa2, fv2 = inertia_mul(c, i, mass, cd_ang, cd_vel)
    fa3, fv3 = motion_cross_force(cd_ang, cd_vel, fa2, fv2)
    own = [(fa1[:, l] + fa3[:, l], fv1[:, l] + fv3[:, l]) for l in range(self.L)]
    cfrc = tree_reverse_sum(s, own)  # dynamics.py:174-179
    cf_ang = torch.stack([f[0] for f in cfrc], 1)[:, self.dof_link]
    cf_vel = torch.stack([f[1] for f in cfrc], 1)[:, self.dof_link]
    cdof_ang, cdof_vel = st.cdof
    return dot(cdof_vel, cf_vel) + dot(cdof_ang, cf_ang)  # dynamics.py:182

  def passive(self, st: OState):
    """dynamics.py:187-196 (fluid branch off)."""
    frc = torch.zeros_like(st.qd)
    for l, t in enumerate(self.sys.link_types):
      if t in "fb":
        continue
      dr, qr = self.dof_ranges[l], self.q_ranges[l]
      frc[:, dr] = -st.q[:, qr] * self.stiffness[dr]
    return frc - self.damping * st.qd

  def to_tau(self, act, q, qd):
    """actuator.py:23-57."""
    a = self.sys.actuator
    N = q.shape[0]
    if self.sys.act_size() == 0:
      return torch.zeros(N, self.nv, dtype=self.dtype)
    T = lambda x: _t(x, self.dtype)  # noqa: E731
    ctrl, frc = T(a.ctrl_range), T(a.force_range)
    qa, qda = q[:, a.q_id], qd[:, a.qd_id]
    act = torch.minimum(torch.maximum(act, ctrl[:, 0]), ctrl[:, 1])
    bias = T(a.gear) * (qa * T(a.bias_q) + qda * T(a.bias_qd))
    force = T(a.gain) * act + bias
    force = torch.minimum(torch.maximum(force, frc[:, 0]), frc[:, 1])
    force = force * T(a.gear)
    tau = torch.zeros(N, self.nv, dtype=self.dtype)
    tau.index_add_(1, torch.as_tensor(a.qd_id), force)
    return tau

  # -- mass.matrix / matrix_inv (mass.py:27-106) -----------------------------

  def mass_matrix(self, st: OState):
    s = self.sys
    c, rot, i, mass = st.cinr
    own = [(c[:, l], rot[:, l], i[:, l], mass[:, l] if mass.dim() > 1 else mass[l].expand(c.shape[0]))
           for l in range(self.L)]
    crb = tree_reverse_sum(s, own)  # mass.py:44-49
    cdof_ang, cdof_vel = st.cdof
    N, nv = st.q.shape[0], self.nv
    mx = torch.zeros(N, nv, nv, dtype=self.dtype)
    for di in range(nv):
      cc, _, ii, mm = crb[int(self.dof_link[di])]
      fa, fv = inertia_mul(cc, ii, mm, cdof_ang[:, di], cdof_vel[:, di])
      mx[:, di, :] = dot(cdof_vel, fv[:, None]) + dot(cdof_ang, fa[:, None])
    # mass.py:64-74: mask to (dof_i, dof_j) with link(j) an ancestor-or-self
    mask = torch.zeros(nv, nv, dtype=self.dtype)
    for li in range(self.L):
      lj = li
      while lj > -1:
        for a in self.dof_ranges[li]:
          for b in self.dof_ranges[lj]:
            mask[a, b] = 1.0
        lj = s.link_parents[lj]
    mx = mx * mask
    mx = torch.tril(mx) + torch.tril(mx, -1).transpose(-1, -2)  # mass.py:78
    return mx + torch.diag(self.armature)

  def matrix_inv(self, st: OState, num_iter: int) -> OState:
    mx = self.mass_matrix(st)
    if num_iter > 0 and self.inverse == "ns":
      inv = inv_approximate(mx, st.mass_mx_inv, num_iter)
    else:
      inv = torch.cholesky_inverse(torch.linalg.cholesky(mx))
    return dataclasses.replace(st, mass_mx=mx, mass_mx_inv=inv)

  # -- contact.get + MJX plane collisions (contact.py:28-67) -----------------

  def collide(self, x_pos, x_rot):
    """Contacts of every collision pair, static count, in pair order."""
    s, N = self.sys, x_pos.shape[0]
    T = lambda v: _t(v, self.dtype)  # noqa: E731
    xp = torch.cat([x_pos, torch.zeros(N, 1, 3, dtype=self.dtype)], 1)
    xr = torch.cat([x_rot, torch.tensor([1.0, 0, 0, 0], dtype=self.dtype).expand(N, 1, 4)], 1)

    def geom_pose(g):  # contact.py:42-53
      b = int(s.geom_bodyid[g]) - 1
      b = b if b >= 0 else self.L
      pos = xp[:, b] + rotate(T(s.geom_pos[g]).expand(N, 3), xr[:, b])
      mat = quat_to_3x3(quat_mul(xr[:, b], T(s.geom_quat[g]).expand(N, 4)))
      return pos, mat

    out = {k: [] for k in ("dist", "pos", "frame", "friction", "solref",
                           "solimp", "link_a", "link_b", "elasticity")}
    def add(p, dist, pos, frame):
      out["dist"].append(dist)
      out["pos"].append(pos)
      out["frame"].append(frame)
      out["friction"].append(T(p.friction).expand(N, 5))
      out["solref"].append(T(p.solref).expand(N, 2))
      out["solimp"].append(T(p.solimp).expand(N, 5))
      out["link_a"].append(int(s.geom_bodyid[p.geom1]) - 1)
      out["link_b"].append(int(s.geom_bodyid[p.geom2]) - 1)
      out["elasticity"].append(0.5 * (float(s.elasticity[p.geom1]) +
                                      float(s.elasticity[p.geom2])))

    for p in self.pairs:
      ppos, pmat = geom_pose(p.geom1)
      gpos, gmat = geom_pose(p.geom2)
      if p.kind >= 2:  # sphere and capsule pairs: one contact each
        r1 = float(s.geom_size[p.geom1][0])
        r2 = float(s.geom_size[p.geom2][0])
        if p.kind == 2:  # sphere-sphere
          add(p, *sphere_sphere(ppos, r1, gpos, r2))
        elif p.kind == 3:  # sphere (geom1) - capsule (geom2)
          seg = gmat[..., :, 2] * float(s.geom_size[p.geom2][1])
          pt = closest_segment_point(gpos - seg, gpos + seg, ppos)
          add(p, *sphere_sphere(ppos, r1, pt, r2))
        else:  # capsule-capsule
          s1 = pmat[..., :, 2] * float(s.geom_size[p.geom1][1])
          s2 = gmat[..., :, 2] * float(s.geom_size[p.geom2][1])
          pt1, pt2 = closest_segment_to_segment_points(ppos - s1, ppos + s1, gpos - s2, gpos + s2)
          add(p, *sphere_sphere(pt1, r1, pt2, r2))
        continue
      n = pmat[..., :, 2]  # plane normal: z column
      r = float(s.geom_size[p.geom2][0])
      if p.kind == 0:  # plane-sphere: frame from orthogonals of n
        b1, b2 = orthogonals(normalize(n)[0])
        frame = torch.stack([n, b1, b2], -2)
        centers = [gpos]
      else:  # plane-capsule: tangent aligned with the capsule axis
        axis = gmat[..., :, 2]
        bb, bn = normalize(axis - n * dot(n, axis)[..., None])
        yv = torch.tensor([0.0, 1.0, 0.0], dtype=self.dtype).expand(N, 3)
        zv = torch.tensor([0.0, 0.0, 1.0], dtype=self.dtype).expand(N, 3)
        alt = torch.where(((-0.5 < n[..., 1]) & (n[..., 1] < 0.5))[..., None], yv, zv)
        bb = torch.where((bn < 0.5)[..., None], alt, bb)
        frame = torch.stack([n, bb, cross(n, bb)], -2)
        seg = axis * float(s.geom_size[p.geom2][1])
        centers = [gpos + seg, gpos - seg]
      for ctr in centers:
        dist = dot(ctr - ppos, n) - r
        pos = ctr - n * (r + 0.5 * dist)[..., None]
        add(p, dist, pos, frame)
    if not out["dist"]:
      return None
    res = {k: torch.stack(v, 1) for k, v in out.items()
           if k not in ("link_a", "link_b", "elasticity")}
    for k in ("link_a", "link_b", "elasticity"):
      res[k] = out[k]
    return res

  # -- constraint.jacobian (constraint.py:29-191) ----------------------------

  @staticmethod
  def imp_aref(params, pos, vel):
    """constraint.py:29-65."""
    timeconst, dampratio, dmin, dmax, width, mid, power = [params[..., k] for k in range(7)]
    imp_x = pos.abs() / width
    imp_a = (1.0 / torch.pow(mid, power - 1)) * torch.pow(imp_x, power)
    imp_b = 1 - (1.0 / torch.pow(1 - mid, power - 1)) * torch.pow(1 - imp_x, power)
    imp_y = torch.where(imp_x < mid, imp_a, imp_b)
    imp = dmin + imp_y * (dmax - dmin)
    imp = torch.minimum(torch.maximum(imp, dmin), dmax)
    imp = torch.where(imp_x > 1.0, dmax, imp)
    b = 2 / (dmax * timeconst)
    k = 1 / (dmax * dmax * timeconst * timeconst * dampratio * dampratio)
    stiffness, damping = params[..., 0], params[..., 1]
    b = torch.where(damping <= 0, -damping / dmax, b)
    k = torch.where(stiffness <= 0, -stiffness / (dmax * dmax), k)
    aref = -b * vel - k * imp * pos
    return imp, aref

  def link_mask(self, link_idx: int):
    """constraint.py:88-96: 1 for link_idx and its ancestors, 0 otherwise."""
    m = torch.zeros(self.L, dtype=self.dtype)
    l = link_idx
    while l >= 0:
      m[l] = 1.0
      l = self.sys.link_parents[l]
    return m

  def point_jacobian(self, st: OState, pos, link_idx: int):
    """constraint.py:68-98 -> (ang, vel) each (N, nv, 3)."""
    cdof_ang, cdof_vel = st.cdof
    m = self.link_mask(link_idx)[self.dof_link][None, :, None]
    ca, cv = cdof_ang * m, cdof_vel * m
    com = st.root_com[:, link_idx if link_idx >= 0 else self.L - 1]
    off = (pos - com)[:, None, :]
    one = torch.tensor([1.0, 0, 0, 0], dtype=self.dtype)
    return tf_do_motion(off, one, ca, cv)

  def jac_contact(self, st: OState):
    """constraint.py:135-176."""
    c = self.collide(st.x_pos, st.x_rot)
    N = st.q.shape[0]
    if c is None:
      z = torch.zeros(N, 0, dtype=self.dtype)
      return torch.zeros(N, 0, self.nv, dtype=self.dtype), z, z, None
    jacs, diags, arefs = [], [], []
    for k in range(c["dist"].shape[1]):
      la, lb = c["link_a"][k], c["link_b"][k]
      pos = c["pos"][:, k]
      _, av = self.point_jacobian(st, pos, la)
      _, bv = self.point_jacobian(st, pos, lb)
      diff = bv - av  # (N, nv, 3)
      frame = c["frame"][:, k]
      mu = c["friction"][:, k, 0]
      rows = []
      for d in (-frame[:, 1], -frame[:, 2]):
        for f in (-mu, mu):
          dirv = d * f[:, None] + frame[:, 0]
          rows.append((diff @ dirv[..., None])[..., 0])
      jac = torch.stack(rows, 1)  # (N, 4, nv)
      dist4 = c["dist"][:, k, None].expand(N, 4)
      params = torch.cat([c["solref"][:, k], c["solimp"][:, k]], -1)[:, None, :].expand(N, 4, 7)
      vel = (jac @ st.qd[..., None])[..., 0]
      imp, aref = self.imp_aref(params, dist4, vel)
      t = self.l_invweight[la] * float(la > -1) + self.l_invweight[lb]
      diag = (t + mu * mu * t)[:, None].expand(N, 4)
      diag = diag * 2 * (mu * mu)[:, None] * (1 - imp) / (imp + 1e-8)
      act = (c["dist"][:, k] < 0).to(self.dtype)
      jacs.append(jac * act[:, None, None])
      diags.append(diag * act[:, None])
      arefs.append(aref * act[:, None])
    return (torch.cat(jacs, 1), torch.cat(diags, 1), torch.cat(arefs, 1), c)

  def jac_limit(self, st: OState):
    """constraint.py:101-132."""
    N = st.q.shape[0]
    if self.limit is None:
      z = torch.zeros(N, 0, dtype=self.dtype)
      return torch.zeros(N, 0, self.nv, dtype=self.dtype), z, z
    q_idx = torch.as_tensor(self.sys.q_idx("123"))
    qd_idx = torch.as_tensor(self.sys.qd_idx("123"))
    pos_min = st.q[:, q_idx] - self.limit[0][qd_idx]
    pos_max = self.limit[1][qd_idx] - st.q[:, q_idx]
    pos = torch.minimum(torch.minimum(pos_min, pos_max), torch.zeros_like(pos_min))
    side = ((pos_min < pos_max).to(self.dtype) * 2 - 1) * (pos < 0).to(self.dtype)
    jac = torch.eye(self.nv, dtype=self.dtype)[qd_idx][None] * side[..., None]
    params = self.solver_params[qd_idx].expand(N, -1, 7)
    vel = (jac @ st.qd[..., None])[..., 0]
    imp, aref = self.imp_aref(params, pos, vel)
    act = (pos < 0).to(self.dtype)
    diag = self.d_invweight[qd_idx] * act * (1 - imp) / (imp + 1e-8)
    aref = aref * act
    return jac, diag, aref

  def jacobian(self, st: OState) -> OState:
    """constraint.py:179-191: contact rows, then limit rows."""
    jc, dc, ac, contact = self.jac_contact(st)
    jl, dl, al = self.jac_limit(st)
    return dataclasses.replace(
        st, con_jac=torch.cat([jc, jl], 1), con_diag=torch.cat([dc, dl], 1),
        con_aref=torch.cat([ac, al], 1), contact=contact)

  # -- constraint.force (constraint.py:194-240) ------------------------------

  def pg_solve(self, a, b, maxiter: int, maxls: int, tol: float = 1e-3):
    """[3P] jaxopt.ProjectedGradient(fun=0.5|Ax+b|^2, relu, maxiter, maxls).

    FISTA with backtracking line search, from x0 = y0 = 0, t = 1, stepsize 1
    (jaxopt defaults: acceleration=True, decrease_factor=0.5, tol=1e-3).  The
    first update always runs; later updates run while error > tol.  A
    line-search trial x+ = relu(y - eta g) is rejected while
    eta (f(x+) - f(y)) > eta <x+ - y, g> + 0.5 |x+ - y|^2 + eps, for at most
    maxls halvings.  Then eta = 1 if eta <= 1e-6 else eta / 0.5.
    """
    N, R = b.shape
    at = a.transpose(-1, -2)
    bmv = lambda m, v: (m @ v[..., None])[..., 0]  # noqa: E731

    def fun(x):
      r = bmv(a, x) + b
      return 0.5 * (r * r).sum(-1), r

    def grad(x):
      f, r = fun(x)
      return f, bmv(at, r)

    relu = torch.relu
    x = torch.zeros(N, R, dtype=self.dtype)
    y = torch.zeros_like(x)
    t = torch.ones(N, dtype=self.dtype)
    eta = torch.ones(N, dtype=self.dtype)
    error = torch.full((N,), float("inf"), dtype=self.dtype)
    running = torch.ones(N, dtype=torch.bool)
    n_iter = torch.zeros(N, dtype=torch.int64)
    n_trials = torch.zeros(N, dtype=torch.int64)
    # the discrete decisions of every update (whether it ran, its line-search
    # trials, the relu pattern of its iterate): the loss is smooth in the
    # inputs wherever none of them changes
    decisions = []

    def ls_cond(xp, eta_, f_y, g):
      diff = xp - y
      f_x, _ = fun(xp)
      lhs = eta_ * (f_x - f_y)
      rhs = eta_ * dot(diff, g) + 0.5 * dot(diff, diff) + self.ls_eps
      return lhs > rhs

    for it in range(maxiter):
      if it > 0:
        running = running & (error > tol)
      if not bool(running.any()):
        break
      f_y, g = grad(y)
      xp = relu(y - eta[:, None] * g)
      cond = ls_cond(xp, eta, f_y, g)
      trials = torch.ones(N, dtype=torch.int64)
      for _ in range(maxls):
        if not bool(cond.any()):
          break
        eta_n = torch.where(cond, eta * 0.5, eta)
        xp_n = relu(y - eta_n[:, None] * g)
        xp = torch.where(cond[:, None], xp_n, xp)
        trials = trials + cond.to(torch.int64)
        eta = eta_n
        cond = cond & ls_cond(xp, eta, f_y, g)
      eta_next = torch.where(eta <= 1e-6, torch.ones_like(eta), eta / 0.5)
      t_next = 0.5 * (1 + torch.sqrt(1 + 4 * t * t))
      y_next = xp + ((t - 1) / t_next)[:, None] * (xp - x)
      _, gx = grad(xp)
      err_next = torch.linalg.norm(relu(xp - gx) - xp, dim=-1)
      rm = running[:, None]
      x = torch.where(rm, xp, x)
      y = torch.where(rm, y_next, y)
      t = torch.where(running, t_next, t)
      eta = torch.where(running, eta_next, eta)
      error = torch.where(running, err_next, error)
      n_iter = n_iter + running.to(torch.int64)
      n_trials = n_trials + torch.where(running, trials, torch.zeros_like(trials))
      decisions.append(torch.cat([running[:, None].to(torch.int64),
                                  torch.where(running, trials, 0)[:, None],
                                  (rm & (xp > 0)).to(torch.int64)], 1).detach())
    return x, {"iters": n_iter, "trials": n_trials, "error": error,
               "decisions": decisions}

  def constraint_force(self, st: OState):
    N = st.q.shape[0]
    if st.con_jac.shape[1] == 0:
      return torch.zeros(N, self.nv, dtype=self.dtype), None
    J, X = st.con_jac, st.mass_mx_inv
    a = J @ X @ J.transpose(-1, -2) + torch.diag_embed(st.con_diag)
    b = ((J @ X) @ st.qf_smooth[..., None])[..., 0] - st.con_aref
    x, info = self.pg_solve(a, b, self.sys.solver_iterations,
                            self.sys.solver_maxls)
    info["force"] = x
    info["a"], info["b"] = a, b
    return (J.transpose(-1, -2) @ x[..., None])[..., 0], info

  # -- integrator.integrate (integrator.py:26-84) ----------------------------

  def integrate(self, st: OState) -> OState:
    dt = self.dt
    if self.sys.matrix_inv_iterations == 0:
      mx = st.mass_mx + torch.diag(self.damping) * dt
      mx_inv = torch.cholesky_inverse(torch.linalg.cholesky(mx))
    else:
      mx_inv = st.mass_mx_inv
    qdd = (mx_inv @ (st.qf_smooth + st.qf_constraint)[..., None])[..., 0]
    qd = st.qd + qdd * dt
    q = st.q.clone()
    for l, t in enumerate(self.sys.link_types):
      qr, dr = self.q_ranges[l], self.dof_ranges[l]
      if t == "f":
        rot, ang = st.q[:, qr[3:7]], qd[:, dr[3:6]]
        ang_norm = torch.linalg.norm(ang, dim=-1) + 1e-8
        axis = ang / ang_norm[:, None]
        angle = dt * ang_norm
        qrot = quat_rot_axis(axis, angle)
        rot = quat_mul(rot, qrot)
        rot = rot / torch.linalg.norm(rot, dim=-1, keepdim=True)
        pos = st.q[:, qr[0:3]] + qd[:, dr[0:3]] * dt
        q[:, qr] = torch.cat([pos, rot], -1)
      else:
        q[:, qr] = st.q[:, qr] + qd[:, dr] * dt
    return dataclasses.replace(st, q=q, qd=qd, qdd=qdd)

  # -- pipeline (pipeline.py:32-94) --------------------------------------------

  def _as(self, v):
    return torch.as_tensor(np.asarray(v) if not torch.is_tensor(v) else v,
                           dtype=self.dtype)

  def init(self, q, qd) -> OState:
    q, qd = self._as(q), self._as(qd)
    x_pos, x_rot, xd_ang, xd_vel = self.kinematics(q, qd)
    N = q.shape[0]
    st = OState(q=q, qd=qd, x_pos=x_pos, x_rot=x_rot, xd_ang=xd_ang,
                xd_vel=xd_vel,
                mass_mx_inv=torch.zeros(N, self.nv, self.nv, dtype=self.dtype))
    st = self.transform_com(st)
    st = self.matrix_inv(st, 0)
    return self.jacobian(st)

  def step(self, st: OState, act) -> OState:
    N = st.q.shape[0]
    if act is None:
      act = torch.zeros(N, self.sys.act_size(), dtype=self.dtype)
    act = self._as(act)
    tau = self.to_tau(act, st.q, st.qd)
    qf_smooth = self.passive(st) - self.inverse_dynamics(st) + tau
    st = dataclasses.replace(st, qf_smooth=qf_smooth)
    qf_c, info = self.constraint_force(st)
    st = dataclasses.replace(st, qf_constraint=qf_c, pg_info=info)
    st = self.integrate(st)
    x_pos, x_rot, xd_ang, xd_vel = self.kinematics(st.q, st.qd)
    st = dataclasses.replace(st, x_pos=x_pos, x_rot=x_rot, xd_ang=xd_ang,
                             xd_vel=xd_vel)
    st = self.transform_com(st)
    st = self.matrix_inv(st, self.sys.matrix_inv_iterations)
    return self.jacobian(st)

  def env_step(self, st: OState, act, n_frames: int) -> OState:
    """envs/base.py:128-137: n_frames substeps with one action."""
    for _ in range(n_frames):
      st = self.step(st, act)
    return st
