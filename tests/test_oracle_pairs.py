"""The oracle's sphere and capsule pair collisions (MJX narrow phase, reached
through contact.py:57), pinned by the reference's sphere-sphere KAT
(contact_test.py:67-75), its world-geom pair count (contact_test.py:116-120),
closed forms, brute-force segment distances and finite differences of the
contact point's relative velocity."""

import numpy as np
import pytest
import torch

from brax_amd import systems as S
import oracle.generalized as G
from test_oracle_invariants import _advance, _unit_q


def _contacts(sys, q):
  o = G.Oracle(sys)
  q = torch.as_tensor(np.atleast_2d(q), dtype=torch.float64)
  x_pos, x_rot, _, _ = o.kinematics(q, torch.zeros(q.shape[0], sys.qd_size(), dtype=torch.float64))
  return o.collide(x_pos, x_rot)


def test_sphere_sphere_kat():
  """contact_test.py:67-75: dist -0.01, pos (0, 0.195, 0), normal +y,
  elasticity 0.15, to 6 decimals."""
  sys = S.two_spheres()
  c = _contacts(sys, sys.init_q)
  assert c["dist"].shape[1] == 1
  np.testing.assert_array_almost_equal(c["dist"][0, 0], -0.01)
  np.testing.assert_array_almost_equal(c["pos"][0, 0], [0.0, 0.195, 0.0])
  np.testing.assert_array_almost_equal(c["frame"][0, 0, 0], [0.0, 1.0, 0.0])
  np.testing.assert_array_almost_equal(c["elasticity"][0], 0.15)
  assert (c["link_a"][0], c["link_b"][0]) == (0, 1)


def test_no_world_self_collision():
  """contact_test.py:116-120: a ball among twelve capsules fixed to the world
  over a floor has 13 contacts (ball-floor and ball-capsule); the world's
  geoms never pair with each other.  The reference's fixture moves the ball
  on two slide joints; the count depends only on the geoms, so the ball
  here is free."""
  b = S.ModelBuilder("arena", timestep=0.01, iterations=4, gravity=(0, 0, 0))
  kw = dict(friction=(0.1, 0.1, 0.1))
  b.add_geom(0, S.GEOM_PLANE, (1, 1, 10), **kw)
  walls = [(-.75, -.65, .75, -.65), (-.75, -.65, -.75, -.2), (-.75, .65, -.75, .2),
           (-.75, .65, .75, .65), (.75, -.65, .75, -.2), (.75, .65, .75, .2),
           (.85, -.2, .85, .2), (-.85, -.2, -.85, .2), (-.75, -.2, -.85, -.2),
           (-.75, .2, -.85, .2), (.75, -.2, .85, -.2), (.75, .2, .85, .2)]
  for x0, y0, x1, y1 in walls:
    b.add_geom(0, S.GEOM_CAPSULE, (0.03,), fromto=(x0, y0, 0.01, x1, y1, 0.02135), **kw)
  ball = b.add_body("ball", 0, pos=(0, 0, 0.02135))
  b.add_free_joint(ball)
  b.add_geom(ball, S.GEOM_SPHERE, (0.02135,), mass=0.046, **kw)
  sys = b.build()
  c = _contacts(sys, sys.init_q)
  assert c["pos"].shape[1] == 13
  assert sum(1 for a, bb in zip(c["link_a"], c["link_b"]) if a < 0 and bb < 0) == 0


def test_pile_pairs_and_roles():
  """Every pair kind occurs, geom1 has the lower type, and a world geom can
  sit on either side (link -1)."""
  sys = S.pile()
  kinds = [p.kind for p in sys.contact_pairs()]
  assert sorted(set(kinds)) == [0, 1, 2, 3, 4]
  c = _contacts(sys, sys.init_q)
  assert c["dist"].shape[1] == sys.num_contacts() == 16
  la, lb = c["link_a"], c["link_b"]
  assert any(a < 0 <= bb for a, bb in zip(la, lb))
  assert any(bb < 0 <= a for a, bb in zip(la, lb))


def test_sphere_capsule_closed_form():
  """A sphere beside the middle of a capsule, and one past its end."""
  c = torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float64)
  ax = torch.tensor([[1.0, 0.0, 0.0]], dtype=torch.float64)
  for p, pt in (([0.3, 0.2, 0.0], [0.3, 0.0, 0.0]), ([1.5, 0.0, 0.1], [1.0, 0.0, 0.0])):
    ps = torch.tensor([p], dtype=torch.float64)
    got = G.closest_segment_point(c - ax, c + ax, ps)
    np.testing.assert_allclose(got[0], pt, atol=2e-6)  # 1e-6 regularized projection
    dist, pos, frame = G.sphere_sphere(ps, 0.15, got, 0.1)
    g = got[0].numpy()
    d = np.linalg.norm(np.array(p) - g)
    np.testing.assert_allclose(dist[0], d - 0.25, atol=1e-12)
    n = (g - np.array(p)) / d
    np.testing.assert_allclose(frame[0, 0], n, atol=1e-12)
    np.testing.assert_allclose(pos[0], np.array(p) + n * (0.15 + 0.5 * (d - 0.25)), atol=1e-12)
    f = frame[0].numpy()
    np.testing.assert_allclose(f @ f.T, np.eye(3), atol=1e-12)
    assert np.linalg.det(f) == pytest.approx(1.0)


def _seg_point_dist(a, b, p):
  ab = b - a
  t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0, 1)
  return np.linalg.norm(p - (a + t * ab))


def _seg_seg_exact(a0, a1, b0, b1):
  """Minimum distance of two segments: the lines' interior critical point
  when it lies on both segments, else the best endpoint-to-segment
  distance."""
  u, v, w = a1 - a0, b1 - b0, a0 - b0
  A, B, C, D, E = u @ u, u @ v, v @ v, u @ w, v @ w
  den = A * C - B * B
  best = min(_seg_point_dist(b0, b1, a0), _seg_point_dist(b0, b1, a1),
             _seg_point_dist(a0, a1, b0), _seg_point_dist(a0, a1, b1))
  if den > 1e-12:
    s, t = (B * E - C * D) / den, (A * E - B * D) / den
    if 0 <= s <= 1 and 0 <= t <= 1:
      best = min(best, np.linalg.norm(w + s * u - t * v))
  return best


def test_segment_closest_points_against_brute_force():
  """MJX's segment-segment closest points lie on the segments, are never
  nearer than the true minimum, and equal it wherever the clipped line
  solution or its one refinement is the minimum (here: all but a few
  percent of random pairs, and every pair whose unclipped line solution is
  interior)."""
  rng = np.random.default_rng(0)
  n = 2000
  a0, a1, b0, b1 = (rng.uniform(-1, 1, (n, 3)) for _ in range(4))
  T = lambda x: torch.tensor(x, dtype=torch.float64)  # noqa: E731
  pa, pb = G.closest_segment_to_segment_points(T(a0), T(a1), T(b0), T(b1))
  pa, pb = pa.numpy(), pb.numpy()
  got = np.linalg.norm(pa - pb, axis=1)
  exact = np.array([_seg_seg_exact(*x) for x in zip(a0, a1, b0, b1)])
  on_a = np.array([_seg_point_dist(x, y, p) for x, y, p in zip(a0, a1, pa)])
  on_b = np.array([_seg_point_dist(x, y, p) for x, y, p in zip(b0, b1, pb)])
  assert on_a.max() < 1e-5 and on_b.max() < 1e-5
  assert (got >= exact - 1e-6).all()
  close = np.abs(got - exact) < 1e-5
  assert close.mean() > 0.97, close.mean()
  # interior line solutions are exact
  u, v, w = a1 - a0, b1 - b0, a0 - b0
  A, B, C = (u * u).sum(1), (u * v).sum(1), (v * v).sum(1)
  D, E = (u * w).sum(1), (v * w).sum(1)
  den = A * C - B * B
  s, t = (B * E - C * D) / den, (A * E - B * D) / den
  interior = (s > 0.01) & (s < 0.99) & (t > 0.01) & (t < 0.99)
  assert interior.sum() > 100
  np.testing.assert_allclose(got[interior], exact[interior], atol=1e-5)


def test_pair_jacobian_rows_are_relative_point_velocities():
  """Each contact row . qd = dir . (velocity of the contact point as a point
  of geom2's link - as a point of geom1's link), by FD; a world side moves
  with nothing."""
  sys = S.pile()
  rng = np.random.default_rng(7)
  q = _unit_q(sys, rng, 1, noise=0.05)[0]
  for l in range(4):  # squeeze the bodies together so pairs overlap
    r = sys.q_ranges()[l]
    q[r[0:3]] = 0.55 * q[r[0:3]] + np.array([0.0, -0.05, 0.0])
  qd = rng.standard_normal(sys.qd_size())
  o = G.Oracle(sys)
  st = o.init(q[None], qd[None])
  c = st.contact
  jac = st.con_jac[0].numpy()
  pos0, rot0, _, _, _ = S.body_frames(sys, q)
  eps = 1e-6
  qp, qm = _advance(sys, q, qd, eps), _advance(sys, q, qd, -eps)
  posp, rotp, _, _, _ = S.body_frames(sys, qp)
  posm, rotm, _, _, _ = S.body_frames(sys, qm)

  def vel(l, p):
    if l < 0:
      return np.zeros(3)
    local = S.quat_to_mat(rot0[l]).T @ (p - pos0[l])
    return ((posp[l] + S.quat_to_mat(rotp[l]) @ local) -
            (posm[l] + S.quat_to_mat(rotm[l]) @ local)) / (2 * eps)

  kinds = []
  pair_of = [p for p in sys.contact_pairs() for _ in range(p.ncon)]
  for k in range(c["dist"].shape[1]):
    if c["dist"][0, k] >= 0:
      assert np.all(jac[4 * k:4 * k + 4] == 0)
      continue
    kinds.append(pair_of[k].kind)
    p = c["pos"][0, k].numpy()
    v = vel(c["link_b"][k], p) - vel(c["link_a"][k], p)
    fr = c["frame"][0, k].numpy()
    mu = c["friction"][0, k, 0].item()
    dirs = [fr[0] + mu * fr[1], fr[0] - mu * fr[1], fr[0] + mu * fr[2], fr[0] - mu * fr[2]]
    for r in range(4):
      assert jac[4 * k + r] @ qd == pytest.approx(dirs[r] @ v, rel=1e-6, abs=1e-7)
  assert {2, 3, 4} <= set(kinds), kinds
