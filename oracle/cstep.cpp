// C++ restatement of the oracle's generalized step, for the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY: the same algorithm as oracle/generalized.py
// (the reference's brax/generalized pipeline with the exact mass-matrix
// inverse, jaxopt's ProjectedGradient and the MJX plane collisions restated
// there), one env at a time, OpenMP over envs, in float or double.  Only the
// tests and bench.py's CPU-baseline leg load it.  Its double build is pinned
// against oracle/generalized.py in float64 (tests/test_oracle_cstep.py).
//
// Differences from oracle/generalized.py that change results only by
// rounding: M^-1 products are Cholesky solves instead of an explicit
// inverse, and the solver works on the active constraint rows only (the
// rows of separated contacts and inactive limits are zero, so they leave
// every projected-gradient iterate unchanged).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int MAXL = 64, MAXV = 64, MAXR = 512;

// flat System, filled by oracle/cstep.py from the System fields (all double)
struct Sys {
  int L, nq, nv, nu, nc, has_limit, mat_inv_iters, solver_iters, solver_maxls;
  double dt, ls_eps, gravity[3];
  const int32_t* ltype;   // [L] 0 free, k = k stacked hinges
  const int32_t* parent;  // [L]
  const int32_t* qadr;    // [L]
  const int32_t* dadr;    // [L]
  const double* lt;       // [L][7] link transform pos, rot
  const double* lj;       // [L][7] joint transform
  const double* li;       // [L][7] inertia transform
  const double* inertia;  // [L][9]
  const double* mass;     // [L]
  const double* l_invw;   // [L]
  const double* motion;   // [nv][6] ang, vel
  const double* dofp;     // [nv][6] armature, stiffness, damping, invweight, lo, hi
  const double* solp;     // [nv][7]
  const int32_t* act_q;   // [nu]
  const int32_t* act_d;   // [nu]
  const double* actp;     // [nu][8] ctrl lo, hi, force lo, hi, gain, gear, bias_q, bias_qd
  // contacts, one record each: kind (0 sphere, 1 capsule), end sign,
  // plane body link, plane pos[3] quat[4], geom link, pos[3], quat[4],
  // radius, half length, mu, solref[2], solimp[5], link_a, link_b
  const double* con;      // [nc][CON]
  // capsule tangent frame: 0 = the axis projected on the plane (the
  // restatement's reading of MJX), 1 = orthogonals(n) as for a sphere (an
  // alternative reading, for sensitivity studies of this unpinned detail)
  int capsule_frame;
};
enum { C_KIND, C_SIGN, C_PL, C_PP, C_PQ = C_PP + 3, C_GL = C_PQ + 4, C_GP, C_GQ = C_GP + 3,
       C_R = C_GQ + 4, C_HL, C_MU, C_SOLREF, C_SOLIMP = C_SOLREF + 2, C_LA = C_SOLIMP + 5,
       C_LB, CON };

template <class R> struct V3 { R x, y, z; };
template <class R> struct Q4 { R w, x, y, z; };
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> V3<R> ld3(const double* p) { return {(R)p[0], (R)p[1], (R)p[2]}; }
template <class R> Q4<R> ld4(const double* p) { return {(R)p[0], (R)p[1], (R)p[2], (R)p[3]}; }

// math.py:25-40
template <class R> V3<R> rotate(V3<R> v, Q4<R> q) {
  const V3<R> u{q.x, q.y, q.z};
  const V3<R> r = u * (R(2) * dot(u, v)) + v * (q.w * q.w - dot(u, u));
  return r + cross(u, v) * (R(2) * q.w);
}
// math.py:86-101
template <class R> Q4<R> qmul(Q4<R> u, Q4<R> v) {
  return {u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z,
          u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y,
          u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x,
          u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w};
}
template <class R> Q4<R> qinv(Q4<R> q) { return {q.w, -q.x, -q.y, -q.z}; }
// math.py:134-148
template <class R> Q4<R> qaxis(V3<R> a, R angle) {
  const R s = std::sin(angle / 2);
  return {std::cos(angle / 2), a.x * s, a.y * s, a.z * s};
}
// math.py:151-165, rows of the rotation matrix
template <class R> void q3x3(Q4<R> q, R m[3][3]) {
  const R d = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z, s = R(2) / d;
  const R xs = q.x * s, ys = q.y * s, zs = q.z * s;
  const R wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
  const R xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
  const R yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
  m[0][0] = 1 - (yy + zz), m[0][1] = xy - wz, m[0][2] = xz + wy;
  m[1][0] = xy + wz, m[1][1] = 1 - (xx + zz), m[1][2] = yz - wx;
  m[2][0] = xz - wy, m[2][1] = yz + wx, m[2][2] = 1 - (xx + yy);
}
// math.py:308-345 (safe_norm / normalize over one vector)
template <class R> R safe_norm(const R* x, int n) {
  bool zero = true;
  for (int k = 0; k < n; ++k) zero = zero && std::fabs(x[k]) <= R(1e-8);
  if (zero) return R(0);
  R s = 0;
  for (int k = 0; k < n; ++k) s += x[k] * x[k];
  return std::sqrt(s);
}
template <class R> V3<R> normalize3(V3<R> v, R* norm = nullptr) {
  R a[3] = {v.x, v.y, v.z};
  const R n = safe_norm(a, 3);
  if (norm) *norm = n;
  const R d = n + (n == R(0) ? R(1e-6) : R(0));
  return {v.x / d, v.y / d, v.z / d};
}
template <class R> Q4<R> normalize4(Q4<R> q) {
  R a[4] = {q.w, q.x, q.y, q.z};
  const R n = safe_norm(a, 4);
  const R d = n + (n == R(0) ? R(1e-6) : R(0));
  return {q.w / d, q.x / d, q.y / d, q.z / d};
}
// base.py:559-579
template <class R> void tf_do_tf(V3<R> p, Q4<R> r, V3<R> p2, Q4<R> r2, V3<R>& po, Q4<R>& ro) {
  po = p + rotate(p2, r);
  ro = qmul(r, r2);
}
// base.py:604-614
template <class R> void cross_motion(V3<R> sa, V3<R> sv, V3<R> ma, V3<R> mv, V3<R>& oa,
                                     V3<R>& ov) {
  oa = cross(sa, ma);
  ov = cross(sa, mv) + cross(sv, ma);
}
// base.py:297-302 with c = mass-weighted COM offset
template <class R> void inertia_mul(V3<R> c, const R i[9], R m, V3<R> a, V3<R> v, V3<R>& fa,
                                    V3<R>& fv) {
  fa = V3<R>{i[0] * a.x + i[1] * a.y + i[2] * a.z, i[3] * a.x + i[4] * a.y + i[5] * a.z,
             i[6] * a.x + i[7] * a.y + i[8] * a.z} + cross(c, v);
  fv = v * m - cross(c, a);
}
// constraint.py:29-65
template <class R> void imp_aref(const double* sr, const double* si, R pos, R vel, R& imp,
                                 R& aref) {
  const R timeconst = (R)sr[0], dampratio = (R)sr[1];
  const R dmin = (R)si[0], dmax = (R)si[1], width = (R)si[2], mid = (R)si[3], power = (R)si[4];
  const R x = std::fabs(pos) / width;
  const R ia = (R(1) / std::pow(mid, power - 1)) * std::pow(x, power);
  const R ib = R(1) - (R(1) / std::pow(R(1) - mid, power - 1)) * std::pow(R(1) - x, power);
  const R y = x < mid ? ia : ib;
  imp = dmin + y * (dmax - dmin);
  imp = imp < dmin ? dmin : (imp > dmax ? dmax : imp);
  if (x > R(1)) imp = dmax;
  R b = R(2) / (dmax * timeconst);
  R k = R(1) / (dmax * dmax * timeconst * timeconst * dampratio * dampratio);
  if (sr[1] <= 0) b = -(R)sr[1] / dmax;
  if (sr[0] <= 0) k = -(R)sr[0] / (dmax * dmax);
  aref = -b * vel - k * imp * pos;
}

// in-place Cholesky M = C C^T (lower), then solves
template <class R> void chol(R* a, int n) {
  for (int j = 0; j < n; ++j) {
    R s = a[j * n + j];
    for (int k = 0; k < j; ++k) s -= a[j * n + k] * a[j * n + k];
    a[j * n + j] = std::sqrt(s);
    for (int i = j + 1; i < n; ++i) {
      R t = a[i * n + j];
      for (int k = 0; k < j; ++k) t -= a[i * n + k] * a[j * n + k];
      a[i * n + j] = t / a[j * n + j];
    }
  }
}
template <class R> void chol_solve(const R* c, int n, R* x) {
  for (int i = 0; i < n; ++i) {
    R s = x[i];
    for (int k = 0; k < i; ++k) s -= c[i * n + k] * x[k];
    x[i] = s / c[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    R s = x[i];
    for (int k = i + 1; k < n; ++k) s -= c[k * n + i] * x[k];
    x[i] = s / c[i * n + i];
  }
}

template <class R>
struct Work {
  V3<R> xp[MAXL], xa[MAXL], xv[MAXL];
  Q4<R> xr[MAXL];
  V3<R> rc[MAXL], cc[MAXL], cda[MAXL], cdv[MAXL];
  R ci[MAXL][9], cm[MAXL];
  V3<R> ca[MAXV], cv[MAXV], cdda[MAXV], cddv[MAXV];
  R M[MAXV * MAXV], C[MAXV * MAXV];
  R J[MAXR * MAXV], dg[MAXR], ar[MAXR];
  R A[MAXR * MAXR], W[MAXV * MAXR];
  R qf[MAXV], qfc[MAXV];
};

// kinematics.py:31-108
template <class R> void kinematics(const Sys& s, const R* q, const R* qd, Work<R>& w) {
  for (int l = 0; l < s.L; ++l) {
    const int qa = s.qadr[l], da = s.dadr[l];
    V3<R> jp{0, 0, 0}, ja{0, 0, 0}, jv{0, 0, 0};
    Q4<R> jr{1, 0, 0, 0};
    if (s.ltype[l] == 0) {
      jp = {q[qa], q[qa + 1], q[qa + 2]};
      jr = {q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]};
      ja = {qd[da + 3], qd[da + 4], qd[da + 5]};
      jv = {qd[da], qd[da + 1], qd[da + 2]};
    } else {
      for (int k = 0; k < s.ltype[l]; ++k) {
        const int d = da + k;
        const V3<R> ang = ld3<R>(s.motion + 6 * d), vel = ld3<R>(s.motion + 6 * d + 3);
        const R qi = q[qa + k];
        const Q4<R> rk = normalize4(qaxis(ang, qi));
        const V3<R> pk = vel * qi, jda = ang * qd[d], jdv = vel * qd[d];
        if (k == 0) {
          jp = pk, jr = rk, ja = jda, jv = jdv;
        } else {
          tf_do_tf(jp, jr, pk, rk, jp, jr);
          ja = ja + rotate(jda, rk);
          jv = jv + rotate(jdv + cross(pk, jda), rk);
        }
      }
    }
    const V3<R> ljp = ld3<R>(s.lj + 7 * l);
    jp = jp + ljp - rotate(ljp, jr);
    tf_do_tf(ld3<R>(s.lt + 7 * l), ld4<R>(s.lt + 7 * l + 3), jp, jr, jp, jr);
    const int p = s.parent[l];
    if (p < 0) {
      w.xp[l] = jp, w.xr[l] = jr, w.xa[l] = rotate(ja, jr), w.xv[l] = jv;
    } else {
      V3<R> pos;
      Q4<R> rot;
      tf_do_tf(w.xp[p], w.xr[p], jp, jr, pos, rot);
      w.xv[l] = w.xv[p] + cross(w.xa[p], pos - w.xp[p]) + rotate(jv, w.xr[p]);
      w.xa[l] = w.xa[p] + rotate(ja, rot);
      w.xp[l] = pos, w.xr[l] = rot;
    }
  }
  for (int l = 0; l < s.L; ++l) w.xr[l] = normalize4(w.xr[l]);
}

// dynamics.py:27-134
template <class R> void transform_com(const Sys& s, const R* q, const R* qd, Work<R>& w) {
  const int L = s.L;
  V3<R> xip[MAXL];
  Q4<R> xir[MAXL];
  int root[MAXL];
  for (int l = 0; l < L; ++l) {
    tf_do_tf(w.xp[l], w.xr[l], ld3<R>(s.li + 7 * l), ld4<R>(s.li + 7 * l + 3), xip[l], xir[l]);
    int r = l;
    while (s.parent[r] >= 0) r = s.parent[r];
    root[l] = r;
  }
  for (int l = 0; l < L; ++l) {
    V3<R> mx{0, 0, 0};
    R ms = 0;
    for (int i = 0; i < L; ++i)
      if (root[i] == root[l]) mx = mx + xip[i] * (R)s.mass[i], ms += (R)s.mass[i];
    w.rc[l] = mx * (R(1) / ms);
  }
  // cinr = (xi - root_com, xi_rot).do(inertia)
  for (int l = 0; l < L; ++l) {
    const V3<R> p = xip[l] - w.rc[l];
    const R m = (R)s.mass[l];
    R r[3][3];
    q3x3(xir[l], r);
    R ri[3][3] = {};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int k = 0; k < 3; ++k) ri[a][b] += r[a][k] * (R)s.inertia[9 * l + 3 * k + b];
    const R pp[3] = {p.x, p.y, p.z};
    const R p2 = dot(p, p);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        R v = 0;
        for (int k = 0; k < 3; ++k) v += ri[a][k] * r[b][k];
        w.ci[l][3 * a + b] = v + ((a == b ? p2 : R(0)) - pp[a] * pp[b]) * m;
      }
    w.cc[l] = p * m;
    w.cm[l] = m;
  }
  // cdof: dof motions about the root COM in world axes
  for (int l = 0; l < L; ++l) {
    const int pidx = s.ltype[l] == 0 ? l : s.parent[l];
    V3<R> jpos = pidx >= 0 ? w.xp[pidx] : V3<R>{0, 0, 0};
    Q4<R> jrot = pidx >= 0 ? w.xr[pidx] : Q4<R>{1, 0, 0, 0};
    tf_do_tf(jpos, jrot, ld3<R>(s.lt + 7 * l), ld4<R>(s.lt + 7 * l + 3), jpos, jrot);
    tf_do_tf(jpos, jrot, ld3<R>(s.lj + 7 * l), ld4<R>(s.lj + 7 * l + 3), jpos, jrot);
    const int da = s.dadr[l], qa = s.qadr[l];
    const int nd = s.ltype[l] == 0 ? 6 : s.ltype[l];
    V3<R> ap{0, 0, 0};
    Q4<R> arot{1, 0, 0, 0};
    for (int k = 0; k < nd; ++k) {
      const int d = da + k;
      const V3<R> ang = ld3<R>(s.motion + 6 * d), vel = ld3<R>(s.motion + 6 * d + 3);
      V3<R> a = ang, v = vel;
      if (s.ltype[l] != 0) {  // stacked hinges (dynamics.py:59-83)
        a = rotate(ang, arot);
        v = rotate(vel, arot) + cross(ap, a);
        const R qi = q[qa + k];
        tf_do_tf(ap, arot, vel * qi, normalize4(qaxis(ang, qi)), ap, arot);
      }
      a = rotate(a, jrot);  // dynamics.py:86-89
      const V3<R> off = w.rc[l] - jpos;
      w.ca[d] = a;
      w.cv[d] = v - cross(off, a);
    }
  }
  // cd (dynamics.py:93-103) and cdofd (105-130)
  for (int l = 0; l < L; ++l) {
    const int p = s.parent[l], da = s.dadr[l];
    const int nd = s.ltype[l] == 0 ? 6 : s.ltype[l];
    V3<R> a = p >= 0 ? w.cda[p] : V3<R>{0, 0, 0}, v = p >= 0 ? w.cdv[p] : V3<R>{0, 0, 0};
    for (int k = 0; k < nd; ++k) a = a + w.ca[da + k] * qd[da + k], v = v + w.cv[da + k] * qd[da + k];
    w.cda[l] = a, w.cdv[l] = v;
  }
  for (int l = 0; l < L; ++l) {
    const int da = s.dadr[l];
    if (s.ltype[l] == 0) {
      V3<R> sa{0, 0, 0}, sv{0, 0, 0};
      for (int k = 0; k < 3; ++k) sa = sa + w.ca[da + k] * qd[da + k], sv = sv + w.cv[da + k] * qd[da + k];
      for (int k = 0; k < 6; ++k) {
        if (k < 3) {
          w.cdda[da + k] = {0, 0, 0}, w.cddv[da + k] = {0, 0, 0};
        } else {
          cross_motion(sa, sv, w.ca[da + k], w.cv[da + k], w.cdda[da + k], w.cddv[da + k]);
        }
      }
      continue;
    }
    const int p = s.parent[l];
    V3<R> sa = p >= 0 ? w.cda[p] : V3<R>{0, 0, 0}, sv = p >= 0 ? w.cdv[p] : V3<R>{0, 0, 0};
    for (int k = 0; k < s.ltype[l]; ++k) {
      const int d = da + k;
      if (k > 0) sa = sa + w.ca[d - 1] * qd[d - 1], sv = sv + w.cv[d - 1] * qd[d - 1];
      cross_motion(sa, sv, w.ca[d], w.cv[d], w.cdda[d], w.cddv[d]);
    }
  }
}

// dynamics.py:137-236, actuator.py:23-57: qf_smooth = passive - bias + tau
template <class R> void smooth_force(const Sys& s, const R* q, const R* qd, const R* act,
                                     Work<R>& w) {
  const int L = s.L;
  V3<R> aa[MAXL], av[MAXL], fa[MAXL], fv[MAXL];
  const V3<R> g{(R)s.gravity[0], (R)s.gravity[1], (R)s.gravity[2]};
  for (int l = 0; l < L; ++l) {
    const int p = s.parent[l], da = s.dadr[l];
    const int nd = s.ltype[l] == 0 ? 6 : s.ltype[l];
    V3<R> a = p >= 0 ? aa[p] : V3<R>{0, 0, 0};
    V3<R> v = p >= 0 ? av[p] : V3<R>{-g.x, -g.y, -g.z};
    for (int k = 0; k < nd; ++k) a = a + w.cdda[da + k] * qd[da + k], v = v + w.cddv[da + k] * qd[da + k];
    aa[l] = a, av[l] = v;
    V3<R> f1a, f1v, f2a, f2v;
    inertia_mul(w.cc[l], w.ci[l], w.cm[l], a, v, f1a, f1v);
    inertia_mul(w.cc[l], w.ci[l], w.cm[l], w.cda[l], w.cdv[l], f2a, f2v);
    // motion cross force (base.py:611-614)
    fa[l] = f1a + cross(w.cda[l], f2a) + cross(w.cdv[l], f2v);
    fv[l] = f1v + cross(w.cda[l], f2v);
  }
  for (int l = L - 1; l >= 0; --l) {
    const int p = s.parent[l];
    if (p >= 0) fa[p] = fa[p] + fa[l], fv[p] = fv[p] + fv[l];
  }
  for (int l = 0; l < L; ++l) {
    const int da = s.dadr[l], qa = s.qadr[l];
    const int nd = s.ltype[l] == 0 ? 6 : s.ltype[l];
    for (int k = 0; k < nd; ++k) {
      const int d = da + k;
      const R bias = dot(w.cv[d], fv[l]) + dot(w.ca[d], fa[l]);
      R pas = s.ltype[l] == 0 ? R(0) : -q[qa + k] * (R)s.dofp[6 * d + 1];
      pas -= (R)s.dofp[6 * d + 2] * qd[d];
      w.qf[d] = pas - bias;
    }
  }
  for (int u = 0; u < s.nu; ++u) {
    const double* ap = s.actp + 8 * u;
    R a = act ? act[u] : R(0);
    a = a < (R)ap[0] ? (R)ap[0] : (a > (R)ap[1] ? (R)ap[1] : a);
    const R bias = (R)ap[5] * (q[s.act_q[u]] * (R)ap[6] + qd[s.act_d[u]] * (R)ap[7]);
    R f = (R)ap[4] * a + bias;
    f = f < (R)ap[2] ? (R)ap[2] : (f > (R)ap[3] ? (R)ap[3] : f);
    w.qf[s.act_d[u]] += f * (R)ap[5];
  }
}

// mass.py:27-83: CRBA, masked to the dof-ancestor pattern, + armature
template <class R> void mass_matrix(const Sys& s, Work<R>& w) {
  const int L = s.L, nv = s.nv;
  V3<R> bc[MAXL];
  R bi[MAXL][9], bm[MAXL];
  for (int l = 0; l < L; ++l) {
    bc[l] = w.cc[l], bm[l] = w.cm[l];
    for (int k = 0; k < 9; ++k) bi[l][k] = w.ci[l][k];
  }
  for (int l = L - 1; l >= 0; --l) {
    const int p = s.parent[l];
    if (p >= 0) {
      bc[p] = bc[p] + bc[l], bm[p] += bm[l];
      for (int k = 0; k < 9; ++k) bi[p][k] += bi[l][k];
    }
  }
  int dlink[MAXV];
  for (int l = 0; l < L; ++l) {
    const int nd = s.ltype[l] == 0 ? 6 : s.ltype[l];
    for (int k = 0; k < nd; ++k) dlink[s.dadr[l] + k] = l;
  }
  for (int i = 0; i < nv; ++i) {
    const int li = dlink[i];
    V3<R> fa, fv;
    inertia_mul(bc[li], bi[li], bm[li], w.ca[i], w.cv[i], fa, fv);
    for (int j = 0; j <= i; ++j) {
      bool anc = false;  // link(j) an ancestor-or-self of link(i)
      for (int a = li; a >= 0; a = s.parent[a]) anc = anc || a == dlink[j];
      const R v = anc ? dot(w.cv[j], fv) + dot(w.ca[j], fa) : R(0);
      w.M[i * nv + j] = v;
      w.M[j * nv + i] = v;
    }
    w.M[i * nv + i] += (R)s.dofp[6 * i];
  }
}

// contact.py:28-67 + MJX plane collisions, constraint.py:68-191
template <class R> int jacobian(const Sys& s, const R* q, const R* qd, Work<R>& w) {
  const int nv = s.nv;
  int dlink[MAXV];
  for (int l = 0; l < s.L; ++l) {
    const int nd = s.ltype[l] == 0 ? 6 : s.ltype[l];
    for (int k = 0; k < nd; ++k) dlink[s.dadr[l] + k] = l;
  }
  auto pose = [&](int link, const double* gp, const double* gq, V3<R>& pos, R m[3][3]) {
    const V3<R> xp = link >= 0 ? w.xp[link] : V3<R>{0, 0, 0};
    const Q4<R> xr = link >= 0 ? w.xr[link] : Q4<R>{1, 0, 0, 0};
    pos = xp + rotate(ld3<R>(gp), xr);
    q3x3(qmul(xr, ld4<R>(gq)), m);
  };
  // velocity part of the point Jacobian of dof d at pos for a link's tree
  auto pjac = [&](int link, int d, V3<R> pos) {
    bool anc = false;
    for (int a = link; a >= 0; a = s.parent[a]) anc = anc || a == dlink[d];
    if (!anc) return V3<R>{0, 0, 0};
    const V3<R> off = pos - w.rc[link];
    return w.cv[d] - cross(off, w.ca[d]);
  };
  int r = 0;
  for (int c = 0; c < s.nc; ++c) {
    const double* cr = s.con + CON * c;
    V3<R> ppos, gpos;
    R pm[3][3], gm[3][3];
    pose((int)cr[C_PL], cr + C_PP, cr + C_PQ, ppos, pm);
    pose((int)cr[C_GL], cr + C_GP, cr + C_GQ, gpos, gm);
    const V3<R> n{pm[0][2], pm[1][2], pm[2][2]};
    const R rad = (R)cr[C_R];
    V3<R> f1, f2, ctr = gpos;
    if (cr[C_KIND] == 0) {  // sphere: frame from orthogonals(normalize(n))
      const V3<R> a = normalize3(n);
      V3<R> b = (R(-0.5) < a.y && a.y < R(0.5)) ? V3<R>{0, 1, 0} : V3<R>{0, 0, 1};
      b = b - a * dot(a, b);
      const bool nz = a.x != 0 || a.y != 0 || a.z != 0;
      b = normalize3(b) * (nz ? R(1) : R(0));
      f1 = b, f2 = cross(a, b);
    } else {  // capsule: tangent along the axis projected on the plane
      const V3<R> axis{gm[0][2], gm[1][2], gm[2][2]};
      R bn;
      V3<R> bb = normalize3(axis - n * dot(n, axis), &bn);
      if (bn < R(0.5)) bb = (R(-0.5) < n.y && n.y < R(0.5)) ? V3<R>{0, 1, 0} : V3<R>{0, 0, 1};
      if (s.capsule_frame == 1) {  // alternative: orthogonals(normalize(n))
        const V3<R> a = normalize3(n);
        V3<R> b = (R(-0.5) < a.y && a.y < R(0.5)) ? V3<R>{0, 1, 0} : V3<R>{0, 0, 1};
        bb = normalize3(b - a * dot(a, b));
      }
      f1 = bb, f2 = cross(n, bb);
      ctr = gpos + axis * ((R)cr[C_HL] * (R)cr[C_SIGN]);
    }
    const R dist = dot(ctr - ppos, n) - rad;
    if (!(dist < 0)) continue;  // inactive rows are zero
    const V3<R> pos = ctr - n * (rad + R(0.5) * dist);
    const int la = (int)cr[C_LA], lb = (int)cr[C_LB];
    const R mu = (R)cr[C_MU];
    R t = (R)s.l_invw[lb] + (la > -1 ? (R)s.l_invw[la] : R(0));
    for (int k = 0; k < 4; ++k) {
      const V3<R> dir = (k < 2 ? f1 : f2) * ((k % 2 == 0) ? mu : -mu) + n;
      R vel = 0;
      for (int d = 0; d < nv; ++d) {
        const V3<R> bv = pjac(lb, d, pos);
        const V3<R> av = la >= 0 ? pjac(la, d, pos) : V3<R>{0, 0, 0};
        const R j = dot(bv - av, dir);
        w.J[r * nv + d] = j;
        vel += j * qd[d];
      }
      R imp, aref;
      imp_aref(cr + C_SOLREF, cr + C_SOLIMP, dist, vel, imp, aref);
      w.dg[r] = (t + mu * mu * t) * R(2) * (mu * mu) * (R(1) - imp) / (imp + R(1e-8));
      w.ar[r] = aref;
      ++r;
    }
  }
  if (s.has_limit) {
    for (int l = 0; l < s.L; ++l) {
      if (s.ltype[l] == 0) continue;
      for (int k = 0; k < s.ltype[l]; ++k) {
        const int d = s.dadr[l] + k, qi = s.qadr[l] + k;
        const R pmin = q[qi] - (R)s.dofp[6 * d + 4], pmax = (R)s.dofp[6 * d + 5] - q[qi];
        R pos = pmin < pmax ? pmin : pmax;
        pos = pos < R(0) ? pos : R(0);
        if (!(pos < 0)) continue;
        const R side = pmin < pmax ? R(1) : R(-1);
        for (int j = 0; j < nv; ++j) w.J[r * nv + j] = j == d ? side : R(0);
        R imp, aref;
        imp_aref(s.solp + 7 * d, s.solp + 7 * d + 2, pos, side * qd[d], imp, aref);
        w.dg[r] = (R)s.dofp[6 * d + 3] * (R(1) - imp) / (imp + R(1e-8));
        w.ar[r] = aref;
        ++r;
      }
    }
  }
  return r;
}

// [3P] jaxopt.ProjectedGradient on 0.5 |A x + b|^2, x >= 0 (see the oracle)
template <class R> void pg_solve(const R* A, const R* b, int n, int maxiter, int maxls, R ls_eps,
                                 R* x) {
  R y[MAXR], xp[MAXR], g[MAXR], r[MAXR], tmp[MAXR];
  auto fun = [&](const R* v, R* res) {
    R f = 0;
    for (int i = 0; i < n; ++i) {
      R s = b[i];
      for (int j = 0; j < n; ++j) s += A[i * n + j] * v[j];
      res[i] = s;
      f += s * s;
    }
    return R(0.5) * f;
  };
  auto grad = [&](const R* v, R* gr) {
    const R f = fun(v, r);
    for (int j = 0; j < n; ++j) {
      R s = 0;
      for (int i = 0; i < n; ++i) s += A[i * n + j] * r[i];
      gr[j] = s;
    }
    return f;
  };
  auto ls_cond = [&](R eta, R f_y) {
    const R fx = fun(xp, tmp);
    R dg = 0, dd = 0;
    for (int i = 0; i < n; ++i) {
      const R d = xp[i] - y[i];
      dg += d * g[i], dd += d * d;
    }
    return eta * (fx - f_y) > eta * dg + R(0.5) * dd + ls_eps;
  };
  for (int i = 0; i < n; ++i) x[i] = y[i] = 0;
  R t = 1, eta = 1, error = INFINITY;
  for (int it = 0; it < maxiter; ++it) {
    if (it > 0 && !(error > R(1e-3))) break;
    const R f_y = grad(y, g);
    for (int i = 0; i < n; ++i) xp[i] = std::fmax(y[i] - eta * g[i], R(0));
    bool cond = ls_cond(eta, f_y);
    for (int ls = 0; ls < maxls && cond; ++ls) {
      eta = eta * R(0.5);
      for (int i = 0; i < n; ++i) xp[i] = std::fmax(y[i] - eta * g[i], R(0));
      cond = ls_cond(eta, f_y);
    }
    const R eta_next = eta <= R(1e-6) ? R(1) : eta / R(0.5);
    const R t_next = R(0.5) * (R(1) + std::sqrt(R(1) + R(4) * t * t));
    R gx[MAXR];
    grad(xp, gx);
    R e2 = 0;
    for (int i = 0; i < n; ++i) {
      const R d = std::fmax(xp[i] - gx[i], R(0)) - xp[i];
      e2 += d * d;
    }
    for (int i = 0; i < n; ++i) {
      y[i] = xp[i] + ((t - R(1)) / t_next) * (xp[i] - x[i]);
      x[i] = xp[i];
    }
    t = t_next, eta = eta_next, error = std::sqrt(e2);
  }
}

// one substep of one env (pipeline.py:64-94 with the exact inverse);
// returns the number of active constraint rows
template <class R> int substep(const Sys& s, R* q, R* qd, const R* act, Work<R>& w) {
  const int nv = s.nv;
  kinematics(s, q, qd, w);
  transform_com(s, q, qd, w);
  mass_matrix(s, w);
  const int nr = jacobian(s, q, qd, w);
  smooth_force(s, q, qd, act, w);
  // the matrix the integrator inverts (integrator.py:55-62)
  for (int k = 0; k < nv * nv; ++k) w.C[k] = w.M[k];
  if (s.mat_inv_iters == 0)
    for (int d = 0; d < nv; ++d) w.C[d * nv + d] += (R)s.dofp[6 * d + 2] * (R)s.dt;
  R Cm[MAXV * MAXV];
  for (int k = 0; k < nv * nv; ++k) Cm[k] = w.M[k];
  chol(Cm, nv);
  for (int k = 0; k < nv; ++k) w.qfc[k] = 0;
  if (nr > 0) {
    // W = M^-1 J^T, A = J W + diag, b = J M^-1 qf - aref (constraint.py:194-240)
    for (int i = 0; i < nr; ++i) {
      R col[MAXV];
      for (int d = 0; d < nv; ++d) col[d] = w.J[i * nv + d];
      chol_solve(Cm, nv, col);
      for (int d = 0; d < nv; ++d) w.W[d * nr + i] = col[d];
    }
    R u[MAXV];
    for (int d = 0; d < nv; ++d) u[d] = w.qf[d];
    chol_solve(Cm, nv, u);
    R b[MAXR], x[MAXR];
    for (int i = 0; i < nr; ++i) {
      R bi = -w.ar[i];
      for (int d = 0; d < nv; ++d) bi += w.J[i * nv + d] * u[d];
      b[i] = bi;
      for (int j = 0; j < nr; ++j) {
        R a = 0;
        for (int d = 0; d < nv; ++d) a += w.J[i * nv + d] * w.W[d * nr + j];
        w.A[i * nr + j] = a + (i == j ? w.dg[i] : R(0));
      }
    }
    pg_solve(w.A, b, nr, s.solver_iters, s.solver_maxls, (R)s.ls_eps, x);
    for (int d = 0; d < nv; ++d) {
      R f = 0;
      for (int i = 0; i < nr; ++i) f += w.J[i * nv + d] * x[i];
      w.qfc[d] = f;
    }
  }
  // integrator.py:26-84
  R qdd[MAXV];
  for (int d = 0; d < nv; ++d) qdd[d] = w.qf[d] + w.qfc[d];
  if (s.mat_inv_iters == 0) chol(w.C, nv), chol_solve(w.C, nv, qdd);
  else chol_solve(Cm, nv, qdd);
  for (int d = 0; d < nv; ++d) qd[d] += qdd[d] * (R)s.dt;
  const R dt = (R)s.dt;
  for (int l = 0; l < s.L; ++l) {
    const int qa = s.qadr[l], da = s.dadr[l];
    if (s.ltype[l] == 0) {
      const V3<R> ang{qd[da + 3], qd[da + 4], qd[da + 5]};
      const R an = std::sqrt(dot(ang, ang)) + R(1e-8);
      const Q4<R> qr = qaxis(ang * (R(1) / an), dt * an);
      Q4<R> rot = qmul(Q4<R>{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]}, qr);
      const R rn = std::sqrt(rot.w * rot.w + rot.x * rot.x + rot.y * rot.y + rot.z * rot.z);
      rot = {rot.w / rn, rot.x / rn, rot.y / rn, rot.z / rn};
      for (int k = 0; k < 3; ++k) q[qa + k] += qd[da + k] * dt;
      q[qa + 3] = rot.w, q[qa + 4] = rot.x, q[qa + 5] = rot.y, q[qa + 6] = rot.z;
    } else {
      for (int k = 0; k < s.ltype[l]; ++k) q[qa + k] += qd[da + k] * dt;
    }
  }
  return nr;
}

template <class R>
int run(const Sys* s, int n, int n_steps, int n_frames, R* q, R* qd, const R* act,
        double* rows, int threads) {
  if (s->L > MAXL || s->nv > MAXV || 4 * s->nc + s->nv > MAXR) return 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
  for (int e = 0; e < n; ++e) {
    static thread_local std::vector<char> buf(sizeof(Work<R>));
    Work<R>& w = *reinterpret_cast<Work<R>*>(buf.data());
    double nr = 0;
    for (int t = 0; t < n_steps; ++t) {
      const R* a = act ? act + ((long long)t * n + e) * s->nu : nullptr;
      for (int f = 0; f < n_frames; ++f)
        nr += substep(*s, q + (long long)e * s->nq, qd + (long long)e * s->nv, a, w);
    }
    if (rows) rows[e] = nr;
  }
  return 0;
}

}  // namespace

extern "C" {

int ocs_sizeof_sys(void) { return (int)sizeof(Sys); }
int ocs_con_record(void) { return CON; }

// n envs, n_steps env steps of n_frames substeps; act (n_steps, n, nu) or
// NULL; q (n, nq), qd (n, nv) updated in place; rows (n) or NULL receives
// each env's active constraint rows summed over its substeps
int ocs_rollout_f32(const Sys* s, int n, int n_steps, int n_frames, float* q, float* qd,
                    const float* act, double* rows, int threads) {
  return run<float>(s, n, n_steps, n_frames, q, qd, act, rows, threads);
}
int ocs_rollout_f64(const Sys* s, int n, int n_steps, int n_frames, double* q, double* qd,
                    const double* act, double* rows, int threads) {
  return run<double>(s, n, n_steps, n_frames, q, qd, act, rows, threads);
}

}  // extern "C"
