// One generalized-coordinates physics substep for one environment, per
// thread.  Restates brax/generalized/pipeline.py:64-94 (and init, :32-61).
//
// The topology T is a constexpr struct (see topologies.inc): every loop over
// links, dofs, limit rows and contact candidates is fully unrolled, so the
// per-env arrays below live in registers with static indices.  Numbers come
// from the float32 model table M (uniform across the wave -> scalar loads).
//
// Result-preserving restructurings relative to the reference (each changes
// results only by rounding):
//  * the step's position-level terms (cinr, cdof, M, contact Jacobians) are
//    computed at the start of the substep from (q, qd) instead of at the end
//    of the previous one; the reference computes the same values from the
//    same (q, qd) (pipeline.py:84-89, used at :79-82 of the next call);
//  * M^-1 comes from an exact sparse factorization M = L^T L (Featherstone's
//    LTL, no fill-in for tree sparsity) instead of a warm-started
//    Newton-Schulz iterate (mass.py:100-102), which converges to the same
//    inverse;
//  * constraint rows whose contact has dist >= 0 are exactly zero in the
//    reference (constraint.py:174): zero rows and columns of A and zero b
//    entries leave the projected-gradient iterates unchanged, so only the
//    active contacts are kept, compacted into slots;
//  * A = J M^-1 J^T + diag(d) is never formed: products A v are evaluated
//    through the tree (J^T v as spatial forces accumulated to the root,
//    M^-1 by the LTL solve, J z as contact-point velocities).
#pragma once

#include <type_traits>

#include "bgx_layout.h"
#include "bgx_math.h"

namespace bgx {

constexpr int WAVE = 64;

// per-env status bits
constexpr int ST_SLOW_PATH = 1;  // more active contacts than fast slots
constexpr int ST_NONFINITE = 2;  // non-finite q/qd after the call

BGX_D V3 ldv3(const float* __restrict__ p) { return V3{p[0], p[1], p[2]}; }
BGX_D Q4 ldq4(const float* __restrict__ p) { return Q4{p[0], p[1], p[2], p[3]}; }
BGX_D Sym3 ldsym(const float* __restrict__ p) {
  return Sym3{p[0], p[1], p[2], p[3], p[4], p[5]};
}
BGX_D float clipf(float x, float lo, float hi) {
  // jp.clip(x, lo, hi) = minimum(maximum(x, lo), hi); NaN propagates
  x = x < lo ? lo : x;
  return x > hi ? hi : x;
}

template <class T>
BGX_D const float* link_rec(const float* __restrict__ M, int l) {
  return M + T::OFF_LINK + l * LINK_STRIDE;
}
template <class T>
BGX_D const float* dof_rec(const float* __restrict__ M, int d) {
  return M + T::OFF_DOF + d * DOF_STRIDE;
}
template <class T>
BGX_D const float* cand_rec(const float* __restrict__ M, int c) {
  return M + T::OFF_CAND + c * CAND_STRIDE;
}

BGX_D V3 v3_(float x, float y, float z) { return V3{x, y, z}; }
BGX_D V3 unit3(int k) {
  return k == 0 ? v3(1.f, 0.f, 0.f) : (k == 1 ? v3(0.f, 1.f, 0.f) : v3(0.f, 0.f, 1.f));
}

// dof motion (brax/base.py DoF.motion): a free joint's six dofs are the
// fixed unit translations then unit rotations (brax/io/mjcf.py:354-356), and
// a hinge has no translational part, so both are compile-time constants
template <class T>
BGX_D V3 dof_ang(const float* __restrict__ M, int d) {
  const int l = T::DOF_LINK[d];
  if (T::LTYPE[l] == 0) {
    const int k = d - T::DADR[l];
    return k < 3 ? v3zero() : unit3(k - 3);
  }
  return ldv3(dof_rec<T>(M, d) + DF_ANG);
}
template <class T>
BGX_D V3 dof_vel(const float* __restrict__ M, int d) {
  const int l = T::DOF_LINK[d];
  if (T::LTYPE[l] == 0) {
    const int k = d - T::DADR[l];
    return k < 3 ? unit3(k) : v3zero();
  }
  if (T::DOF_HINGE[d]) return v3zero();
  return ldv3(dof_rec<T>(M, d) + DF_VEL);
}

template <class T>
constexpr bool hosts_contact(int l) {
  for (int c = 0; c < T::NCAND; ++c)
    if (T::CAND_LINK[c] == l || T::CAND_LINKA[c] == l) return true;
  return false;
}

// ---------------------------------------------------------------------------
// kinematics.forward (kinematics.py:31-108); XD: also compute xd

template <class T, bool XD, class S>
BGX_D void forward_kinematics(const float* __restrict__ M, const S* q, const S* qd,
                              V3T<S>* xp, Q4T<S>* xr, V3T<S>* xa, V3T<S>* xv) {
  V3T<S> jp[T::L], ja[T::L], jv[T::L];
  Q4T<S> jr[T::L];
  const S zero = lift<S>(0.f);
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const float* lk = link_rec<T>(M, l);
    const int qa = T::QADR[l], da = T::DADR[l];
    V3T<S> p{zero, zero, zero}, a{zero, zero, zero}, v{zero, zero, zero};
    Q4T<S> r = lift4<S>(qident());
    if (T::LTYPE[l] == 0) {  // free joint (kinematics.py:48-51)
      p = V3T<S>{q[qa], q[qa + 1], q[qa + 2]};
      r = Q4T<S>{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]};
      if (XD) {
        a = V3T<S>{qd[da + 3], qd[da + 4], qd[da + 5]};
        v = V3T<S>{qd[da], qd[da + 1], qd[da + 2]};
      }
    } else {  // 1-3 stacked axis joints (kinematics.py:52-81)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k < T::LTYPE[l]) {
          const V3 ang = dof_ang<T>(M, da + k), vel = dof_vel<T>(M, da + k);
          const S qk = q[qa + k];
          const Q4T<S> rk = qnormalize(qrot_axis(lift3<S>(ang), qk));
          const V3T<S> pk = vel * qk;
          if (k == 0) {
            p = pk, r = rk;
          } else {
            p = p + rotate(pk, r);
            r = qmul(r, rk);
          }
          if (XD) {
            const S qdk = qd[da + k];
            const V3T<S> jda = ang * qdk, jdv = vel * qdk;
            if (k == 0) {
              a = jda, v = jdv;
            } else {
              a = a + rotate(jda, rk);
              v = v + rotate(jdv + cross(pk, jda), rk);
            }
          }
        }
      }
    }
    // joint anchor offset and link transform (kinematics.py:85-87)
    const V3T<S> jpos = lift3<S>(ldv3(lk + LK_JPOS));
    p = (p + jpos) - rotate(jpos, r);
    const V3T<S> tpos = lift3<S>(ldv3(lk + LK_TPOS));
    const Q4T<S> trot = lift4<S>(ldq4(lk + LK_TROT));
    p = tpos + rotate(p, trot);
    r = qmul(trot, r);
    jp[l] = p, jr[l] = r, ja[l] = a, jv[l] = v;
  }
  // world frames, root to leaves (kinematics.py:89-104)
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const int par = T::PARENT[l];
    if (par < 0) {
      xp[l] = jp[l];
      xr[l] = jr[l];
      if (XD) {
        xa[l] = rotate(ja[l], jr[l]);
        xv[l] = jv[l];
      }
    } else {
      xp[l] = xp[par] + rotate(jp[l], xr[par]);
      xr[l] = qmul(xr[par], jr[l]);
      if (XD) {
        V3T<S> vel = xv[par] + cross(xa[par], xp[l] - xp[par]);
        xv[l] = vel + rotate(jv[l], xr[par]);
        xa[l] = xa[par] + rotate(ja[l], xr[l]);
      }
    }
  }
#pragma unroll
  for (int l = 0; l < T::L; ++l) xr[l] = qnormalize(xr[l]);  // :106
}

// ---------------------------------------------------------------------------
// per-substep tree quantities

constexpr int popc(unsigned x) {
  int n = 0;
  while (x) {
    n += (int)(x & 1u);
    x >>= 1;
  }
  return n;
}
// index of entry (i, j), j <= i, of the lower dof-ancestor pattern: rows in
// order, in each row the ancestor columns ascending, then the diagonal
template <class T>
constexpr int mpat_row(int i) {
  int s = 0;
  for (int r = 0; r < i; ++r) s += popc(T::DANC[r]) + 1;
  return s;
}
template <class T>
struct MPat {
  struct Table {
    int v[T::NV][T::NV];
  };
  static constexpr Table make() {
    Table t{};
    for (int i = 0; i < T::NV; ++i)
      for (int j = 0; j <= i; ++j)
        t.v[i][j] = mpat_row<T>(i) +
                    (j == i ? popc(T::DANC[i]) : popc(T::DANC[i] & ((1u << j) - 1u)));
    return t;
  }
  static constexpr Table table = make();
};
// (a table lookup, so that unrolled loops fold it to a constant)
template <class T>
BGX_D constexpr int mpat_idx(int i, int j) {
  return MPat<T>::table.v[i][j];
}

template <class T>
struct Tree {
  static constexpr int NR = T::LM_LDS ? 1 : T::NV;
  V3 ca[T::NV], cv[T::NV];  // cdof (ang, vel)
  // LTL factor on the lower dof-ancestor pattern: in registers, or (LM_LDS
  // topologies) packed by mpat_idx at lml[idx * lms] (LDS or private memory)
  float Lm[NR][NR];
  float* lml = nullptr;
  int lms = WAVE;
  float inv[T::NV];         // 1 / L_kk
  V3 rc[T::L];              // root COM of each link's tree
  BGX_D float& lm(int i, int j) {
    if constexpr (T::LM_LDS) return lml[mpat_idx<T>(i, j) * lms];
    else return Lm[i][j];
  }
  BGX_D float lm(int i, int j) const {
    if constexpr (T::LM_LDS) return lml[mpat_idx<T>(i, j) * lms];
    else return Lm[i][j];
  }
};

// LDS of the step kernel per lane (floats, stride WAVE): link scratch (6 per
// link), the fast-path slot rows (4 vectors x 4 rows per slot; also a second
// link vector while the rows are unused), then for LM_LDS topologies the LTL
// factor and the 4 limit-row solver vectors
template <class T>
constexpr int lds_rows_area() {
  return 4 * 4 * T::KC > T::L * 6 ? 4 * 4 * T::KC : T::L * 6;
}
template <class T>
constexpr int lds_lm_offset() {
  return T::L * 6 + lds_rows_area<T>();
}
template <class T>
constexpr int lds_floats() {
  return lds_lm_offset<T>() + (T::LM_LDS ? mpat_row<T>(T::NV) + 4 * T::NLIM : 0);
}

// unroll factor of the dof-indexed nested loops: fully unrolled (registers,
// static indices) up to 20 dofs; above, rolled loops keep compile time and
// code size bounded (the factor then lives in scratch memory)
template <class T>
constexpr int unroll_nv() {
  return T::NV <= 24 ? 32 : 1;
}

// M = L^T L in place on the lower part of M (Featherstone's LTL: leaves
// first, no fill-in outside the dof-ancestor pattern)
template <class T>
BGX_D void ltl_factor(Tree<T>& t) {
  constexpr int U = unroll_nv<T>();
#pragma unroll U
  for (int k = T::NV - 1; k >= 0; --k) {
    const float lkk = fsqrt(t.lm(k, k));
    t.lm(k, k) = lkk;
    t.inv[k] = rcp(lkk);
#pragma unroll U
    for (int j = 0; j < k; ++j)
      if ((T::DANC[k] >> j) & 1u) t.lm(k, j) *= t.inv[k];
#pragma unroll U
    for (int i = 0; i < k; ++i) {
      if ((T::DANC[k] >> i) & 1u) {
        const float a = t.lm(k, i);
        t.lm(i, i) -= a * a;
#pragma unroll U
        for (int j = 0; j < i; ++j)
          if ((T::DANC[i] >> j) & 1u) t.lm(i, j) -= a * t.lm(k, j);
      }
    }
  }
}

// M z = tau with M = L^T L (in place on z)
template <class T>
BGX_D void ltl_solve(const Tree<T>& t, float* z) {
  constexpr int U = unroll_nv<T>();
#pragma unroll U
  for (int i = T::NV - 1; i >= 0; --i) {
    z[i] *= t.inv[i];
#pragma unroll U
    for (int j = 0; j < i; ++j)
      if ((T::DANC[i] >> j) & 1u) z[j] -= t.lm(i, j) * z[i];
  }
#pragma unroll U
  for (int i = 0; i < T::NV; ++i) {
#pragma unroll U
    for (int j = 0; j < i; ++j)
      if ((T::DANC[i] >> j) & 1u) z[i] -= t.lm(i, j) * z[j];
    z[i] *= t.inv[i];
  }
}

// dofs of link l and of its ancestors (bitmask)
template <class T>
constexpr unsigned link_dof_mask(int l) {
  const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
  const int last = T::DADR[l] + nd - 1;
  return T::DANC[last] | (1u << last);
}

// spatial velocity of every link from generalized velocities z (the forward
// scan of dynamics.py:93-103 with cdof)
template <class T, class S, class Z>
BGX_D void link_vel(const V3T<S>* ca, const V3T<S>* cv, const Z* z, V3T<S>* va,
                    V3T<S>* vv) {
  const S zero = lift<S>(0.f);
  const V3T<S> z3{zero, zero, zero};
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const int par = T::PARENT[l];
    V3T<S> a = par < 0 ? z3 : va[par];
    V3T<S> v = par < 0 ? z3 : vv[par];
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = T::DADR[l] + k;
        a = a + ca[d] * z[d];
        v = v + cv[d] * z[d];
      }
    }
    va[l] = a, vv[l] = v;
  }
}
template <class T>
BGX_D void link_velocities(const Tree<T>& t, const float* z, V3* va, V3* vv) {
  link_vel<T>(t.ca, t.cv, z, va, vv);
}

// generalized force of per-link spatial forces (fa: torque about the root
// COM, fv: force), accumulated leaf to root (dynamics.py:174-182); each
// link's dofs are projected as soon as its subtree sum is final
template <class T, class S>
BGX_D void project_forces_on(const V3T<S>* ca, const V3T<S>* cv, const V3T<S>* fa,
                             const V3T<S>* fv, S* tau) {
  const S zero = lift<S>(0.f);
  V3T<S> acc_a[T::L], acc_v[T::L];
#pragma unroll
  for (int l = 0; l < T::L; ++l)
    acc_a[l] = V3T<S>{zero, zero, zero}, acc_v[l] = V3T<S>{zero, zero, zero};
#pragma unroll
  for (int l = T::L - 1; l >= 0; --l) {
    const V3T<S> a = fa[l] + acc_a[l], v = fv[l] + acc_v[l];
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = T::DADR[l] + k;
        tau[d] = dot(cv[d], v) + dot(ca[d], a);
      }
    }
    const int par = T::PARENT[l];
    if (par >= 0) {
      acc_a[par] = acc_a[par] + a;
      acc_v[par] = acc_v[par] + v;
    }
  }
}
template <class T>
BGX_D void project_forces(const Tree<T>& t, const V3* fa, const V3* fv, float* tau) {
  project_forces_on<T>(t.ca, t.cv, fa, fv, tau);
}

// ---------------------------------------------------------------------------
// dynamics pieces, generic in the scalar (float for the step, Dual for the
// adjoint's tangent pass)

// root COM of each link's tree (dynamics.py:37-43: the mass-weighted link
// COMs of every link with the same root), into rc[l]
template <class T, class S>
BGX_D void root_coms(const float* __restrict__ M, const V3T<S>* xp, const Q4T<S>* xr,
                     V3T<S>* rc) {
  const S zero = lift<S>(0.f);
#pragma unroll
  for (int r = 0; r < T::L; ++r) {
    if (T::PARENT[r] >= 0) continue;
    V3T<S> sm{zero, zero, zero};
    float ms = 0.f;
#pragma unroll
    for (int l = 0; l < T::L; ++l) {
      if (T::ROOT[l] != r) continue;
      const float* lk = link_rec<T>(M, l);
      const float m = lk[LK_MASS];
      sm = sm + (xp[l] + rotate(lift3<S>(ldv3(lk + LK_IPOS)), xr[l])) * m;
      ms = ms + m;
    }
    const float ims = rcp(ms);
    const V3T<S> c{sm.x * ims, sm.y * ims, sm.z * ims};
#pragma unroll
    for (int l = 0; l < T::L; ++l)
      if (T::ROOT[l] == r) rc[l] = c;
  }
}

// cinr (dynamics.py:44-55): per link, the mass-weighted COM offset from its
// tree's root COM cc = m (xi - rc) and the rotational inertia about it ci
template <class T, class S>
BGX_D void com_inertia(const float* __restrict__ M, const V3T<S>* xp, const Q4T<S>* xr,
                       const V3T<S>* rc, V3T<S>* cc, Sym3T<S>* ci) {
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const float* lk = link_rec<T>(M, l);
    const V3T<S> xip = xp[l] + rotate(lift3<S>(ldv3(lk + LK_IPOS)), xr[l]);
    const Q4T<S> xir = qmul(xr[l], lift4<S>(ldq4(lk + LK_IROT)));
    const float m = lk[LK_MASS];
    const V3T<S> p = xip - rc[l];
    S R[3][3];
    quat_to_3x3(xir, R);
    const float* i0 = lk + LK_I;
    const Sym3T<S> I0{lift<S>(i0[0]), lift<S>(i0[1]), lift<S>(i0[2]),
                      lift<S>(i0[3]), lift<S>(i0[4]), lift<S>(i0[5])};
    S RI[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const V3T<S> c = symmul(I0, V3T<S>{R[i][0], R[i][1], R[i][2]});
      RI[i][0] = c.x, RI[i][1] = c.y, RI[i][2] = c.z;
    }
    auto rir = [&](int i, int j) {
      return RI[i][0] * R[j][0] + RI[i][1] * R[j][1] + RI[i][2] * R[j][2];
    };
    const S pp = dot(p, p);
    ci[l] = Sym3T<S>{rir(0, 0) + (pp - p.x * p.x) * m, rir(1, 1) + (pp - p.y * p.y) * m,
                     rir(2, 2) + (pp - p.z * p.z) * m, rir(0, 1) - p.x * p.y * m,
                     rir(0, 2) - p.x * p.z * m, rir(1, 2) - p.y * p.z * m};
    cc[l] = p * m;
  }
}

// cdof: dof motions about their tree's root COM in world axes
// (dynamics.py:56-89)
template <class T, class S>
BGX_D void dof_motions(const float* __restrict__ M, const S* q, const V3T<S>* xp,
                       const Q4T<S>* xr, const V3T<S>* rc, V3T<S>* ca, V3T<S>* cv) {
  const S zero = lift<S>(0.f);
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const float* lk = link_rec<T>(M, l);
    const int par = T::PARENT[l], da = T::DADR[l], qa = T::QADR[l];
    V3T<S> Pp{zero, zero, zero};
    Q4T<S> Pr = lift4<S>(qident());
    if (T::LTYPE[l] == 0) {
      Pp = xp[l], Pr = xr[l];
    } else if (par >= 0) {
      Pp = xp[par], Pr = xr[par];
    }
    const V3T<S> t1p = Pp + rotate(lift3<S>(ldv3(lk + LK_TPOS)), Pr);
    const Q4T<S> jr = qmul(Pr, lift4<S>(ldq4(lk + LK_TROT)));
    const V3T<S> jp = t1p + rotate(lift3<S>(ldv3(lk + LK_JPOS)), jr);
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
    V3T<S> accp{zero, zero, zero};
    Q4T<S> accr = lift4<S>(qident());
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = da + k;
        const V3 ang = dof_ang<T>(M, d), vel = dof_vel<T>(M, d);
        V3T<S> a = lift3<S>(ang), v = lift3<S>(vel);
        if (T::LTYPE[l] != 0) {  // stacked joints (dynamics.py:59-83)
          a = rotate(lift3<S>(ang), accr);
          v = rotate(lift3<S>(vel), accr) + cross(accp, a);
          const Q4T<S> rk = qnormalize(qrot_axis(lift3<S>(ang), q[qa + k]));
          accp = accp + rotate(vel * q[qa + k], accr);
          accr = qmul(accr, rk);
        }
        a = rotate(a, jr);  // dynamics.py:86-89
        ca[d] = a;
        cv[d] = v - cross(rc[l] - jp, a);
      }
    }
  }
}

// inverse dynamics bias (dynamics.py:105-184): RNEA with the root at -g and
// cdofd formed on the fly; cda, cdv are the link velocities at qd
template <class T, class S>
BGX_D void rnea_bias(const float* __restrict__ M, const V3T<S>* ca, const V3T<S>* cv,
                     const V3T<S>* cda, const V3T<S>* cdv, const V3T<S>* cc,
                     const Sym3T<S>* ci, const S* qd, S* bias) {
  const S zero = lift<S>(0.f);
  const V3T<S> z3{zero, zero, zero};
  const V3T<S> grav = lift3<S>(ldv3(M + H_GRAVITY));
  V3T<S> aa[T::L], av[T::L], fa[T::L], fv[T::L];
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const int par = T::PARENT[l], da = T::DADR[l];
    V3T<S> a = par < 0 ? z3 : aa[par];
    V3T<S> v = par < 0 ? -grav : av[par];
    if (T::LTYPE[l] == 0) {
      // free joint: cdofd = (sum of the translational cdof qd) x cdof for
      // the rotational dofs, zero for the translational ones (:107-112)
      V3T<S> sa = z3, sv = z3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sa = sa + ca[da + k] * qd[da + k];
        sv = sv + cv[da + k] * qd[da + k];
      }
#pragma unroll
      for (int k = 3; k < 6; ++k) {
        V3T<S> oa, ov;
        cross_motion(sa, sv, ca[da + k], cv[da + k], oa, ov);
        a = a + oa * qd[da + k];
        v = v + ov * qd[da + k];
      }
    } else {
      V3T<S> sa = par >= 0 ? cda[par] : z3;
      V3T<S> sv = par >= 0 ? cdv[par] : z3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k < T::LTYPE[l]) {
          const int d = da + k;
          if (k > 0) {
            sa = sa + ca[d - 1] * qd[d - 1];
            sv = sv + cv[d - 1] * qd[d - 1];
          }
          V3T<S> oa, ov;
          cross_motion(sa, sv, ca[d], cv[d], oa, ov);
          a = a + oa * qd[d];
          v = v + ov * qd[d];
        }
      }
    }
    aa[l] = a, av[l] = v;
    const float m = link_rec<T>(M, l)[LK_MASS];
    V3T<S> f1a, f1v, f2a, f2v, f3a, f3v;
    inertia_mul(cc[l], ci[l], m, a, v, f1a, f1v);
    inertia_mul(cc[l], ci[l], m, cda[l], cdv[l], f2a, f2v);
    cross_force(cda[l], cdv[l], f2a, f2v, f3a, f3v);
    fa[l] = f1a + f3a, fv[l] = f1v + f3v;
  }
  project_forces_on<T>(ca, cv, fa, fv, bias);
}

// qf_smooth = passive - bias + tau (dynamics.py:187-236, actuator.py:23-57);
// qf holds the bias on entry
template <class T, class S>
BGX_D void smooth_forces(const float* __restrict__ M, const S* q, const S* qd,
                         const float* act, S* qf) {
#pragma unroll
  for (int d = 0; d < T::NV; ++d) {
    const float* df = dof_rec<T>(M, d);
    S p = -df[DF_DAMPING] * qd[d];
    const int l = T::DOF_LINK[d];
    if (T::LTYPE[l] != 0)
      p = -q[T::QADR[l] + (d - T::DADR[l])] * df[DF_STIFFNESS] + p;
    qf[d] = p - qf[d];
  }
#pragma unroll
  for (int a = 0; a < T::NU; ++a) {
    const float* ac = M + T::OFF_ACT + a * ACT_STRIDE;
    const int qa = T::ACT_Q[a], da = T::ACT_D[a];
    const float ctl = clipf(act[a], ac[AC_CTRL_LO], ac[AC_CTRL_HI]);
    const S b = ac[AC_GEAR] * (q[qa] * ac[AC_BIAS_Q] + qd[da] * ac[AC_BIAS_QD]);
    S f = ac[AC_GAIN] * ctl + b;
    const float fv = val(f);
    if (fv < ac[AC_FORCE_LO]) f = lift<S>(ac[AC_FORCE_LO]);
    if (fv > ac[AC_FORCE_HI]) f = lift<S>(ac[AC_FORCE_HI]);
    qf[da] += f * ac[AC_GEAR];
  }
}

// mass.matrix (mass.py:27-83): composite rigid-body inertias (cc, ci turn into
// subtree sums in place), then M_ij = cdof_j . (I_c(i) cdof_i) on the
// dof-ancestor pattern, lower part, with armature on the diagonal
template <class T, class S, class Put>
BGX_D void crba(const float* __restrict__ M, V3T<S>* cc, Sym3T<S>* ci,
                const V3T<S>* ca, const V3T<S>* cv, Put put) {
  float km[T::L];
#pragma unroll
  for (int l = 0; l < T::L; ++l) km[l] = link_rec<T>(M, l)[LK_MASS];
#pragma unroll
  for (int l = T::L - 1; l >= 0; --l) {
    const int par = T::PARENT[l];
    if (par >= 0) {
      cc[par] = cc[par] + cc[l];
      ci[par] = symadd(ci[par], ci[l]);
      km[par] = km[par] + km[l];
    }
  }
  constexpr int U = unroll_nv<T>();
#pragma unroll U
  for (int i = 0; i < T::NV; ++i) {
    const int l = T::DOF_LINK[i];
    V3T<S> fa, fv;
    inertia_mul(cc[l], ci[l], km[l], ca[i], cv[i], fa, fv);
    put(i, i, dot(cv[i], fv) + dot(ca[i], fa) + dof_rec<T>(M, i)[DF_ARMATURE]);
#pragma unroll U
    for (int j = 0; j < i; ++j)
      if ((T::DANC[i] >> j) & 1u) put(i, j, dot(cv[j], fv) + dot(ca[j], fa));
  }
}

// ---------------------------------------------------------------------------
// impedance and reference acceleration (constraint.py:29-65); im is an
// impedance record (bgx_layout.h) with the parameter-only factors
// precomputed on the host

template <class S>
BGX_D S impedance(const float* __restrict__ im, S pos) {
  const float dmin = im[IM_DMIN], dmax = im[IM_DMAX], mid = im[IM_MID],
              power = im[IM_POWER];
  const S x = fabs_(pos) * im[IM_INVWIDTH];
  S xa, xb;
  if (power == 2.f) {
    xa = x * x;
    xb = (1.f - x) * (1.f - x);
  } else {
    xa = pow_(x, power);
    xb = pow_(1.f - x, power);
  }
  const S ia = im[IM_ACOEF] * xa;
  const S ib = 1.f - im[IM_BCOEF] * xb;
  S imp = dmin + (val(x) < mid ? ia : ib) * (dmax - dmin);
  // jp.clip(imp, dmin, dmax), then imp = dmax beyond the width
  if (val(imp) < dmin) imp = lift<S>(dmin);
  if (val(imp) > dmax) imp = lift<S>(dmax);
  return val(x) > 1.f ? lift<S>(dmax) : imp;
}

// (1 - imp) / (imp + 1e-8): the diagonal regularizer factor
template <class S>
BGX_D S imp_ratio(S imp) {
  return (1.f - imp) * rcp(imp + 1e-8f);
}

// ---------------------------------------------------------------------------
// contact geometry of candidate c: MJX plane-sphere / plane-capsule

template <class T, class S>
BGX_D void candidate_center(const float* __restrict__ M, int c, const V3T<S>* xp,
                            const Q4T<S>* xr, V3T<S>& ctr, V3T<S>& axis) {
  const int l = T::CAND_LINK[c];
  const float* cd = cand_rec<T>(M, c);
  // geom world pose (contact.py:42-53)
  const V3T<S> gp = xp[l] + rotate(lift3<S>(ldv3(cd + CD_GPOS)), xr[l]);
  ctr = gp;
  const S zero = lift<S>(0.f);
  axis = V3T<S>{zero, zero, zero};
  if (T::CAND_KIND[c] == KIND_CAPSULE) {
    const Q4T<S> gq = qmul(xr[l], lift4<S>(ldq4(cd + CD_GQUAT)));
    axis = quat_zaxis(gq);
    ctr = gp + axis * (cd[CD_HALFLEN] * cd[CD_SIGN]);
  }
}

template <class T, class S>
BGX_D S candidate_dist(const float* __restrict__ M, int c, V3T<S> ctr) {
  const float* cd = cand_rec<T>(M, c);
  if (T::PLANE_Z) return (ctr.z - cd[CD_PPOS + 2]) - cd[CD_RADIUS];
  return dot(ctr - lift3<S>(ldv3(cd + CD_PPOS)), lift3<S>(ldv3(cd + CD_PN))) -
         cd[CD_RADIUS];
}

// contact point and frame of an active candidate (MJX plane collisions):
// pos = center - n (radius + dist / 2); sphere: the host-computed frame;
// capsule: t1 along the capsule axis projected on the plane, or a fixed axis
// when the capsule is near-normal (fallback < 0: decide here; 0 / 1: use the
// given decision); t2 = n x t1.  Returns whether the fallback was taken.
template <class T, class S>
BGX_D bool contact_frame(const float* __restrict__ M, int c, V3T<S> ctr, V3T<S> axis,
                         S dist, int fallback, V3T<S>& pos, V3T<S>& t1, V3T<S>& t2) {
  const float* cd = cand_rec<T>(M, c);
  const V3 nf = T::PLANE_Z ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN);
  const V3T<S> n = lift3<S>(nf);
  pos = ctr - n * (cd[CD_RADIUS] + 0.5f * dist);
  t1 = lift3<S>(ldv3(cd + CD_T1)), t2 = lift3<S>(ldv3(cd + CD_T2));
  bool fb = false;
  if (T::CAND_KIND[c] == KIND_CAPSULE) {
    S bn;
    const V3T<S> b = vnormalize(axis - n * dot(n, axis), &bn);
    const V3 alt = (-0.5f < nf.y && nf.y < 0.5f) ? v3(0.f, 1.f, 0.f) : v3(0.f, 0.f, 1.f);
    fb = fallback < 0 ? (val(bn) < 0.5f) : (fallback != 0);
    t1 = fb ? lift3<S>(alt) : b;
    t2 = cross(n, t1);
  }
  return fb;
}

// ---------------------------------------------------------------------------
// sphere and capsule pairs (MJX collision_primitive sphere_sphere,
// sphere_capsule, capsule_capsule and the segment helpers of its math
// module).  Geom 1 has the lower type; the normal points from geom 1 to
// geom 2.  A pair candidate's links: CAND_LINK holds a body link; when
// geom 2 is a world geom (CAND_FLIP) that link carries geom 1, else geom 2,
// and CAND_LINKA carries geom 1 (-1: the world).

template <class S>
BGX_D S clamp_s(S x, S lo, S hi) {
  return val(x) < val(lo) ? lo : (val(x) > val(hi) ? hi : x);
}

// the point of segment a-b closest to p (projection regularized by 1e-6)
template <class S>
BGX_D V3T<S> seg_point(V3T<S> a, V3T<S> b, V3T<S> p) {
  const V3T<S> ab = b - a;
  const S t = dot(p - a, ab) * rcp(dot(ab, ab) + 1e-6f);
  return a + ab * clamp_s(t, lift<S>(0.f), lift<S>(1.f));
}

// closest points of segments a0-a1 and b0-b1: the lines' closest points
// clipped to the segments, then the nearer of the two back-projections
template <class S>
BGX_D void seg_seg(V3T<S> a0, V3T<S> a1, V3T<S> b0, V3T<S> b1, V3T<S>& pa, V3T<S>& pb) {
  S la, lb;
  const V3T<S> da = vnormalize(a1 - a0, &la), db = vnormalize(b1 - b0, &lb);
  const S ha = la * 0.5f, hb = lb * 0.5f;
  const V3T<S> am = a0 + da * ha, bm = b0 + db * hb;
  const V3T<S> tr = am - bm;
  const S dd = dot(da, db), dat = dot(da, tr), dbt = dot(db, tr);
  const S ta0 = (dd * dbt - dat) * rcp((1.f - dd * dd) + 1e-6f);
  const S tb0 = dbt + ta0 * dd;
  const V3T<S> ba = am + da * clamp_s(ta0, lift<S>(0.f) - ha, ha);
  const V3T<S> bb = bm + db * clamp_s(tb0, lift<S>(0.f) - hb, hb);
  const V3T<S> na = seg_point(a0, a1, bb), nb = seg_point(b0, b1, ba);
  const V3T<S> ea = bb - na, eb = ba - nb;
  const bool first = val(dot(ea, ea)) < val(dot(eb, eb));
  pa = first ? na : ba;
  pb = first ? bb : nb;
}

// contact of spheres (p1, r1), (p2, r2): n from 1 to 2 (x for coincident
// centres), dist = |p2 - p1| - r1 - r2, pos = p1 + n (r1 + dist / 2), frame
// [a, b, a x b] with a = n normalized and b from orthogonals(a) (helper axis
// y when |a.y| < 1/2, else z; fallback < 0: decide here, else 0 / 1 = y / z).
// Returns whether z was the helper.
template <class S>
BGX_D bool sphere_pair(V3T<S> p1, float r1, V3T<S> p2, float r2, int fallback, S& dist,
                       V3T<S>& pos, V3T<S>& n, V3T<S>& t1, V3T<S>& t2) {
  S nn;
  V3T<S> u = vnormalize(p2 - p1, &nn);
  if (val(nn) == 0.f) u = lift3<S>(v3(1.f, 0.f, 0.f));
  dist = nn - (r1 + r2);
  pos = p1 + u * (r1 + dist * 0.5f);
  n = vnormalize(u, (S*)nullptr);
  const bool fb = fallback < 0 ? !(-0.5f < val(n.y) && val(n.y) < 0.5f) : fallback != 0;
  const V3T<S> h = lift3<S>(fb ? v3(0.f, 0.f, 1.f) : v3(0.f, 1.f, 0.f));
  t1 = vnormalize(h - n * dot(n, h), (S*)nullptr);
  t2 = cross(n, t1);
  return fb;
}

// world pose of a candidate's geom (contact.py:42-53): at (pos, quat) in
// link l's frame, or at that world pose for l < 0; axis: its z axis
template <class S>
BGX_D V3T<S> geom_pos(const float* cd, int off, int l, const V3T<S>* xp, const Q4T<S>* xr) {
  if (l < 0) return lift3<S>(ldv3(cd + off));
  return xp[l] + rotate(lift3<S>(ldv3(cd + off)), xr[l]);
}
template <class S>
BGX_D V3T<S> geom_axis(const float* cd, int off, int l, const Q4T<S>* xr) {
  if (l < 0) return lift3<S>(quat_zaxis(ldq4(cd + off)));
  return quat_zaxis(qmul(xr[l], lift4<S>(ldq4(cd + off))));
}

template <class T>
constexpr int pair_link1(int c) { return T::CAND_FLIP[c] ? T::CAND_LINK[c] : T::CAND_LINKA[c]; }
template <class T>
constexpr int pair_link2(int c) { return T::CAND_FLIP[c] ? -1 : T::CAND_LINK[c]; }

// contact of pair candidate c in the reference's orientation (normal from
// geom 1 to geom 2); returns the frame's helper-axis choice
template <class T, class S>
BGX_D bool pair_contact(const float* __restrict__ M, int c, const V3T<S>* xp, const Q4T<S>* xr,
                        int fallback, S& dist, V3T<S>& pos, V3T<S>& n, V3T<S>& t1,
                        V3T<S>& t2) {
  const float* cd = cand_rec<T>(M, c);
  const int kind = T::CAND_KIND[c], l1 = pair_link1<T>(c), l2 = pair_link2<T>(c);
  V3T<S> p1 = geom_pos<S>(cd, CD_G1POS, l1, xp, xr);
  V3T<S> p2 = geom_pos<S>(cd, CD_GPOS, l2, xp, xr);
  if (kind == KIND_SC) {
    const V3T<S> sg = geom_axis<S>(cd, CD_GQUAT, l2, xr) * cd[CD_HALFLEN];
    p2 = seg_point(p2 - sg, p2 + sg, p1);
  } else if (kind == KIND_CC) {
    const V3T<S> s1 = geom_axis<S>(cd, CD_G1QUAT, l1, xr) * cd[CD_HL1];
    const V3T<S> s2 = geom_axis<S>(cd, CD_GQUAT, l2, xr) * cd[CD_HALFLEN];
    V3T<S> c1, c2;
    seg_seg(p1 - s1, p1 + s1, p2 - s2, p2 + s2, c1, c2);
    p1 = c1, p2 = c2;
  }
  return sphere_pair(p1, cd[CD_R1], p2, cd[CD_RADIUS], fallback, dist, pos, n, t1, t2);
}

// ---------------------------------------------------------------------------
// position update of the semi-implicit Euler step (integrator.py:26-44)

// free-joint rotation: rot (x) quat_rot_axis(w / (|w| + 1e-8), dt (|w| + 1e-8)),
// renormalized
template <class S>
BGX_D Q4T<S> integrate_rot(Q4T<S> rot, V3T<S> ang, float dt) {
  const S an = fsqrt(dot(ang, ang)) + 1e-8f;
  const S ian = rcp(an);
  const Q4T<S> qr = qrot_axis(V3T<S>{ang.x * ian, ang.y * ian, ang.z * ian}, dt * an);
  const Q4T<S> r = qmul(rot, qr);
  const S rn = rsq(r.w * r.w + r.x * r.x + r.y * r.y + r.z * r.z);
  return Q4T<S>{r.w * rn, r.x * rn, r.y * rn, r.z * rn};
}

// q += qd dt for axis joints and the free-joint position; rotation as above
template <class T>
BGX_D void integrate_q(float* q, const float* qd, float dt) {
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const int qa = T::QADR[l], da = T::DADR[l];
    if (T::LTYPE[l] == 0) {
      const Q4 rot = integrate_rot(Q4{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]},
                                   v3(qd[da + 3], qd[da + 4], qd[da + 5]), dt);
      q[qa + 3] = rot.w, q[qa + 4] = rot.x, q[qa + 5] = rot.y, q[qa + 6] = rot.z;
#pragma unroll
      for (int k = 0; k < 3; ++k) q[qa + k] = q[qa + k] + qd[da + k] * dt;
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (k < T::LTYPE[l]) q[qa + k] = q[qa + k] + qd[da + k] * dt;
    }
  }
}

// ---------------------------------------------------------------------------
// contact slots and projected-gradient row storage

// geometry of one active contact: link, lever arm r = pos - root COM, and the
// pyramid directions n + s mu t_k (constraint.py:160-164).  On the z-up floor
// n = z and t1, t2 are horizontal, so mu t1 = (ax, ay, 0), mu t2 = (-ay, ax, 0)
template <bool PZ>
struct SlotGeo;
template <>
struct SlotGeo<true> {
  int link;
  V3 r;
  float ax, ay;
};
template <>
struct SlotGeo<false> {
  int link;
  V3 r, d0, d1, d2;
};

// a contact between two links (sphere / capsule pairs): rows act on the
// velocity of the contact point on `link` minus that on `linka` (-1: the
// world), with lever arms r and ra from each link's root COM
struct SlotGeoP {
  int link, linka;
  V3 r, ra, d0, d1, d2;
};
template <class T>
using SlotOf = std::conditional_t<T::PAIRS != 0, SlotGeoP, SlotGeo<T::PLANE_Z != 0>>;

BGX_D void slot_set(SlotGeo<true>& g, int link, V3 r, V3 n, V3 d1, V3 d2) {
  (void)n, (void)d2;
  g.link = link, g.r = r, g.ax = d1.x, g.ay = d1.y;
}
BGX_D void slot_set(SlotGeo<false>& g, int link, V3 r, V3 n, V3 d1, V3 d2) {
  g.link = link, g.r = r, g.d0 = n, g.d1 = d1, g.d2 = d2;
}
// J_c^T v for the 4 rows of a contact: the world force at the contact point
BGX_D V3 slot_force(const SlotGeo<true>& g, float v0, float v1, float v2, float v3) {
  const float a = v0 - v1, b = v2 - v3;
  return v3_(g.ax * a - g.ay * b, g.ay * a + g.ax * b, (v0 + v1) + (v2 + v3));
}
BGX_D V3 slot_force(const SlotGeo<false>& g, float v0, float v1, float v2, float v3) {
  return g.d0 * ((v0 + v1) + (v2 + v3)) + g.d1 * (v0 - v1) + g.d2 * (v2 - v3);
}
BGX_D V3 slot_force(const SlotGeoP& g, float v0, float v1, float v2, float v3) {
  return g.d0 * ((v0 + v1) + (v2 + v3)) + g.d1 * (v0 - v1) + g.d2 * (v2 - v3);
}
// (n.vp, mu t1.vp, mu t2.vp) for a contact-point velocity vp
BGX_D void slot_dots(const SlotGeo<true>& g, V3 vp, float& e0, float& e1, float& e2) {
  e0 = vp.z;
  e1 = g.ax * vp.x + g.ay * vp.y;
  e2 = g.ax * vp.y - g.ay * vp.x;
}
BGX_D void slot_dots(const SlotGeo<false>& g, V3 vp, float& e0, float& e1, float& e2) {
  e0 = dot(g.d0, vp), e1 = dot(g.d1, vp), e2 = dot(g.d2, vp);
}
BGX_D void slot_dots(const SlotGeoP& g, V3 vp, float& e0, float& e1, float& e2) {
  e0 = dot(g.d0, vp), e1 = dot(g.d1, vp), e2 = dot(g.d2, vp);
}
// velocity of slot g's contact point from the link velocities (ang, vel) at
// f + 6 l WAVE: on link, minus on linka for a two-link slot
template <class G>
BGX_D V3 slot_point_vel(const G& g, const float* f) {
  const float* a = f + g.link * 6 * WAVE;
  V3 v = v3(a[3 * WAVE], a[4 * WAVE], a[5 * WAVE]) + cross(v3(a[0], a[WAVE], a[2 * WAVE]), g.r);
  if constexpr (std::is_same_v<G, SlotGeoP>) {
    if (g.linka >= 0) {
      const float* b = f + g.linka * 6 * WAVE;
      v = v - (v3(b[3 * WAVE], b[4 * WAVE], b[5 * WAVE]) +
               cross(v3(b[0], b[WAVE], b[2 * WAVE]), g.ra));
    }
  }
  return v;
}
// adds force w at the contact point of slot g to the link forces (torque,
// force) at f + 6 l WAVE: +w on link, -w on linka
template <class G>
BGX_D void slot_add_force(const G& g, V3 w, float* f) {
  const V3 ta = cross(g.r, w);
  float* a = f + g.link * 6 * WAVE;
  a[0] += ta.x, a[WAVE] += ta.y, a[2 * WAVE] += ta.z;
  a[3 * WAVE] += w.x, a[4 * WAVE] += w.y, a[5 * WAVE] += w.z;
  if constexpr (std::is_same_v<G, SlotGeoP>) {
    if (g.linka >= 0) {
      const V3 tb = cross(g.ra, w);
      float* b = f + g.linka * 6 * WAVE;
      b[0] -= tb.x, b[WAVE] -= tb.y, b[2 * WAVE] -= tb.z;
      b[3 * WAVE] -= w.x, b[4 * WAVE] -= w.y, b[5 * WAVE] -= w.z;
    }
  }
}

// per-row solver vectors of the slots: x, y (FISTA iterates), g (gradient),
// r (residual).  Fast path: LDS, element e of this lane at lds[e * WAVE];
// overflow path: a private array.  (b stays in registers.)
constexpr int RV_X = 0, RV_Y = 1, RV_G = 2, RV_R = 3, RV_N = 4;
template <int KS>
struct LdsRows {
  float* base;
  BGX_D float& at(int v, int s, int k) const {
    return base[((v * KS + s) * 4 + k) * WAVE];
  }
};
// limit-row solver vectors x, y, g, r (k = 0..3): registers, or LDS for
// LM_LDS topologies (element (k, i) of this lane at base[(k NLIM + i) WAVE])
template <class T, bool LDS>
struct LimVecs;
template <class T>
struct LimVecs<T, false> {
  float v[4][T::NLIM > 0 ? T::NLIM : 1];
  BGX_D float& at(int k, int i) { return v[k][i]; }
};
template <class T>
struct LimVecs<T, true> {
  float* base;
  BGX_D float& at(int k, int i) { return base[(k * T::NLIM + i) * WAVE]; }
};

template <int KS>
struct PrivRows {
  float a[RV_N][KS][4];
  BGX_D float& at(int v, int s, int k) { return a[v][s][k]; }
};

// ---------------------------------------------------------------------------
// optional record of a substep's intermediates (the adjoint's tape): with a
// slot per candidate, the tape holds what the reverse sweep needs -- the
// tree (cdof, LTL factor), the contact slots, the unconstrained and
// constraint accelerations, every projected-gradient iterate and the
// products M^-1 J^T y_k, M^-1 J^T r_k of each iteration.  The step
// kernels pass NoTape, which compiles to nothing.

constexpr int ADJ_MAXITER = 8;  // most solver iterations the tape holds

template <class T>
struct TapeLayout {
  static constexpr int NV = T::NV, NQ = T::NQ;
  static constexpr int KS = T::NCAND > 0 ? T::NCAND : 1;  // one slot per candidate
  static constexpr int NRC = 4 * KS;                      // contact rows
  static constexpr int NLR = T::NLIM > 0 ? T::NLIM : 1;
  static constexpr int NR = NRC + NLR;                    // contact rows, then limit rows
  static constexpr int NMB = mpat_row<T>(T::NV);
  static constexpr int CSLOT = 0;              // [KS] slot of each candidate, -1 inactive
  static constexpr int CFB = CSLOT + KS;       // [KS] capsule frame fallback taken
  static constexpr int NC = CFB + KS;          // active contacts
  static constexpr int ITERS = NC + 1;         // solver iterations run
  static constexpr int CA = ITERS + 1;         // [3 NV] cdof angular
  static constexpr int CV = CA + 3 * NV;       // [3 NV] cdof linear
  static constexpr int LM = CV + 3 * NV;       // [NMB] LTL factor (pattern order)
  static constexpr int INV = LM + NMB;         // [NV] 1 / L_kk
  static constexpr int SLINK = INV + NV;       // [KS] link of each slot
  static constexpr int SR = SLINK + KS;        // [3 KS] lever arm
  static constexpr int SD0 = SR + 3 * KS;      // [3 KS] normal
  static constexpr int SD1 = SD0 + 3 * KS;     // [3 KS] mu t1
  static constexpr int SD2 = SD1 + 3 * KS;     // [3 KS] mu t2
  static constexpr int SDIAG = SD2 + 3 * KS;   // [KS] regularizer of the slot's rows
  static constexpr int LSIDE = SDIAG + KS;     // [NLR] limit row sign
  static constexpr int LDIAG = LSIDE + NLR;    // [NLR] limit row regularizer
  static constexpr int U = LDIAG + NLR;        // [NV] M^-1 qf_smooth
  static constexpr int W = U + NV;             // [NV] M^-1 J^T x
  static constexpr int QD1 = W + NV;           // [NV] qd after the substep
  static constexpr int ETA = QD1 + NV;         // [ADJ_MAXITER] accepted step sizes
  static constexpr int BETA = ETA + ADJ_MAXITER;      // [ADJ_MAXITER] momentum factors
  static constexpr int YS = BETA + ADJ_MAXITER;       // [ADJ_MAXITER][NR] y_k
  static constexpr int RS = YS + ADJ_MAXITER * NR;    // [ADJ_MAXITER][NR] A y_k + b
  static constexpr int XS = RS + ADJ_MAXITER * NR;    // [ADJ_MAXITER][NR] x_{k+1}
  static constexpr int UY = XS + ADJ_MAXITER * NR;    // [ADJ_MAXITER][NV] M^-1 J^T y_k
  static constexpr int UR = UY + ADJ_MAXITER * NV;    // [ADJ_MAXITER][NV] M^-1 J^T r_k
  // two-link slots (topologies with sphere / capsule pairs)
  static constexpr int KSP = T::PAIRS ? KS : 0;
  static constexpr int SLINKA = UR + ADJ_MAXITER * NV;  // [KSP] second link (-1: none)
  static constexpr int SRA = SLINKA + KSP;              // [3 KSP] its lever arm
  static constexpr int SIZE = SRA + 3 * KSP;
};

struct NoTape {
  static constexpr bool ON = false;
};

// tape in global memory, field-major: field f of this env at p[f * n]
struct RecPtr {
  static constexpr bool ON = true;
  float* p;
  long long n;
  BGX_D float& operator[](int f) const { return p[(long long)f * n]; }
};

// an integer field of a tape (a count, slot or link index): the float
// rounded toward zero and clamped to [lo, hi], NaN to lo.  The forward
// writes in-range values; the clamp keeps a damaged tape from indexing
// outside the arrays its readers address.
BGX_D int tape_int(float v, int lo, int hi) {
  return v >= (float)lo ? (v <= (float)hi ? (int)v : hi) : lo;
}

// pyramid directions of a slot in the generic form (n, mu t1, mu t2)
BGX_D void slot_dirs(const SlotGeo<true>& g, V3& d0, V3& d1, V3& d2) {
  d0 = v3(0.f, 0.f, 1.f), d1 = v3(g.ax, g.ay, 0.f), d2 = v3(-g.ay, g.ax, 0.f);
}
BGX_D void slot_dirs(const SlotGeo<false>& g, V3& d0, V3& d1, V3& d2) {
  d0 = g.d0, d1 = g.d1, d2 = g.d2;
}
BGX_D void slot_dirs(const SlotGeoP& g, V3& d0, V3& d1, V3& d2) {
  d0 = g.d0, d1 = g.d1, d2 = g.d2;
}
template <class Tp>
BGX_D void tape_v3(const Tp& tape, int f, V3 v) {
  tape[f] = v.x, tape[f + 1] = v.y, tape[f + 2] = v.z;
}

// ---------------------------------------------------------------------------
// one substep: pipeline.step (pipeline.py:64-94) with KS contact slots.
// xp, xr: link frames at (q, qd) (kinematics.forward); lds: this lane's
// link scratch (6 floats per link) followed by space for 54 more floats

template <class T, int KS, class Rows, class Tape = NoTape>
BGX_D bool substep_impl(const float* __restrict__ M, float* q, float* qd,
                        const float* act, const V3* xp, const Q4* xr, const V3* rcl,
                        float* lds, Rows& rows, float* dbg, Tape tape = Tape{}) {
  using TL = TapeLayout<T>;
  constexpr int NV = T::NV, L = T::L;
  constexpr int NL = T::NLIM > 0 ? T::NLIM : 1;
  constexpr bool PZ = T::PLANE_Z != 0;
  Tree<T> t;
  if constexpr (T::LM_LDS) t.lml = lds + lds_lm_offset<T>() * WAVE;
  float* lds2 = lds + L * 6 * WAVE;  // second link-vector area (rows space)
#pragma unroll
  for (int l = 0; l < L; ++l) t.rc[l] = rcl[l];

  // ---- contacts (contact.py + MJX collisions) into slots
  using SG = SlotOf<T>;
  SG sg[KS];
  float sl_diag[KS], sl_bq[KS], sl_kt[KS];
  int n_c = 0;
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    const float* cd = cand_rec<T>(M, c);
    const bool pair = T::PAIRS && T::CAND_KIND[c] >= KIND_SS;
    float dist;
    V3 ctr, axis, pos, n, t1, t2;
    bool fb = false;
    if (pair) {
      fb = pair_contact<T>(M, c, xp, xr, -1, dist, pos, n, t1, t2);
    } else {
      candidate_center<T>(M, c, xp, xr, ctr, axis);
      dist = candidate_dist<T>(M, c, ctr);
    }
    if constexpr (Tape::ON) tape[TL::CSLOT + c] = -1.f, tape[TL::CFB + c] = 0.f;
    if (dist < 0.f) {
      const int l = T::CAND_LINK[c];
      if (!pair) {
        fb = contact_frame<T>(M, c, ctr, axis, dist, -1, pos, t1, t2);
        n = PZ ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN);
      }
      if constexpr (Tape::ON) tape[TL::CSLOT + c] = (float)n_c, tape[TL::CFB + c] = fb ? 1.f : 0.f;
      const float mu = cd[CD_MU];
      SG g;
      if constexpr (T::PAIRS) {
        // rows act on (geom 2's link - geom 1's link): a world geom 2 turns
        // them onto geom 1's link with the directions negated
        const float sgn = T::CAND_FLIP[c] ? -1.f : 1.f;
        const int la = T::CAND_LINKA[c];
        g.link = l, g.linka = la;
        g.r = pos - t.rc[l];
        g.ra = la >= 0 ? pos - t.rc[la >= 0 ? la : 0] : v3zero();
        g.d0 = n * sgn, g.d1 = t1 * (mu * sgn), g.d2 = t2 * (mu * sgn);
      } else {
        slot_set(g, l, pos - t.rc[l], n, t1 * mu, t2 * mu);
      }
      const float* im = cd + CD_IMP;
      const float imp = impedance(im, dist);
      const float tw = cd[CD_TW];
      const float diag = (tw + mu * mu * tw) * (2.f * mu * mu * imp_ratio(imp));
      const float kt = im[IM_K] * imp * dist;
      const float bq = im[IM_B];
      const int s = n_c;
#pragma unroll
      for (int ss = 0; ss < KS; ++ss) {
        if (ss <= c && ss == s) {
          sg[ss] = g;
          sl_diag[ss] = diag;
          sl_bq[ss] = bq;
          sl_kt[ss] = kt;
        }
      }
      n_c += 1;
    }
  }
  // more active contacts than slots: leave the state untouched and let the
  // caller take the path with a slot per candidate
  if (n_c > KS) return false;
  if constexpr (Tape::ON) {
    tape[TL::NC] = (float)n_c;
#pragma unroll
    for (int s = 0; s < KS; ++s)
      if (s < n_c) {
        V3 d0, d1, d2;
        slot_dirs(sg[s], d0, d1, d2);
        tape[TL::SLINK + s] = (float)sg[s].link;
        tape_v3(tape, TL::SR + 3 * s, sg[s].r);
        if constexpr (T::PAIRS) {
          tape[TL::SLINKA + s] = (float)sg[s].linka;
          tape_v3(tape, TL::SRA + 3 * s, sg[s].ra);
        }
        tape_v3(tape, TL::SD0 + 3 * s, d0);
        tape_v3(tape, TL::SD1 + 3 * s, d1);
        tape_v3(tape, TL::SD2 + 3 * s, d2);
        tape[TL::SDIAG + s] = sl_diag[s];
      }
  }

  // ---- dynamics.transform_com (dynamics.py:27-89): cinr, cdof
  V3 cc[L];
  Sym3 ci[L];
  com_inertia<T>(M, xp, xr, t.rc, cc, ci);
  dof_motions<T>(M, q, xp, xr, t.rc, t.ca, t.cv);

  // ---- cd: link velocities at qd (dynamics.py:90-103), also kept in lds2
  V3 cda[L], cdv[L];
  {
    float qv[NV];
#pragma unroll
    for (int d = 0; d < NV; ++d) qv[d] = qd[d];
    link_velocities<T>(t, qv, cda, cdv);
  }
#pragma unroll
  for (int l = 0; l < L; ++l)
    if (hosts_contact<T>(l)) {
      float* f = lds2 + l * 6 * WAVE;
      f[0] = cda[l].x, f[WAVE] = cda[l].y, f[2 * WAVE] = cda[l].z;
      f[3 * WAVE] = cdv[l].x, f[4 * WAVE] = cdv[l].y, f[5 * WAVE] = cdv[l].z;
    }

  // ---- dynamics.inverse: RNEA bias (dynamics.py:105-184), cdofd on the fly;
  // then passive forces and actuators: qf_smooth = passive - bias + tau
  float qf[NV];
  rnea_bias<T>(M, t.ca, t.cv, cda, cdv, cc, ci, qd, qf);
  smooth_forces<T>(M, q, qd, act, qf);
  if (dbg) {
#pragma unroll
    for (int d = 0; d < NV; ++d) dbg[d] = qf[d];
  }

  // ---- mass.matrix: CRBA on the dof-ancestor pattern (mass.py:27-83), then
  // the LTL factorization, leaves first
  crba<T>(M, cc, ci, t.ca, t.cv, [&](int i, int j, float v) { t.lm(i, j) = v; });
  ltl_factor<T>(t);

  // ---- unconstrained acceleration u = M^-1 qf_smooth
  float u[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) u[d] = qf[d];
  ltl_solve<T>(t, u);
  if constexpr (Tape::ON) {
#pragma unroll
    for (int d = 0; d < NV; ++d) {
      tape_v3(tape, TL::CA + 3 * d, t.ca[d]);
      tape_v3(tape, TL::CV + 3 * d, t.cv[d]);
      tape[TL::INV + d] = t.inv[d];
      tape[TL::U + d] = u[d];
      tape[TL::LM + mpat_idx<T>(d, d)] = t.lm(d, d);
#pragma unroll
      for (int j = 0; j < d; ++j)
        if ((T::DANC[d] >> j) & 1u) tape[TL::LM + mpat_idx<T>(d, j)] = t.lm(d, j);
    }
  }

  // ---- contact rows: b = J u - aref, aref = -b J qd - k imp dist
  // (constraint.py:58-65, :205-210); link velocities from lds (u) and lds2 (qd)
  {
    V3 ua[L], uv[L];
    link_velocities<T>(t, u, ua, uv);
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (hosts_contact<T>(l)) {
        float* f = lds + l * 6 * WAVE;
        f[0] = ua[l].x, f[WAVE] = ua[l].y, f[2 * WAVE] = ua[l].z;
        f[3 * WAVE] = uv[l].x, f[4 * WAVE] = uv[l].y, f[5 * WAVE] = uv[l].z;
      }
  }
  float sb[KS][4];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    if (s < n_c) {
      const V3 vu = slot_point_vel(sg[s], lds);
      const V3 vq = slot_point_vel(sg[s], lds2);
      float e0u, e1u, e2u, e0q, e1q, e2q;
      slot_dots(sg[s], vu, e0u, e1u, e2u);
      slot_dots(sg[s], vq, e0q, e1q, e2q);
      const float bq = sl_bq[s], kt = sl_kt[s];
      sb[s][0] = (e0u + e1u) + (bq * (e0q + e1q) + kt);
      sb[s][1] = (e0u - e1u) + (bq * (e0q - e1q) + kt);
      sb[s][2] = (e0u + e2u) + (bq * (e0q + e2q) + kt);
      sb[s][3] = (e0u - e2u) + (bq * (e0q - e2q) + kt);
    }
  }
#pragma unroll
  for (int s = 0; s < KS; ++s)
    if (s < n_c)
#pragma unroll
      for (int k = 0; k < 4; ++k) rows.at(RV_X, s, k) = rows.at(RV_Y, s, k) = 0.f;

  // ---- limit rows (constraint.py:101-132), one per non-free dof
  float lside[NL], ldiag[NL], lb[NL];
  bool any_lim = false;
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i) {
    const int d = T::LIM_D[i], qi = T::LIM_Q[i];
    const float* df = dof_rec<T>(M, d);
    const float pmin = q[qi] - df[DF_LO];
    const float pmax = df[DF_HI] - q[qi];
    const float pos = fminf(fminf(pmin, pmax), 0.f);
    const bool act = pos < 0.f;
    const float side = act ? (pmin < pmax ? 1.f : -1.f) : 0.f;
    const float* im = df + DF_IMP;
    const float imp = impedance(im, pos);
    const float vel = side * qd[d];
    const float aref = act ? (-im[IM_B] * vel - im[IM_K] * imp * pos) : 0.f;
    lside[i] = side;
    ldiag[i] = act ? df[DF_INVWEIGHT] * imp_ratio(imp) : 0.f;
    if constexpr (Tape::ON) tape[TL::LSIDE + i] = side, tape[TL::LDIAG + i] = ldiag[i];
    lb[i] = side * u[d] - aref;
    any_lim = any_lim || act;
  }

  // ---- products with A = J M^-1 J^T + diag, evaluated through the tree
  auto apply_jt = [&](auto in_slot, auto in_lim, float* tau) {
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (hosts_contact<T>(l)) {
#pragma unroll
        for (int k = 0; k < 6; ++k) lds[(l * 6 + k) * WAVE] = 0.f;
      }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (s < n_c) {
        const V3 w = slot_force(sg[s], in_slot(s, 0), in_slot(s, 1),
                                in_slot(s, 2), in_slot(s, 3));
        slot_add_force(sg[s], w, lds);
      }
    }
    V3 fa[L], fv[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (hosts_contact<T>(l)) {
        const float* f = lds + l * 6 * WAVE;
        fa[l] = v3(f[0], f[WAVE], f[2 * WAVE]);
        fv[l] = v3(f[3 * WAVE], f[4 * WAVE], f[5 * WAVE]);
      } else {
        fa[l] = v3zero(), fv[l] = v3zero();
      }
    }
    project_forces<T>(t, fa, fv, tau);
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) tau[T::LIM_D[i]] += lside[i] * in_lim(i);
  };

  auto no_z = [](const float*) {};
  // out = A in; z_out receives M^-1 J^T in
  auto apply_a = [&](auto in_slot, auto in_lim, auto out_slot, auto out_lim, auto z_out) {
    float z[NV];
    apply_jt(in_slot, in_lim, z);
    ltl_solve<T>(t, z);
    z_out(z);
    {
      V3 va[L], vv[L];
      link_velocities<T>(t, z, va, vv);
#pragma unroll
      for (int l = 0; l < L; ++l)
        if (hosts_contact<T>(l)) {
          float* f = lds + l * 6 * WAVE;
          f[0] = va[l].x, f[WAVE] = va[l].y, f[2 * WAVE] = va[l].z;
          f[3 * WAVE] = vv[l].x, f[4 * WAVE] = vv[l].y, f[5 * WAVE] = vv[l].z;
        }
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (s < n_c) {
        float e0, e1, e2;
        slot_dots(sg[s], slot_point_vel(sg[s], lds), e0, e1, e2);
        const float dg = sl_diag[s];
        out_slot(s, 0, e0 + e1 + dg * in_slot(s, 0));
        out_slot(s, 1, e0 - e1 + dg * in_slot(s, 1));
        out_slot(s, 2, e0 + e2 + dg * in_slot(s, 2));
        out_slot(s, 3, e0 - e2 + dg * in_slot(s, 3));
      }
    }
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i)
      out_lim(i, lside[i] * z[T::LIM_D[i]] + ldiag[i] * in_lim(i));
  };

  // ---- jaxopt.ProjectedGradient: FISTA + backtracking line search from 0
  LimVecs<T, T::LM_LDS != 0> lv;
  if constexpr (T::LM_LDS)
    lv.base = lds + (lds_lm_offset<T>() + mpat_row<T>(T::NV)) * WAVE;
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i)
#pragma unroll
    for (int k = 0; k < 4; ++k) lv.at(k, i) = 0.f;

  const bool has_rows = (n_c > 0) || any_lim;
  const int maxiter = Tape::ON && (int)M[H_MAXITER] > ADJ_MAXITER
                          ? ADJ_MAXITER : (int)M[H_MAXITER];
  const int maxls = (int)M[H_MAXLS];
  const float tol = M[H_TOL], lseps = M[H_LSEPS];
  float tk = 1.f, eta = 1.f, err = INFINITY;
  bool running = has_rows;
  int iters = 0, trials = 0;
  for (int it = 0; it < maxiter; ++it) {
    if (it > 0) running = running && (err > tol);
    if (!running) break;
    // gradient at y: r = A y + b, g = A^T r
    float fy = 0.f;
    apply_a([&](int s, int k) { return rows.at(RV_Y, s, k); },
            [&](int i) { return lv.at(1, i); },
            [&](int s, int k, float o) {
              const float r = o + sb[s][k];
              rows.at(RV_R, s, k) = r;
              fy += 0.5f * r * r;
            },
            [&](int i, float o) {
              const float r = o + lb[i];
              lv.at(3, i) = r;
              fy += 0.5f * r * r;
            },
            [&](const float* z) {
              if constexpr (Tape::ON)
#pragma unroll
                for (int d = 0; d < NV; ++d) tape[TL::UY + it * NV + d] = z[d];
            });
    apply_a([&](int s, int k) { return rows.at(RV_R, s, k); },
            [&](int i) { return lv.at(3, i); },
            [&](int s, int k, float o) { rows.at(RV_G, s, k) = o; },
            [&](int i, float o) { lv.at(2, i) = o; },
            [&](const float* z) {
              if constexpr (Tape::ON)
#pragma unroll
                for (int d = 0; d < NV; ++d) tape[TL::UR + it * NV + d] = z[d];
            });
    if constexpr (Tape::ON) {  // y_k and r_k = A y_k + b
      const int o = it * TL::NR;
#pragma unroll
      for (int s = 0; s < KS; ++s)
        if (s < n_c)
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            tape[TL::YS + o + 4 * s + k] = rows.at(RV_Y, s, k);
            tape[TL::RS + o + 4 * s + k] = rows.at(RV_R, s, k);
          }
#pragma unroll
      for (int i = 0; i < T::NLIM; ++i) {
        tape[TL::YS + o + TL::NRC + i] = lv.at(1, i);
        tape[TL::RS + o + TL::NRC + i] = lv.at(3, i);
      }
    }
    // line search: trial x+ = relu(y - eta g), residual into r
    bool cond = true;
    for (int ls = 0; ls <= maxls && cond; ++ls) {
      if (ls > 0) eta *= 0.5f;
      const float e_ = eta;
      float fx = 0.f, dg = 0.f, dd = 0.f;
      apply_a([&](int s, int k) {
                return fmaxf(rows.at(RV_Y, s, k) - e_ * rows.at(RV_G, s, k), 0.f);
              },
              [&](int i) { return fmaxf(lv.at(1, i) - e_ * lv.at(2, i), 0.f); },
              [&](int s, int k, float o) {
                const float r = o + sb[s][k];
                rows.at(RV_R, s, k) = r;
                fx += 0.5f * r * r;
                const float yv = rows.at(RV_Y, s, k), gv = rows.at(RV_G, s, k);
                const float df = fmaxf(yv - e_ * gv, 0.f) - yv;
                dg += df * gv;
                dd += df * df;
              },
              [&](int i, float o) {
                const float r = o + lb[i];
                lv.at(3, i) = r;
                fx += 0.5f * r * r;
                const float df = fmaxf(lv.at(1, i) - e_ * lv.at(2, i), 0.f) - lv.at(1, i);
                dg += df * lv.at(2, i);
                dd += df * df;
              },
              no_z);
      trials += 1;
      cond = e_ * (fx - fy) > e_ * dg + 0.5f * dd + lseps;
    }
    // error at x+: |relu(x+ - grad f(x+)) - x+| with grad = A^T (A x+ + b)
    float e2 = 0.f;
    const float e_ = eta;
    apply_a([&](int s, int k) { return rows.at(RV_R, s, k); },
            [&](int i) { return lv.at(3, i); },
            [&](int s, int k, float o) {
              const float xp = fmaxf(rows.at(RV_Y, s, k) - e_ * rows.at(RV_G, s, k), 0.f);
              const float dv = fmaxf(xp - o, 0.f) - xp;
              e2 += dv * dv;
            },
            [&](int i, float o) {
              const float xp = fmaxf(lv.at(1, i) - e_ * lv.at(2, i), 0.f);
              const float dv = fmaxf(xp - o, 0.f) - xp;
              e2 += dv * dv;
            },
            no_z);
    err = fsqrt(e2);
    const float tn = 0.5f * (1.f + fsqrt(1.f + 4.f * tk * tk));
    const float beta = (tk - 1.f) * rcp(tn);
#pragma unroll
    for (int s = 0; s < KS; ++s)
      if (s < n_c)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float xp = fmaxf(rows.at(RV_Y, s, k) - e_ * rows.at(RV_G, s, k), 0.f);
          rows.at(RV_Y, s, k) = xp + beta * (xp - rows.at(RV_X, s, k));
          rows.at(RV_X, s, k) = xp;
          if constexpr (Tape::ON) tape[TL::XS + it * TL::NR + 4 * s + k] = xp;
        }
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) {
      const float xp = fmaxf(lv.at(1, i) - e_ * lv.at(2, i), 0.f);
      lv.at(1, i) = xp + beta * (xp - lv.at(0, i));
      lv.at(0, i) = xp;
      if constexpr (Tape::ON) tape[TL::XS + it * TL::NR + TL::NRC + i] = xp;
    }
    if constexpr (Tape::ON) tape[TL::ETA + it] = e_, tape[TL::BETA + it] = beta;
    eta = eta <= 1e-6f ? 1.f : eta * 2.f;
    tk = tn;
    iters += 1;
  }

  // ---- qdd = M^-1 (qf_smooth + J^T x) (integrator.py:58-64)
  float qdd[NV];
  if constexpr (Tape::ON) tape[TL::ITERS] = (float)iters;
  if (has_rows) {
    apply_jt([&](int s, int k) { return rows.at(RV_X, s, k); },
             [&](int i) { return lv.at(0, i); }, qdd);
    ltl_solve<T>(t, qdd);
#pragma unroll
    for (int d = 0; d < NV; ++d) {
      if constexpr (Tape::ON) tape[TL::W + d] = qdd[d];
      qdd[d] = u[d] + qdd[d];
    }
  } else {
#pragma unroll
    for (int d = 0; d < NV; ++d) {
      if constexpr (Tape::ON) tape[TL::W + d] = 0.f;
      qdd[d] = u[d];
    }
  }
  if (dbg) {  // per-env probe of the last substep: qdd, counts
#pragma unroll
    for (int d = 0; d < NV; ++d) dbg[NV + d] = qdd[d];
    dbg[2 * NV] = (float)n_c;
    dbg[2 * NV + 1] = (float)iters;
    dbg[2 * NV + 2] = (float)trials;
  }

  // ---- semi-implicit Euler (integrator.py:26-84)
  const float dt = M[H_DT];
#pragma unroll
  for (int d = 0; d < NV; ++d) qd[d] = qd[d] + qdd[d] * dt;
  if constexpr (Tape::ON) {
#pragma unroll
    for (int d = 0; d < NV; ++d) tape[TL::QD1 + d] = qd[d];
  }
  integrate_q<T>(q, qd, dt);
  return true;
}

// kinematics, root COM, then substep_impl with KS slots
template <class T, int KS, class Rows, class Tape = NoTape>
BGX_D bool substep_ks(const float* __restrict__ M, float* q, float* qd,
                      const float* act, float* lds, Rows& rows, float* dbg,
                      Tape tape = Tape{}) {
  static_assert(T::L > 0, "empty topology");
  V3 xp[T::L];
  Q4 xr[T::L];
  forward_kinematics<T, false, float>(M, q, qd, xp, xr, nullptr, nullptr);
  V3 rc[T::L];
  root_coms<T>(M, xp, xr, rc);  // dynamics.py:37-43
  return substep_impl<T, KS>(M, q, qd, act, xp, xr, rc, lds, rows, dbg, tape);
}

// an opaque copy of the model pointer (offset by the table's zero slot, which
// the compiler cannot fold): loads through it are not merged with loads
// through the original
BGX_D const float* launder(const float* M) { return M + (int)M[H_ZERO]; }

// one substep: the fast path keeps the active contacts in KC slots (rows in
// LDS); when more candidates touch the floor it returns before changing the
// state and the overflow path redoes the substep with a slot per candidate
template <class T>
BGX_D int substep(const float* __restrict__ M, float* q, float* qd,
                  const float* act, float* lds, float* dbg) {
  LdsRows<T::KC> rows{lds + T::L * 6 * WAVE};
  if (substep_ks<T, T::KC>(M, q, qd, act, lds, rows, dbg)) return 0;
  if constexpr (T::KC < T::NCAND) {
    PrivRows<T::NCAND> prow;
    substep_ks<T, T::NCAND>(launder(M), q, qd, act, lds, prow, dbg);
  }
  return ST_SLOW_PATH;
}

}  // namespace bgx
