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

template <class T>
constexpr bool hosts_contact(int l) {
  for (int c = 0; c < T::NCAND; ++c)
    if (T::CAND_LINK[c] == l) return true;
  return false;
}

// ---------------------------------------------------------------------------
// kinematics.forward (kinematics.py:31-108); XD: also compute xd

template <class T, bool XD>
BGX_D void forward_kinematics(const float* __restrict__ M, const float* q,
                              const float* qd, V3* xp, Q4* xr, V3* xa, V3* xv) {
  V3 jp[T::L], ja[T::L], jv[T::L];
  Q4 jr[T::L];
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const float* lk = link_rec<T>(M, l);
    const int qa = T::QADR[l], da = T::DADR[l];
    V3 p = v3zero(), a = v3zero(), v = v3zero();
    Q4 r = qident();
    if (T::LTYPE[l] == 0) {  // free joint (kinematics.py:48-51)
      p = v3(q[qa], q[qa + 1], q[qa + 2]);
      r = Q4{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]};
      a = v3(qd[da + 3], qd[da + 4], qd[da + 5]);
      v = v3(qd[da], qd[da + 1], qd[da + 2]);
    } else {  // 1-3 stacked axis joints (kinematics.py:52-81)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k < T::LTYPE[l]) {
          const float* df = dof_rec<T>(M, da + k);
          const V3 ang = ldv3(df + DF_ANG), vel = ldv3(df + DF_VEL);
          const float qk = q[qa + k], qdk = qd[da + k];
          const Q4 rk = qnormalize(qrot_axis(ang, qk));
          const V3 pk = vel * qk, jda = ang * qdk, jdv = vel * qdk;
          if (k == 0) {
            p = pk, r = rk, a = jda, v = jdv;
          } else {
            p = p + rotate(pk, r);
            r = qmul(r, rk);
            a = a + rotate(jda, rk);
            v = v + rotate(jdv + cross(pk, jda), rk);
          }
        }
      }
    }
    // joint anchor offset and link transform (kinematics.py:85-87)
    const V3 jpos = ldv3(lk + LK_JPOS);
    p = (p + jpos) - rotate(jpos, r);
    const V3 tpos = ldv3(lk + LK_TPOS);
    const Q4 trot = ldq4(lk + LK_TROT);
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
        V3 vel = xv[par] + cross(xa[par], xp[l] - xp[par]);
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

template <class T>
struct Tree {
  V3 ca[T::NV], cv[T::NV];  // cdof (ang, vel)
  float Lm[T::NV][T::NV];   // LTL factor, lower part on the dof-ancestor pattern
  float inv[T::NV];         // 1 / L_kk
  V3 rc[T::L];              // root COM of each link's tree
};

// M z = tau with M = L^T L (in place on z)
template <class T>
BGX_D void ltl_solve(const Tree<T>& t, float* z) {
#pragma unroll
  for (int i = T::NV - 1; i >= 0; --i) {
    z[i] *= t.inv[i];
#pragma unroll
    for (int j = 0; j < i; ++j)
      if ((T::DANC[i] >> j) & 1u) z[j] -= t.Lm[i][j] * z[i];
  }
#pragma unroll
  for (int i = 0; i < T::NV; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j)
      if ((T::DANC[i] >> j) & 1u) z[i] -= t.Lm[i][j] * z[j];
    z[i] *= t.inv[i];
  }
}

// spatial velocity of every link from generalized velocities z (the forward
// scan of dynamics.py:93-103 with cdof)
template <class T>
BGX_D void link_velocities(const Tree<T>& t, const float* z, V3* va, V3* vv) {
#pragma unroll
  for (int l = 0; l < T::L; ++l) {
    const int par = T::PARENT[l];
    V3 a = par < 0 ? v3zero() : va[par];
    V3 v = par < 0 ? v3zero() : vv[par];
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = T::DADR[l] + k;
        a = a + t.ca[d] * z[d];
        v = v + t.cv[d] * z[d];
      }
    }
    va[l] = a, vv[l] = v;
  }
}

// generalized force of per-link spatial forces (fa: torque about the root
// COM, fv: force), accumulated leaf to root (dynamics.py:174-182)
template <class T>
BGX_D void project_forces(const Tree<T>& t, V3* fa, V3* fv, float* tau) {
#pragma unroll
  for (int l = T::L - 1; l >= 0; --l) {
    const int par = T::PARENT[l];
    if (par >= 0) {
      fa[par] = fa[par] + fa[l];
      fv[par] = fv[par] + fv[l];
    }
  }
#pragma unroll
  for (int d = 0; d < T::NV; ++d) {
    const int l = T::DOF_LINK[d];
    tau[d] = dot(t.cv[d], fv[l]) + dot(t.ca[d], fa[l]);
  }
}

// ---------------------------------------------------------------------------
// impedance and reference acceleration (constraint.py:29-65)

BGX_D void imp_aref(const float* sp, float pos, float& imp, float& bcoef,
                    float& kcoef) {
  const float timeconst = sp[0], dampratio = sp[1], dmin = sp[2],
              dmax = sp[3], width = sp[4], mid = sp[5], power = sp[6];
  const float imp_x = fabsf(pos) / width;
  const float imp_a = (1.f / powf(mid, power - 1.f)) * powf(imp_x, power);
  const float imp_b =
      1.f - (1.f / powf(1.f - mid, power - 1.f)) * powf(1.f - imp_x, power);
  const float imp_y = imp_x < mid ? imp_a : imp_b;
  float im = dmin + imp_y * (dmax - dmin);
  im = clipf(im, dmin, dmax);
  im = imp_x > 1.f ? dmax : im;
  float b = 2.f / (dmax * timeconst);
  float k = 1.f / (dmax * dmax * timeconst * timeconst * dampratio * dampratio);
  b = dampratio <= 0.f ? -dampratio / dmax : b;
  k = timeconst <= 0.f ? -timeconst / (dmax * dmax) : k;
  imp = im, bcoef = b, kcoef = k;
}

// ---------------------------------------------------------------------------
// contact geometry of candidate c: MJX plane-sphere / plane-capsule

template <class T>
BGX_D void candidate_center(const float* __restrict__ M, int c, const V3* xp,
                            const Q4* xr, V3& ctr, V3& axis) {
  const int l = T::CAND_LINK[c];
  const float* cd = cand_rec<T>(M, c);
  // geom world pose (contact.py:42-53)
  const V3 gp = xp[l] + rotate(ldv3(cd + CD_GPOS), xr[l]);
  ctr = gp;
  axis = v3zero();
  if (T::CAND_KIND[c] == KIND_CAPSULE) {
    const Q4 gq = qmul(xr[l], ldq4(cd + CD_GQUAT));
    axis = quat_zaxis(gq);
    ctr = gp + axis * (cd[CD_HALFLEN] * cd[CD_SIGN]);
  }
}

template <class T>
BGX_D float candidate_dist(const float* __restrict__ M, int c, V3 ctr) {
  const float* cd = cand_rec<T>(M, c);
  return dot(ctr - ldv3(cd + CD_PPOS), ldv3(cd + CD_PN)) - cd[CD_RADIUS];
}

// ---------------------------------------------------------------------------
// constraint solve (constraint.py:194-240 with jaxopt ProjectedGradient)
// and integration (integrator.py:46-84) with KS contact slots

template <class T, int KS>
BGX_D void solve_and_integrate(const float* __restrict__ M, float* q,
                               float* qd, const Tree<T>& t, const V3* xp,
                               const Q4* xr, const V3* cda, const V3* cdv,
                               const float* qf_smooth, float* lds,
                               float* dbg) {
  constexpr int NV = T::NV, L = T::L;
  constexpr int NL = T::NLIM > 0 ? T::NLIM : 1;

  // unconstrained acceleration u = M^-1 qf_smooth and its link velocities
  float u[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) u[d] = qf_smooth[d];
  ltl_solve<T>(t, u);
  V3 ua[L], uv[L];
  link_velocities<T>(t, u, ua, uv);

  // ---- contact slots (constraint.py:135-176)
  int sl_link[KS];
  V3 sl_r[KS], sl_d0[KS], sl_d1[KS], sl_d2[KS];
  float sl_diag[KS], sl_b[KS][4];
  int n_c = 0;
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    V3 ctr, axis;
    candidate_center<T>(M, c, xp, xr, ctr, axis);
    const float dist = candidate_dist<T>(M, c, ctr);
    if (dist < 0.f) {
      const int l = T::CAND_LINK[c];
      const float* cd = cand_rec<T>(M, c);
      const V3 n = ldv3(cd + CD_PN);
      const float rad = cd[CD_RADIUS];
      const V3 pos = ctr - n * (rad + 0.5f * dist);
      V3 t1 = ldv3(cd + CD_T1), t2 = ldv3(cd + CD_T2);
      if (T::CAND_KIND[c] == KIND_CAPSULE) {
        // frame aligned with the capsule axis, fallback when near-normal
        float bn;
        V3 b = vnormalize(axis - n * dot(n, axis), &bn);
        const V3 alt = (-0.5f < n.y && n.y < 0.5f) ? v3(0.f, 1.f, 0.f)
                                                    : v3(0.f, 0.f, 1.f);
        t1 = bn < 0.5f ? alt : b;
        t2 = cross(n, t1);
      }
      const float mu = cd[CD_MU];
      const V3 r = pos - t.rc[l];
      // row velocities J qd and J M^-1 qf_smooth from link velocities
      const V3 vq = cdv[l] + cross(cda[l], r);
      const V3 vu = uv[l] + cross(ua[l], r);
      const V3 d1 = t1 * mu, d2 = t2 * mu;
      float sp[7];
#pragma unroll
      for (int k = 0; k < 2; ++k) sp[k] = cd[CD_SOLREF + k];
#pragma unroll
      for (int k = 0; k < 5; ++k) sp[2 + k] = cd[CD_SOLIMP + k];
      float imp, bc, kc;
      imp_aref(sp, dist, imp, bc, kc);
      const float tw = cd[CD_TW];
      const float diag = (tw + mu * mu * tw) * (2.f * mu * mu * (1.f - imp) / (imp + 1e-8f));
      const float e0q = dot(n, vq), e1q = dot(d1, vq), e2q = dot(d2, vq);
      const float e0u = dot(n, vu), e1u = dot(d1, vu), e2u = dot(d2, vu);
      const float velr[4] = {e0q + e1q, e0q - e1q, e0q + e2q, e0q - e2q};
      const float jmu[4] = {e0u + e1u, e0u - e1u, e0u + e2u, e0u - e2u};
      float bb[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float aref = -bc * velr[k] - kc * imp * dist;
        bb[k] = jmu[k] - aref;
      }
      const int s = n_c;
#pragma unroll
      for (int ss = 0; ss < KS; ++ss) {
        if (ss <= c && ss == s) {
          sl_link[ss] = l;
          sl_r[ss] = r;
          sl_d0[ss] = n;
          sl_d1[ss] = d1;
          sl_d2[ss] = d2;
          sl_diag[ss] = diag;
#pragma unroll
          for (int k = 0; k < 4; ++k) sl_b[ss][k] = bb[k];
        }
      }
      n_c += 1;
    }
  }
  if (n_c > KS) n_c = KS;  // callers choose KS >= active count

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
    float imp, bc, kc;
    float sp[7];
#pragma unroll
    for (int k = 0; k < 2; ++k) sp[k] = df[DF_SOLREF + k];
#pragma unroll
    for (int k = 0; k < 5; ++k) sp[2 + k] = df[DF_SOLIMP + k];
    imp_aref(sp, pos, imp, bc, kc);
    const float vel = side * qd[d];
    const float aref = act ? (-bc * vel - kc * imp * pos) : 0.f;
    lside[i] = side;
    ldiag[i] = act ? df[DF_INVWEIGHT] * (1.f - imp) / (imp + 1e-8f) : 0.f;
    lb[i] = side * u[d] - aref;
    any_lim = any_lim || act;
  }

  // ---- products with A = J M^-1 J^T + diag, evaluated through the tree
  // link scratch in LDS: element e of this lane at lds[e * WAVE]
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
        const float v0 = in_slot(s, 0), v1 = in_slot(s, 1);
        const float v2 = in_slot(s, 2), v3_ = in_slot(s, 3);
        const V3 w = sl_d0[s] * (v0 + v1 + v2 + v3_) + sl_d1[s] * (v0 - v1) +
                     sl_d2[s] * (v2 - v3_);
        const V3 ta = cross(sl_r[s], w);
        float* f = lds + sl_link[s] * 6 * WAVE;
        f[0 * WAVE] += ta.x;
        f[1 * WAVE] += ta.y;
        f[2 * WAVE] += ta.z;
        f[3 * WAVE] += w.x;
        f[4 * WAVE] += w.y;
        f[5 * WAVE] += w.z;
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

  auto apply_a = [&](auto in_slot, auto in_lim, auto out_slot, auto out_lim) {
    float z[NV];
    apply_jt(in_slot, in_lim, z);
    ltl_solve<T>(t, z);
    V3 va[L], vv[L];
    link_velocities<T>(t, z, va, vv);
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (hosts_contact<T>(l)) {
        float* f = lds + l * 6 * WAVE;
        f[0] = va[l].x, f[WAVE] = va[l].y, f[2 * WAVE] = va[l].z;
        f[3 * WAVE] = vv[l].x, f[4 * WAVE] = vv[l].y, f[5 * WAVE] = vv[l].z;
      }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (s < n_c) {
        const float* f = lds + sl_link[s] * 6 * WAVE;
        const V3 a = v3(f[0], f[WAVE], f[2 * WAVE]);
        const V3 v = v3(f[3 * WAVE], f[4 * WAVE], f[5 * WAVE]);
        const V3 vp = v + cross(a, sl_r[s]);
        const float e0 = dot(sl_d0[s], vp), e1 = dot(sl_d1[s], vp),
                    e2 = dot(sl_d2[s], vp);
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
  float sx[KS][4], sy[KS][4], sg[KS][4], sr[KS][4];
  float lx[NL], ly[NL], lg[NL], lr[NL];
#pragma unroll
  for (int s = 0; s < KS; ++s)
#pragma unroll
    for (int k = 0; k < 4; ++k) sx[s][k] = sy[s][k] = sg[s][k] = sr[s][k] = 0.f;
#pragma unroll
  for (int i = 0; i < NL; ++i) lx[i] = ly[i] = lg[i] = lr[i] = 0.f;

  const bool has_rows = (n_c > 0) || any_lim;
  const int maxiter = (int)M[H_MAXITER], maxls = (int)M[H_MAXLS];
  const float tol = M[H_TOL], lseps = M[H_LSEPS];
  float tk = 1.f, eta = 1.f, err = INFINITY;
  bool running = has_rows;
  int iters = 0, trials = 0;
  for (int it = 0; it < maxiter; ++it) {
    if (it > 0) running = running && (err > tol);
    if (!running) break;
    // gradient at y: r = A y + b, g = A^T r
    float fy = 0.f;
    apply_a([&](int s, int k) { return sy[s][k]; },
            [&](int i) { return ly[i]; },
            [&](int s, int k, float o) {
              const float r = o + sl_b[s][k];
              sr[s][k] = r;
              fy += 0.5f * r * r;
            },
            [&](int i, float o) {
              const float r = o + lb[i];
              lr[i] = r;
              fy += 0.5f * r * r;
            });
    apply_a([&](int s, int k) { return sr[s][k]; },
            [&](int i) { return lr[i]; },
            [&](int s, int k, float o) { sg[s][k] = o; },
            [&](int i, float o) { lg[i] = o; });
    // line search: trial x+ = relu(y - eta g), evaluated into sr/lr
    bool cond = true;
    for (int ls = 0; ls <= maxls && cond; ++ls) {
      if (ls > 0) eta *= 0.5f;
      const float e_ = eta;
      float fx = 0.f, dg = 0.f, dd = 0.f;
      apply_a([&](int s, int k) { return fmaxf(sy[s][k] - e_ * sg[s][k], 0.f); },
              [&](int i) { return fmaxf(ly[i] - e_ * lg[i], 0.f); },
              [&](int s, int k, float o) {
                const float r = o + sl_b[s][k];
                sr[s][k] = r;
                fx += 0.5f * r * r;
                const float xp = fmaxf(sy[s][k] - e_ * sg[s][k], 0.f);
                const float df = xp - sy[s][k];
                dg += df * sg[s][k];
                dd += df * df;
              },
              [&](int i, float o) {
                const float r = o + lb[i];
                lr[i] = r;
                fx += 0.5f * r * r;
                const float xp = fmaxf(ly[i] - e_ * lg[i], 0.f);
                const float df = xp - ly[i];
                dg += df * lg[i];
                dd += df * df;
              });
      trials += 1;
      cond = e_ * (fx - fy) > e_ * dg + 0.5f * dd + lseps;
    }
    // error at x+: |relu(x+ - grad f(x+)) - x+| with grad = A^T (A x+ + b)
    float e2 = 0.f;
    const float e_ = eta;
    apply_a([&](int s, int k) { return sr[s][k]; },
            [&](int i) { return lr[i]; },
            [&](int s, int k, float o) {
              const float xp = fmaxf(sy[s][k] - e_ * sg[s][k], 0.f);
              const float dv = fmaxf(xp - o, 0.f) - xp;
              e2 += dv * dv;
            },
            [&](int i, float o) {
              const float xp = fmaxf(ly[i] - e_ * lg[i], 0.f);
              const float dv = fmaxf(xp - o, 0.f) - xp;
              e2 += dv * dv;
            });
    err = sqrtf(e2);
    const float tn = 0.5f * (1.f + sqrtf(1.f + 4.f * tk * tk));
    const float beta = (tk - 1.f) / tn;
#pragma unroll
    for (int s = 0; s < KS; ++s)
      if (s < n_c)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float xp = fmaxf(sy[s][k] - e_ * sg[s][k], 0.f);
          sy[s][k] = xp + beta * (xp - sx[s][k]);
          sx[s][k] = xp;
        }
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) {
      const float xp = fmaxf(ly[i] - e_ * lg[i], 0.f);
      ly[i] = xp + beta * (xp - lx[i]);
      lx[i] = xp;
    }
    eta = eta <= 1e-6f ? 1.f : eta / 0.5f;
    tk = tn;
    iters += 1;
  }

  // ---- qdd = M^-1 (qf_smooth + J^T x) (integrator.py:58-64)
  float qdd[NV];
  if (has_rows) {
    apply_jt([&](int s, int k) { return sx[s][k]; },
             [&](int i) { return lx[i]; }, qdd);
    ltl_solve<T>(t, qdd);
#pragma unroll
    for (int d = 0; d < NV; ++d) qdd[d] = u[d] + qdd[d];
  } else {
#pragma unroll
    for (int d = 0; d < NV; ++d) qdd[d] = u[d];
  }
  if (dbg) {  // per-env probe of the last substep: qf_smooth, qdd, counts
    int o = 0;
#pragma unroll
    for (int d = 0; d < NV; ++d) dbg[o++] = qf_smooth[d];
#pragma unroll
    for (int d = 0; d < NV; ++d) dbg[o++] = qdd[d];
    dbg[o++] = (float)n_c;
    dbg[o++] = (float)iters;
    dbg[o++] = (float)trials;
  }

  // ---- semi-implicit Euler (integrator.py:26-84)
  const float dt = M[H_DT];
#pragma unroll
  for (int d = 0; d < NV; ++d) qd[d] = qd[d] + qdd[d] * dt;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int qa = T::QADR[l], da = T::DADR[l];
    if (T::LTYPE[l] == 0) {
      const V3 ang = v3(qd[da + 3], qd[da + 4], qd[da + 5]);
      const float an = sqrtf(dot(ang, ang)) + 1e-8f;
      const Q4 qr = qrot_axis(v3(ang.x / an, ang.y / an, ang.z / an), dt * an);
      Q4 rot = qmul(Q4{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]}, qr);
      const float rn = sqrtf(rot.w * rot.w + rot.x * rot.x + rot.y * rot.y + rot.z * rot.z);
      q[qa + 3] = rot.w / rn;
      q[qa + 4] = rot.x / rn;
      q[qa + 5] = rot.y / rn;
      q[qa + 6] = rot.z / rn;
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
// one substep: pipeline.step (pipeline.py:64-94)

template <class T>
BGX_D int substep(const float* __restrict__ M, float* q, float* qd,
                  const float* act, float* lds, float* dbg) {
  constexpr int NV = T::NV, L = T::L;
  Tree<T> t;

  // kinematics at (q, qd)
  V3 xp[L];
  Q4 xr[L];
  forward_kinematics<T, false>(M, q, qd, xp, xr, nullptr, nullptr);

  // dynamics.transform_com (dynamics.py:27-134)
  V3 xip[L];
  Q4 xir[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const float* lk = link_rec<T>(M, l);
    xip[l] = xp[l] + rotate(ldv3(lk + LK_IPOS), xr[l]);
    xir[l] = qmul(xr[l], ldq4(lk + LK_IROT));
  }
#pragma unroll
  for (int l = 0; l < L; ++l) {
    if (T::ROOT[l] == l) {
      V3 s = v3zero();
      float ms = 0.f;
#pragma unroll
      for (int l2 = 0; l2 < L; ++l2)
        if (T::ROOT[l2] == l) {
          const float m = link_rec<T>(M, l2)[LK_MASS];
          s = s + xip[l2] * m;
          ms = ms + m;
        }
      t.rc[l] = v3(s.x / ms, s.y / ms, s.z / ms);
    }
  }
#pragma unroll
  for (int l = 0; l < L; ++l) t.rc[l] = t.rc[T::ROOT[l]];
  V3 cc[L];
  Sym3 ci[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const float* lk = link_rec<T>(M, l);
    const float m = lk[LK_MASS];
    const V3 p = xip[l] - t.rc[l];
    float R[3][3];
    quat_to_3x3(xir[l], R);
    const Sym3 I0 = ldsym(lk + LK_I);
    // R I R^T
    float RI[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const V3 row = v3(R[a][0], R[a][1], R[a][2]);
      const V3 c = symmul(I0, row);
      RI[a][0] = c.x, RI[a][1] = c.y, RI[a][2] = c.z;
    }
    auto rir = [&](int a, int b) {
      return RI[a][0] * R[b][0] + RI[a][1] * R[b][1] + RI[a][2] * R[b][2];
    };
    const float pp = dot(p, p);
    ci[l] = Sym3{rir(0, 0) + (pp - p.x * p.x) * m, rir(1, 1) + (pp - p.y * p.y) * m,
                 rir(2, 2) + (pp - p.z * p.z) * m, rir(0, 1) - p.x * p.y * m,
                 rir(0, 2) - p.x * p.z * m, rir(1, 2) - p.y * p.z * m};
    cc[l] = p * m;
  }

  // cdof: dof motions in the root-COM frame (dynamics.py:47-89)
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const float* lk = link_rec<T>(M, l);
    const int par = T::PARENT[l], da = T::DADR[l], qa = T::QADR[l];
    V3 Pp = v3zero();
    Q4 Pr = qident();
    if (T::LTYPE[l] == 0) {
      Pp = xp[l], Pr = xr[l];
    } else if (par >= 0) {
      Pp = xp[par], Pr = xr[par];
    }
    const V3 t1p = Pp + rotate(ldv3(lk + LK_TPOS), Pr);
    const Q4 jr = qmul(Pr, ldq4(lk + LK_TROT));
    const V3 jp = t1p + rotate(ldv3(lk + LK_JPOS), jr);
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
    V3 accp = v3zero();
    Q4 accr = qident();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = da + k;
        const float* df = dof_rec<T>(M, d);
        const V3 ang = ldv3(df + DF_ANG), vel = ldv3(df + DF_VEL);
        V3 a = ang, v = vel;
        if (T::LTYPE[l] != 0) {
          a = rotate(ang, accr);
          v = rotate(vel, accr) + cross(accp, a);
          const Q4 rk = qnormalize(qrot_axis(ang, q[qa + k]));
          accp = accp + rotate(vel * q[qa + k], accr);
          accr = qmul(accr, rk);
        }
        a = rotate(a, jr);
        const V3 off = t.rc[l] - jp;
        t.ca[d] = a;
        t.cv[d] = v - cross(off, a);
      }
    }
  }

  // cd: link velocities in the root-COM frame (dynamics.py:90-103)
  V3 cda[L], cdv[L];
  {
    float qv[NV];
#pragma unroll
    for (int d = 0; d < NV; ++d) qv[d] = qd[d];
    link_velocities<T>(t, qv, cda, cdv);
  }

  // cdofd (dynamics.py:105-130)
  V3 dda[NV], ddv[NV];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int par = T::PARENT[l], da = T::DADR[l];
    if (T::LTYPE[l] == 0) {
      V3 sa = v3zero(), sv = v3zero();
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sa = sa + t.ca[da + k] * qd[da + k];
        sv = sv + t.cv[da + k] * qd[da + k];
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (k < 3) {
          dda[da + k] = v3zero(), ddv[da + k] = v3zero();
        } else {
          cross_motion(sa, sv, t.ca[da + k], t.cv[da + k], dda[da + k], ddv[da + k]);
        }
      }
    } else {
      V3 sa = par >= 0 ? cda[par] : v3zero();
      V3 sv = par >= 0 ? cdv[par] : v3zero();
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k < T::LTYPE[l]) {
          const int d = da + k;
          if (k > 0) {
            sa = sa + t.ca[d - 1] * qd[d - 1];
            sv = sv + t.cv[d - 1] * qd[d - 1];
          }
          cross_motion(sa, sv, t.ca[d], t.cv[d], dda[d], ddv[d]);
        }
      }
    }
  }

  // dynamics.inverse: RNEA bias (dynamics.py:137-184)
  float bias[NV];
  {
    const V3 grav = ldv3(M + H_GRAVITY);
    V3 aa[L], av[L], fa[L], fv[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int par = T::PARENT[l];
      V3 a = par < 0 ? v3zero() : aa[par];
      V3 v = par < 0 ? -grav : av[par];
      const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (k < nd) {
          const int d = T::DADR[l] + k;
          a = a + dda[d] * qd[d];
          v = v + ddv[d] * qd[d];
        }
      }
      aa[l] = a, av[l] = v;
      const float m = link_rec<T>(M, l)[LK_MASS];
      V3 f1a, f1v, f2a, f2v, f3a, f3v;
      inertia_mul(cc[l], ci[l], m, a, v, f1a, f1v);
      inertia_mul(cc[l], ci[l], m, cda[l], cdv[l], f2a, f2v);
      cross_force(cda[l], cdv[l], f2a, f2v, f3a, f3v);
      fa[l] = f1a + f3a, fv[l] = f1v + f3v;
    }
    project_forces<T>(t, fa, fv, bias);
  }

  // passive forces, actuators, qf_smooth (dynamics.py:187-236, actuator.py)
  float qf[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) {
    const float* df = dof_rec<T>(M, d);
    qf[d] = -df[DF_DAMPING] * qd[d];
  }
#pragma unroll
  for (int l = 0; l < L; ++l) {
    if (T::LTYPE[l] != 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (k < T::LTYPE[l]) {
          const int d = T::DADR[l] + k;
          qf[d] = -q[T::QADR[l] + k] * dof_rec<T>(M, d)[DF_STIFFNESS] + qf[d];
        }
    }
  }
  float tau[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) tau[d] = 0.f;
#pragma unroll
  for (int a = 0; a < T::NU; ++a) {
    const float* ac = M + T::OFF_ACT + a * ACT_STRIDE;
    const int qa = T::ACT_Q[a], da = T::ACT_D[a];
    const float ctl = clipf(act[a], ac[AC_CTRL_LO], ac[AC_CTRL_HI]);
    const float b = ac[AC_GEAR] * (q[qa] * ac[AC_BIAS_Q] + qd[da] * ac[AC_BIAS_QD]);
    float f = ac[AC_GAIN] * ctl + b;
    f = clipf(f, ac[AC_FORCE_LO], ac[AC_FORCE_HI]);
    tau[da] += f * ac[AC_GEAR];
  }
#pragma unroll
  for (int d = 0; d < NV; ++d) qf[d] = qf[d] - bias[d] + tau[d];

  // mass.matrix: CRBA on the dof-ancestor pattern (mass.py:27-83)
  {
    V3 kc[L];
    Sym3 ki[L];
    float km[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      kc[l] = cc[l], ki[l] = ci[l], km[l] = link_rec<T>(M, l)[LK_MASS];
    }
#pragma unroll
    for (int l = L - 1; l >= 0; --l) {
      const int par = T::PARENT[l];
      if (par >= 0) {
        kc[par] = kc[par] + kc[l];
        ki[par] = symadd(ki[par], ki[l]);
        km[par] = km[par] + km[l];
      }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int l = T::DOF_LINK[i];
      V3 fa, fv;
      inertia_mul(kc[l], ki[l], km[l], t.ca[i], t.cv[i], fa, fv);
      t.Lm[i][i] = dot(t.cv[i], fv) + dot(t.ca[i], fa) + dof_rec<T>(M, i)[DF_ARMATURE];
#pragma unroll
      for (int j = 0; j < i; ++j)
        if ((T::DANC[i] >> j) & 1u) t.Lm[i][j] = dot(t.cv[j], fv) + dot(t.ca[j], fa);
    }
  }
  // LTL factorization, leaves first
#pragma unroll
  for (int k = NV - 1; k >= 0; --k) {
    const float lkk = sqrtf(t.Lm[k][k]);
    t.Lm[k][k] = lkk;
    t.inv[k] = 1.f / lkk;
#pragma unroll
    for (int j = 0; j < k; ++j)
      if ((T::DANC[k] >> j) & 1u) t.Lm[k][j] *= t.inv[k];
#pragma unroll
    for (int i = 0; i < k; ++i) {
      if ((T::DANC[k] >> i) & 1u) {
        const float a = t.Lm[k][i];
        t.Lm[i][i] -= a * a;
#pragma unroll
        for (int j = 0; j < i; ++j)
          if ((T::DANC[i] >> j) & 1u) t.Lm[i][j] -= a * t.Lm[k][j];
      }
    }
  }

  // count active contacts to pick the slot capacity
  int n_act = 0;
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    V3 ctr, axis;
    candidate_center<T>(M, c, xp, xr, ctr, axis);
    n_act += candidate_dist<T>(M, c, ctr) < 0.f ? 1 : 0;
  }
  int status = 0;
  if (n_act <= T::KC) {
    solve_and_integrate<T, T::KC>(M, q, qd, t, xp, xr, cda, cdv, qf, lds, dbg);
  } else {
    solve_and_integrate<T, (T::NCAND > T::KC ? T::NCAND : T::KC)>(
        M, q, qd, t, xp, xr, cda, cdv, qf, lds, dbg);
    status |= ST_SLOW_PATH;
  }
  return status;
}

}  // namespace bgx
