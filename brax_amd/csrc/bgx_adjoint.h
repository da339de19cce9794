// Adjoint (vector-Jacobian product) of one generalized-coordinates substep.
//
// The reference differentiates pipeline.step with jax.grad through the
// unrolled projected-gradient solver (jaxopt implicit_diff=False,
// brax/generalized/constraint.py:229-238): the discrete decisions -- which
// contacts and limits are active, how many solver iterations and line-search
// halvings run, where the relu projection clamps -- are constants of the
// derivative, and the smooth arithmetic between them is differentiated.
// This file computes the same derivative in two stages:
//
//  A (one thread per env): recompute the substep, keeping the solver
//    iterates, then run the reverse sweep of the integrator, the solves with
//    M = L^T L and the unrolled FISTA iterations by hand.  Output: cotangents
//    of the position- and velocity-dependent ingredients -- qf_smooth, the
//    mass matrix (a symmetric matrix on the dof-ancestor pattern), the
//    contact rows of J, the diagonal regularizer d and aref -- plus the
//    action gradient and the velocity pass-through of qd' = qd + dt qdd.
//  B (one thread per env and input coordinate): run the same formulas that
//    build those ingredients from (q, qd) in dual numbers along one input
//    coordinate and contract the tangents with the cotangents of stage A:
//    d/de <cotangents, ingredients(q + e dq, qd + e dqd)>.
//
// Record of stage A for stage B: one float per field per env, field-major
// (field f of env e at rec[f * n_envs + e]).
#pragma once

#include "bgx_step.h"

namespace bgx {

constexpr int ADJ_MAXITER = 8;  // most solver iterations the adjoint unrolls

template <class T>
struct AdjRec {
  static constexpr int NV = T::NV, NQ = T::NQ;
  static constexpr int KS = T::NCAND > 0 ? T::NCAND : 1;  // one slot per candidate
  static constexpr int NRC = 4 * KS;                      // contact rows
  static constexpr int NLR = T::NLIM;                     // limit rows
  static constexpr int NR = NRC + (NLR > 0 ? NLR : 1);
  static constexpr int NMB = mpat_row<T>(T::NV);
  static constexpr int QFB = 0;               // [NV]  cotangent of qf_smooth
  static constexpr int MB = QFB + NV;         // [NMB] cotangent of M (symmetrized)
  static constexpr int JB = MB + NMB;         // [NRC][NV] cotangent of J contact rows
  static constexpr int DB = JB + NRC * NV;    // [NR]  cotangent of d
  static constexpr int AB = DB + NR;          // [NR]  cotangent of aref
  static constexpr int GQ1 = AB + NR;         // [NQ]  cotangent of q'
  static constexpr int QD1 = GQ1 + NQ;        // [NV]  qd'
  static constexpr int GQDT = QD1 + NV;       // [NV]  qd' cotangent incl. the q' update
  static constexpr int CSLOT = GQDT + NV;     // [KS]  slot of each candidate, -1 inactive
  static constexpr int CFB = CSLOT + KS;      // [KS]  capsule frame fallback taken
  static constexpr int SIZE = CFB + KS;
};

struct RecPtr {
  float* p;
  long long n;
  BGX_D float& operator[](int f) const { return p[(long long)f * n]; }
};

// ---------------------------------------------------------------------------
// stage A

template <class T>
BGX_D int adjoint_a(const float* __restrict__ M, const float* q_in, const float* qd_in,
                    const float* act, const float* gq1, const float* gqd1, RecPtr rec,
                    float* gact) {
  using R = AdjRec<T>;
  constexpr int NV = T::NV, NQ = T::NQ, L = T::L, KS = R::KS, NR = R::NR;
  constexpr int NRC = R::NRC;
  constexpr bool PZ = T::PLANE_Z != 0;
  float q[NQ], qd[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = q_in[k];
#pragma unroll
  for (int k = 0; k < NV; ++k) qd[k] = qd_in[k];

  // ---- recompute the substep, as substep_impl with a slot per candidate
  V3 xp[L];
  Q4 xr[L];
  forward_kinematics<T, false, float>(M, q, qd, xp, xr, nullptr, nullptr);
  const V3 rc = root_com<T>(M, xp, xr);
  Tree<T> t;
  float lm_store[T::LM_LDS ? R::NMB : 1];
  if constexpr (T::LM_LDS) t.lml = lm_store, t.lms = 1;
#pragma unroll
  for (int l = 0; l < L; ++l) t.rc[l] = rc;

  SlotGeo<PZ> sg[KS];
  float sdiag[KS], sbq[KS], skt[KS];
  int n_c = 0;
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    V3 ctr, axis;
    candidate_center<T>(M, c, xp, xr, ctr, axis);
    const float dist = candidate_dist<T>(M, c, ctr);
    int slot = -1, fb = 0;
    if (dist < 0.f) {
      const int l = T::CAND_LINK[c];
      const float* cd = cand_rec<T>(M, c);
      V3 pos, t1, t2;
      fb = contact_frame<T>(M, c, ctr, axis, dist, -1, pos, t1, t2) ? 1 : 0;
      const V3 n = PZ ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN);
      const float mu = cd[CD_MU];
      SlotGeo<PZ> g;
      slot_set(g, l, pos - rc, n, t1 * mu, t2 * mu);
      const float* im = cd + CD_IMP;
      const float imp = impedance(im, dist);
      const float tw = cd[CD_TW];
      const float diag = (tw + mu * mu * tw) * (2.f * mu * mu * imp_ratio(imp));
#pragma unroll
      for (int ss = 0; ss < KS; ++ss) {
        if (ss <= c && ss == n_c) {
          sg[ss] = g;
          sdiag[ss] = diag;
          sbq[ss] = im[IM_B];
          skt[ss] = im[IM_K] * imp * dist;
        }
      }
      slot = n_c;
      n_c += 1;
    }
    rec[R::CSLOT + c] = (float)slot;
    rec[R::CFB + c] = (float)fb;
  }

  V3 cc[L];
  Sym3 ci[L];
  com_inertia<T>(M, xp, xr, rc, cc, ci);
  dof_motions<T>(M, q, xp, xr, rc, t.ca, t.cv);
  V3 cda[L], cdv[L];
  link_vel<T>(t.ca, t.cv, qd, cda, cdv);
  float qf[NV];
  rnea_bias<T>(M, t.ca, t.cv, cda, cdv, cc, ci, qd, qf);
  smooth_forces<T>(M, q, qd, act, qf);
  crba<T>(M, cc, ci, t.ca, t.cv, [&](int i, int j, float v) { t.lm(i, j) = v; });
  ltl_factor<T>(t);
  float u[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) u[d] = qf[d];
  ltl_solve<T>(t, u);

  // rows: contact rows 4 s + k for slots s < n_c, then limit rows NRC + i
  float lside[NR], rdiag[NR], b[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) lside[r] = rdiag[r] = b[r] = 0.f;

  auto J_apply = [&](const float* z, float* out) {
    V3 va[L], vv[L];
    link_vel<T>(t.ca, t.cv, z, va, vv);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      float e0 = 0.f, e1 = 0.f, e2 = 0.f;
      if (s < n_c) {
        V3 a = va[0], v = vv[0];
#pragma unroll
        for (int l = 1; l < L; ++l)
          if (sg[s].link == l) a = va[l], v = vv[l];
        slot_dots(sg[s], v + cross(a, sg[s].r), e0, e1, e2);
      }
      out[4 * s] = e0 + e1;
      out[4 * s + 1] = e0 - e1;
      out[4 * s + 2] = e0 + e2;
      out[4 * s + 3] = e0 - e2;
    }
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) out[NRC + i] = lside[NRC + i] * z[T::LIM_D[i]];
  };
  auto JT_apply = [&](const float* x, float* tau) {
    V3 fa[L], fv[L];
#pragma unroll
    for (int l = 0; l < L; ++l) fa[l] = v3zero(), fv[l] = v3zero();
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (s < n_c) {
        const V3 w = slot_force(sg[s], x[4 * s], x[4 * s + 1], x[4 * s + 2], x[4 * s + 3]);
        const V3 ta = cross(sg[s].r, w);
#pragma unroll
        for (int l = 0; l < L; ++l)
          if (hosts_contact<T>(l) && sg[s].link == l) fa[l] = fa[l] + ta, fv[l] = fv[l] + w;
      }
    }
    project_forces<T>(t, fa, fv, tau);
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) tau[T::LIM_D[i]] += lside[NRC + i] * x[NRC + i];
  };
  // out = A x, ux = M^-1 J^T x
  auto A_apply = [&](const float* x, float* out, float* ux) {
    JT_apply(x, ux);
    ltl_solve<T>(t, ux);
    J_apply(ux, out);
#pragma unroll
    for (int r = 0; r < NR; ++r) out[r] = out[r] + rdiag[r] * x[r];
  };

  {
    float ju[NR], jq[NR];
    J_apply(u, ju);
    J_apply(qd, jq);
#pragma unroll
    for (int s = 0; s < KS; ++s)
      if (s < n_c)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          b[4 * s + k] = ju[4 * s + k] + (sbq[s] * jq[4 * s + k] + skt[s]);
          rdiag[4 * s + k] = sdiag[s];
        }
  }
  bool any_lim = false;
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i) {
    const int d = T::LIM_D[i], qi = T::LIM_Q[i];
    const float* df = dof_rec<T>(M, d);
    const float pmin = q[qi] - df[DF_LO];
    const float pmax = df[DF_HI] - q[qi];
    const float pos = fminf(fminf(pmin, pmax), 0.f);
    const bool on = pos < 0.f;
    const float side = on ? (pmin < pmax ? 1.f : -1.f) : 0.f;
    const float* im = df + DF_IMP;
    const float imp = impedance(im, pos);
    const float vel = side * qd[d];
    const float aref = on ? (-im[IM_B] * vel - im[IM_K] * imp * pos) : 0.f;
    lside[NRC + i] = side;
    rdiag[NRC + i] = on ? df[DF_INVWEIGHT] * imp_ratio(imp) : 0.f;
    b[NRC + i] = side * u[d] - aref;
    any_lim = any_lim || on;
  }

  // ---- projected gradient, keeping y_k, r_k = A y_k + b, x_{k+1}, eta_k, beta_k
  float X[NR], Y[NR], G[NR], Rr[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) X[r] = Y[r] = G[r] = Rr[r] = 0.f;
  float Ys[ADJ_MAXITER][NR], Rs[ADJ_MAXITER][NR], Xs[ADJ_MAXITER][NR];
  float etas[ADJ_MAXITER], betas[ADJ_MAXITER];
  const bool has_rows = (n_c > 0) || any_lim;
  const int maxiter = (int)M[H_MAXITER], maxls = (int)M[H_MAXLS];
  if (has_rows && maxiter > ADJ_MAXITER) return -1;
  const float tol = M[H_TOL], lseps = M[H_LSEPS];
  float tk = 1.f, eta = 1.f, err = INFINITY;
  bool running = has_rows;
  int iters = 0;
  for (int it = 0; it < maxiter; ++it) {
    if (it > 0) running = running && (err > tol);
    if (!running) break;
    float tmp[NV], ar[NR];
    A_apply(Y, ar, tmp);
    float fy = 0.f;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      Rr[r] = ar[r] + b[r];
      fy += 0.5f * Rr[r] * Rr[r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) Ys[it][r] = Y[r], Rs[it][r] = Rr[r];
    A_apply(Rr, G, tmp);
    bool cond = true;
    for (int ls = 0; ls <= maxls && cond; ++ls) {
      if (ls > 0) eta *= 0.5f;
      const float e_ = eta;
      float xt[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) xt[r] = fmaxf(Y[r] - e_ * G[r], 0.f);
      A_apply(xt, ar, tmp);
      float fx = 0.f, dg = 0.f, dd = 0.f;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const float rr = ar[r] + b[r];
        Rr[r] = rr;
        fx += 0.5f * rr * rr;
        const float df = xt[r] - Y[r];
        dg += df * G[r];
        dd += df * df;
      }
      cond = e_ * (fx - fy) > e_ * dg + 0.5f * dd + lseps;
    }
    const float e_ = eta;
    float xn[NR], gx[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) xn[r] = fmaxf(Y[r] - e_ * G[r], 0.f);
    A_apply(Rr, gx, tmp);
    float e2 = 0.f;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const float dv = fmaxf(xn[r] - gx[r], 0.f) - xn[r];
      e2 += dv * dv;
    }
    err = fsqrt(e2);
    const float tn = 0.5f * (1.f + fsqrt(1.f + 4.f * tk * tk));
    const float beta = (tk - 1.f) * rcp(tn);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      Xs[it][r] = xn[r];
      Y[r] = xn[r] + beta * (xn[r] - X[r]);
      X[r] = xn[r];
    }
    etas[it] = e_;
    betas[it] = beta;
    eta = eta <= 1e-6f ? 1.f : eta * 2.f;
    tk = tn;
    iters += 1;
  }

  // qdd = u + w, w = M^-1 J^T x; qd' = qd + qdd dt
  float w[NV];
  if (has_rows) {
    JT_apply(X, w);
    ltl_solve<T>(t, w);
  } else {
#pragma unroll
    for (int d = 0; d < NV; ++d) w[d] = 0.f;
  }
  const float dt = M[H_DT];
  float qd1[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) qd1[d] = qd[d] + (u[d] + w[d]) * dt;

  // ---- reverse sweep
  // q' = f(q, qd'): cotangent of qd' from the position update
  float gqdt[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) gqdt[d] = gqd1[d];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int qa = T::QADR[l], da = T::DADR[l];
    if (T::LTYPE[l] == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) gqdt[da + k] += dt * gq1[qa + k];
      const Q4T<Dual> rot = lift4<Dual>(Q4{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]});
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const V3T<Dual> ang{dual(qd1[da + 3], k == 0 ? 1.f : 0.f),
                            dual(qd1[da + 4], k == 1 ? 1.f : 0.f),
                            dual(qd1[da + 5], k == 2 ? 1.f : 0.f)};
        const Q4T<Dual> r1 = integrate_rot(rot, ang, dt);
        gqdt[da + 3 + k] += gq1[qa + 3] * r1.w.d + gq1[qa + 4] * r1.x.d +
                            gq1[qa + 5] * r1.y.d + gq1[qa + 6] * r1.z.d;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (k < T::LTYPE[l]) gqdt[da + k] += dt * gq1[qa + k];
    }
  }

  float Mb[R::NMB];
#pragma unroll
  for (int k = 0; k < R::NMB; ++k) Mb[k] = 0.f;
  // Mbar -= sym(a c^T) on the pattern
  auto mbar_sub = [&](const float* a, const float* c) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      Mb[mpat_idx<T>(i, i)] -= a[i] * c[i];
#pragma unroll
      for (int j = 0; j < i; ++j)
        if ((T::DANC[i] >> j) & 1u) Mb[mpat_idx<T>(i, j)] -= a[i] * c[j] + a[j] * c[i];
    }
  };
  // Jbar += a c^T on the contact rows (kept in the record)
#pragma unroll
  for (int s = 0; s < KS; ++s)
    if (s < n_c)
#pragma unroll
      for (int k = 0; k < 4 * NV; ++k) rec[R::JB + 4 * s * NV + k] = 0.f;
  auto jbar_add = [&](const float* a, const float* c) {
#pragma unroll
    for (int s = 0; s < KS; ++s)
      if (s < n_c)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float ar = a[4 * s + k];
#pragma unroll
          for (int d = 0; d < NV; ++d) rec[R::JB + (4 * s + k) * NV + d] += ar * c[d];
        }
  };
  float Db[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) Db[r] = 0.f;

  // qdd = u + w: u-bar = qdd-bar; w = M^-1 J^T x
  float ub[NV], z[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) ub[d] = z[d] = dt * gqdt[d];
  float xb[NR], yb[NR], bb[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) xb[r] = yb[r] = bb[r] = 0.f;
  if (has_rows) {
    ltl_solve<T>(t, z);
    mbar_sub(z, w);
    J_apply(z, xb);
    jbar_add(X, z);
  }
  // unrolled FISTA, last iteration first
  for (int k = iters - 1; k >= 0; --k) {
    const float bk = betas[k], ek = etas[k];
    float wb[NR], gb[NR], rb[NR], ay[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const float xt = xb[r] + (1.f + bk) * yb[r];
      xb[r] = -bk * yb[r];
      wb[r] = Xs[k][r] > 0.f ? xt : 0.f;
      gb[r] = -ek * wb[r];
    }
    float ug[NV], ur[NV], urb[NV];
    A_apply(gb, rb, ug);  // r-bar = A g-bar
    float rk[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) rk[r] = Rs[k][r];
    JT_apply(rk, ur);
    ltl_solve<T>(t, ur);
    jbar_add(gb, ur);
    jbar_add(rk, ug);
    mbar_sub(ug, ur);
#pragma unroll
    for (int r = 0; r < NR; ++r) Db[r] += gb[r] * rk[r];
    A_apply(rb, ay, urb);  // y-bar = w-bar + A r-bar
    if (k > 0) {
      float yk[NR], uy[NV];
#pragma unroll
      for (int r = 0; r < NR; ++r) yk[r] = Ys[k][r];
      JT_apply(yk, uy);
      ltl_solve<T>(t, uy);
      jbar_add(rb, uy);
      jbar_add(yk, urb);
      mbar_sub(urb, uy);
#pragma unroll
      for (int r = 0; r < NR; ++r) Db[r] += rb[r] * yk[r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      yb[r] = wb[r] + ay[r];
      bb[r] += rb[r];
    }
  }
  // b = J u - aref
  {
    float jtb[NV];
    JT_apply(bb, jtb);
#pragma unroll
    for (int d = 0; d < NV; ++d) ub[d] += jtb[d];
    jbar_add(bb, u);
  }
  // u = M^-1 qf
  float qfb[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) qfb[d] = ub[d];
  ltl_solve<T>(t, qfb);
  mbar_sub(qfb, u);
  // actuators: qf[da] += clip(gain clip(act) + bias, force) gear
#pragma unroll
  for (int a = 0; a < T::NU; ++a) {
    const float* ac = M + T::OFF_ACT + a * ACT_STRIDE;
    const int qa = T::ACT_Q[a], da = T::ACT_D[a];
    const bool c_in = act[a] >= ac[AC_CTRL_LO] && act[a] <= ac[AC_CTRL_HI];
    const float ctl = clipf(act[a], ac[AC_CTRL_LO], ac[AC_CTRL_HI]);
    const float f = ac[AC_GAIN] * ctl +
                    ac[AC_GEAR] * (q[qa] * ac[AC_BIAS_Q] + qd[da] * ac[AC_BIAS_QD]);
    const bool f_in = f >= ac[AC_FORCE_LO] && f <= ac[AC_FORCE_HI];
    if (c_in && f_in) gact[a] += qfb[da] * ac[AC_GEAR] * ac[AC_GAIN];
  }

#pragma unroll
  for (int d = 0; d < NV; ++d) rec[R::QFB + d] = qfb[d];
#pragma unroll
  for (int k = 0; k < R::NMB; ++k) rec[R::MB + k] = Mb[k];
#pragma unroll
  for (int r = 0; r < NR; ++r) rec[R::DB + r] = Db[r], rec[R::AB + r] = -bb[r];
#pragma unroll
  for (int k = 0; k < NQ; ++k) rec[R::GQ1 + k] = gq1[k];
#pragma unroll
  for (int d = 0; d < NV; ++d) rec[R::QD1 + d] = qd1[d], rec[R::GQDT + d] = gqdt[d];
  return n_c > T::KC ? ST_SLOW_PATH : 0;
}

// ---------------------------------------------------------------------------
// stage B: d/de of <cotangents, ingredients> along input coordinate dir
// (dir < NQ: q[dir]; else qd[dir - NQ]); returns the input cotangent

template <class T>
BGX_D float adjoint_b(const float* __restrict__ M, const float* q_in, const float* qd_in,
                      const float* act, RecPtr rec, int dir) {
  using R = AdjRec<T>;
  using S = Dual;
  constexpr int NV = T::NV, NQ = T::NQ, L = T::L, NRC = R::NRC;
  constexpr bool PZ = T::PLANE_Z != 0;
  S q[NQ], qd[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = dual(q_in[k], k == dir ? 1.f : 0.f);
#pragma unroll
  for (int k = 0; k < NV; ++k) qd[k] = dual(qd_in[k], NQ + k == dir ? 1.f : 0.f);

  V3T<S> xp[L];
  Q4T<S> xr[L];
  forward_kinematics<T, false, S>(M, q, qd, xp, xr, nullptr, nullptr);
  const V3T<S> rc = root_com<T>(M, xp, xr);
  V3T<S> cc[L], ca[NV], cv[NV];
  Sym3T<S> ci[L];
  com_inertia<T>(M, xp, xr, rc, cc, ci);
  dof_motions<T>(M, q, xp, xr, rc, ca, cv);
  V3T<S> cda[L], cdv[L];
  link_vel<T>(ca, cv, qd, cda, cdv);
  S qf[NV];
  rnea_bias<T>(M, ca, cv, cda, cdv, cc, ci, qd, qf);
  smooth_forces<T>(M, q, qd, act, qf);
  float phi = 0.f;
#pragma unroll
  for (int d = 0; d < NV; ++d) phi += rec[R::QFB + d] * qf[d].d;
  {
    S Ml[NV][NV];
    crba<T>(M, cc, ci, ca, cv, [&](int i, int j, S v) { Ml[i][j] = v; });
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      phi += rec[R::MB + mpat_idx<T>(i, i)] * Ml[i][i].d;
#pragma unroll
      for (int j = 0; j < i; ++j)
        if ((T::DANC[i] >> j) & 1u) phi += rec[R::MB + mpat_idx<T>(i, j)] * Ml[i][j].d;
    }
  }

  // contact rows of the active candidates
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    const int s = (int)rec[R::CSLOT + c];
    if (s >= 0) {
      const int l = T::CAND_LINK[c];
      const float* cd = cand_rec<T>(M, c);
      V3T<S> ctr, axis, pos, t1, t2;
      candidate_center<T>(M, c, xp, xr, ctr, axis);
      const S dist = candidate_dist<T>(M, c, ctr);
      contact_frame<T>(M, c, ctr, axis, dist, (int)rec[R::CFB + c], pos, t1, t2);
      const V3T<S> n = lift3<S>(PZ ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN));
      const float mu = cd[CD_MU];
      const V3T<S> r = pos - rc;
      const float* im = cd + CD_IMP;
      const S imp = impedance(im, dist);
      const float tw = cd[CD_TW];
      const S diag = (tw + mu * mu * tw) * (2.f * mu * mu * imp_ratio(imp));
      const S kt = im[IM_K] * imp * dist;
      const float bq = im[IM_B];
      const V3T<S> vq = cdv[l] + cross(cda[l], r);
      const S e0 = dot(n, vq), e1 = mu * dot(t1, vq), e2 = mu * dot(t2, vq);
      const int r0 = 4 * s;
      phi += rec[R::AB + r0] * -(bq * (e0.d + e1.d) + kt.d);
      phi += rec[R::AB + r0 + 1] * -(bq * (e0.d - e1.d) + kt.d);
      phi += rec[R::AB + r0 + 2] * -(bq * (e0.d + e2.d) + kt.d);
      phi += rec[R::AB + r0 + 3] * -(bq * (e0.d - e2.d) + kt.d);
      phi += ((rec[R::DB + r0] + rec[R::DB + r0 + 1]) +
              (rec[R::DB + r0 + 2] + rec[R::DB + r0 + 3])) * diag.d;
      const unsigned mask = link_dof_mask<T>(T::CAND_LINK[c]);
#pragma unroll
      for (int d = 0; d < NV; ++d) {
        if ((mask >> d) & 1u) {
          const V3T<S> pv = cv[d] + cross(ca[d], r);
          const float f0 = dot(n, pv).d, f1 = mu * dot(t1, pv).d, f2 = mu * dot(t2, pv).d;
          const int jb = R::JB + r0 * NV + d;
          phi += rec[jb] * (f0 + f1) + rec[jb + NV] * (f0 - f1) +
                 rec[jb + 2 * NV] * (f0 + f2) + rec[jb + 3 * NV] * (f0 - f2);
        }
      }
    }
  }

  // limit rows
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i) {
    const int d = T::LIM_D[i], qi = T::LIM_Q[i];
    const float* df = dof_rec<T>(M, d);
    const S pmin = q[qi] - df[DF_LO];
    const S pmax = df[DF_HI] - q[qi];
    const S pos = min_(min_(pmin, pmax), lift<S>(0.f));
    if (pos.v < 0.f) {
      const float side = pmin.v < pmax.v ? 1.f : -1.f;
      const float* im = df + DF_IMP;
      const S imp = impedance(im, pos);
      const S diag = df[DF_INVWEIGHT] * imp_ratio(imp);
      const S aref = -im[IM_B] * (side * qd[d]) - im[IM_K] * imp * pos;
      phi += rec[R::DB + NRC + i] * diag.d + rec[R::AB + NRC + i] * aref.d;
    }
  }

  // position update q' = f(q, qd') at fixed qd'
  if (dir < NQ) {
    const float dt = M[H_DT];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int qa = T::QADR[l], da = T::DADR[l];
      if (T::LTYPE[l] == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) phi += rec[R::GQ1 + qa + k] * q[qa + k].d;
        const V3T<S> ang = lift3<S>(v3(rec[R::QD1 + da + 3], rec[R::QD1 + da + 4],
                                       rec[R::QD1 + da + 5]));
        const Q4T<S> r1 = integrate_rot(Q4T<S>{q[qa + 3], q[qa + 4], q[qa + 5], q[qa + 6]},
                                        ang, dt);
        phi += rec[R::GQ1 + qa + 3] * r1.w.d + rec[R::GQ1 + qa + 4] * r1.x.d +
               rec[R::GQ1 + qa + 5] * r1.y.d + rec[R::GQ1 + qa + 6] * r1.z.d;
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (k < T::LTYPE[l]) phi += rec[R::GQ1 + qa + k] * q[qa + k].d;
      }
    }
    return phi;
  }
  return phi + rec[R::GQDT + (dir - NQ)];
}

}  // namespace bgx
