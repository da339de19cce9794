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
//  A1 (one thread per env): the step's own substep (substep_impl, a slot per
//    candidate) recording its tape: tree, contact slots, accelerations and
//    every projected-gradient iterate.
//  A2 (one thread per env): the reverse sweep of the integrator, the solves
//    with M = L^T L and the unrolled FISTA iterations, by hand, from the
//    tape.  Output: cotangents of the position- and velocity-dependent
//    ingredients -- qf_smooth, the mass matrix (a symmetric matrix on the
//    dof-ancestor pattern), the contact rows of J, the diagonal regularizer
//    d and aref -- plus the action gradient and the velocity pass-through of
//    qd' = qd + dt qdd.
//  B (one thread per env and input coordinate): run the same formulas that
//    build those ingredients from (q, qd) in dual numbers along one input
//    coordinate and contract the tangents with the cotangents of stage A:
//    d/de <cotangents, ingredients(q + e dq, qd + e dqd)>.
//
// Record of stages A1/A2 for stage B: one float per field per env,
// field-major (field f of env e at rec[f * n_envs + e]).
#pragma once

#include "bgx_step.h"

namespace bgx {

// record of one env: the forward's tape (TapeLayout), then the cotangents
// stage A2 writes for stage B
template <class T>
struct AdjRec : TapeLayout<T> {
  using TL = TapeLayout<T>;
  static constexpr int QFB = TL::SIZE;               // [NV]  cotangent of qf_smooth
  static constexpr int MB = QFB + TL::NV;            // [NMB] cotangent of M (symmetrized)
  static constexpr int JB = MB + TL::NMB;            // [NRC][NV] cotangent of J contact rows
  static constexpr int DB = JB + TL::NRC * TL::NV;   // [NR]  cotangent of d
  static constexpr int AB = DB + TL::NR;             // [NR]  cotangent of aref
  static constexpr int GQ1 = AB + TL::NR;            // [NQ]  cotangent of q'
  static constexpr int GQDT = GQ1 + TL::NQ;          // [NV]  qd' cotangent incl. the q' update
  static constexpr int SIZE = GQDT + TL::NV;
};

BGX_D V3 ld3(const RecPtr& rec, int f) { return v3(rec[f], rec[f + 1], rec[f + 2]); }

// ---------------------------------------------------------------------------
// stage A2: reverse sweep from the tape (stage A1 is substep_impl recording
// its tape with a slot per candidate).  Slot loops are rolled (dynamic slot
// index, tape in global memory): compact code for any candidate count.

template <class T>
BGX_D void adjoint_rev(const float* __restrict__ M, const float* q, const float* qd,
                       const float* act, const float* gq1, const float* gqd1, RecPtr rec,
                       float* gact) {
  using R = AdjRec<T>;
  constexpr int NV = T::NV, L = T::L, NR = R::NR, NRC = R::NRC;
  const int n_c = (int)rec[R::NC], iters = (int)rec[R::ITERS];

  Tree<T> t;
  float lm_store[T::LM_LDS ? R::NMB : 1];
  if constexpr (T::LM_LDS) t.lml = lm_store, t.lms = 1;
#pragma unroll
  for (int d = 0; d < NV; ++d) {
    t.ca[d] = ld3(rec, R::CA + 3 * d);
    t.cv[d] = ld3(rec, R::CV + 3 * d);
    t.inv[d] = rec[R::INV + d];
    t.lm(d, d) = rec[R::LM + mpat_idx<T>(d, d)];
#pragma unroll
    for (int j = 0; j < d; ++j)
      if ((T::DANC[d] >> j) & 1u) t.lm(d, j) = rec[R::LM + mpat_idx<T>(d, j)];
  }
  float lside[R::NLR], ldiag[R::NLR];
  bool any_lim = false;
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i) {
    lside[i] = rec[R::LSIDE + i];
    ldiag[i] = rec[R::LDIAG + i];
    any_lim = any_lim || lside[i] != 0.f;
  }
  const bool has_rows = n_c > 0 || any_lim;

  // row vectors: contact rows 4 s + k (s < n_c), then limit rows NRC + i
  auto J_apply = [&](const float* z, float* out) {
    V3 va[L], vv[L];
    link_vel<T>(t.ca, t.cv, z, va, vv);
#pragma unroll 1
    for (int s = 0; s < n_c; ++s) {
      const int l = (int)rec[R::SLINK + s];
      V3 a = va[0], v = vv[0];
#pragma unroll
      for (int k = 1; k < L; ++k)
        if (l == k) a = va[k], v = vv[k];
      const V3 vp = v + cross(a, ld3(rec, R::SR + 3 * s));
      const float e0 = dot(ld3(rec, R::SD0 + 3 * s), vp);
      const float e1 = dot(ld3(rec, R::SD1 + 3 * s), vp);
      const float e2 = dot(ld3(rec, R::SD2 + 3 * s), vp);
      out[4 * s] = e0 + e1;
      out[4 * s + 1] = e0 - e1;
      out[4 * s + 2] = e0 + e2;
      out[4 * s + 3] = e0 - e2;
    }
#pragma unroll 1
    for (int r = 4 * n_c; r < NRC; ++r) out[r] = 0.f;
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) out[NRC + i] = lside[i] * z[T::LIM_D[i]];
  };
  auto JT_apply = [&](const float* x, float* tau) {
    V3 fa[L], fv[L];
#pragma unroll
    for (int l = 0; l < L; ++l) fa[l] = v3zero(), fv[l] = v3zero();
#pragma unroll 1
    for (int s = 0; s < n_c; ++s) {
      const float x0 = x[4 * s], x1 = x[4 * s + 1], x2 = x[4 * s + 2], x3 = x[4 * s + 3];
      const V3 w = ld3(rec, R::SD0 + 3 * s) * ((x0 + x1) + (x2 + x3)) +
                   ld3(rec, R::SD1 + 3 * s) * (x0 - x1) + ld3(rec, R::SD2 + 3 * s) * (x2 - x3);
      const V3 ta = cross(ld3(rec, R::SR + 3 * s), w);
      const int l = (int)rec[R::SLINK + s];
#pragma unroll
      for (int k = 0; k < L; ++k)
        if (hosts_contact<T>(k) && l == k) fa[k] = fa[k] + ta, fv[k] = fv[k] + w;
    }
    project_forces<T>(t, fa, fv, tau);
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) tau[T::LIM_D[i]] += lside[i] * x[NRC + i];
  };
  auto rdiag = [&](int r) {
    return r < NRC ? (r < 4 * n_c ? rec[R::SDIAG + r / 4] : 0.f) : ldiag[r - NRC];
  };
  // out = A x, ux = M^-1 J^T x
  auto A_apply = [&](const float* x, float* out, float* ux) {
    JT_apply(x, ux);
    ltl_solve<T>(t, ux);
    J_apply(ux, out);
#pragma unroll 1
    for (int r = 0; r < NR; ++r) out[r] = out[r] + rdiag(r) * x[r];
  };
  auto row = [&](int base, int k, float* out) {  // a tape row vector of iteration k
#pragma unroll 1
    for (int r = 0; r < NR; ++r) out[r] = rec[base + k * NR + r];
  };

  // ---- q' = f(q, qd'): cotangent of qd' from the position update
  const float dt = M[H_DT];
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
        const V3T<Dual> ang{dual(rec[R::QD1 + da + 3], k == 0 ? 1.f : 0.f),
                            dual(rec[R::QD1 + da + 4], k == 1 ? 1.f : 0.f),
                            dual(rec[R::QD1 + da + 5], k == 2 ? 1.f : 0.f)};
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
  // Jbar += a c^T on the active contact rows (in the record)
#pragma unroll 1
  for (int r = 0; r < 4 * n_c; ++r)
#pragma unroll
    for (int d = 0; d < NV; ++d) rec[R::JB + r * NV + d] = 0.f;
  auto jbar_add = [&](const float* a, const float* c) {
#pragma unroll 1
    for (int r = 0; r < 4 * n_c; ++r) {
      const float ar = a[r];
#pragma unroll
      for (int d = 0; d < NV; ++d) rec[R::JB + r * NV + d] += ar * c[d];
    }
  };
  float Db[NR], xb[NR], yb[NR], bb[NR];
#pragma unroll 1
  for (int r = 0; r < NR; ++r) Db[r] = xb[r] = yb[r] = bb[r] = 0.f;

  // ---- qdd = u + w, w = M^-1 J^T x
  float ub[NV], z[NV], u[NV], w[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) {
    ub[d] = z[d] = dt * gqdt[d];
    u[d] = rec[R::U + d];
    w[d] = rec[R::W + d];
  }
  if (has_rows && iters > 0) {
    float x[NR];
    row(R::XS, iters - 1, x);
    ltl_solve<T>(t, z);
    mbar_sub(z, w);
    J_apply(z, xb);
    jbar_add(x, z);
  }
  // ---- unrolled FISTA, last iteration first:
  //   r_k = A y_k + b, g_k = A r_k, x_{k+1} = relu(y_k - eta_k g_k),
  //   y_{k+1} = x_{k+1} + beta_k (x_{k+1} - x_k)
#pragma unroll 1
  for (int k = iters - 1; k >= 0; --k) {
    const float bk = rec[R::BETA + k], ek = rec[R::ETA + k];
    float wb[NR], gb[NR], rb[NR], ay[NR], xk1[NR], rk[NR];
    row(R::XS, k, xk1);
    row(R::RS, k, rk);
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
      const float xt = xb[r] + (1.f + bk) * yb[r];
      xb[r] = -bk * yb[r];
      wb[r] = xk1[r] > 0.f ? xt : 0.f;
      gb[r] = -ek * wb[r];
    }
    float ug[NV], ur[NV], urb[NV];
    A_apply(gb, rb, ug);  // r-bar = A g-bar
    JT_apply(rk, ur);
    ltl_solve<T>(t, ur);
    jbar_add(gb, ur);
    jbar_add(rk, ug);
    mbar_sub(ug, ur);
#pragma unroll 1
    for (int r = 0; r < NR; ++r) Db[r] += gb[r] * rk[r];
    A_apply(rb, ay, urb);  // y-bar = w-bar + A r-bar
    if (k > 0) {  // y_0 = 0
      float yk[NR], uy[NV];
      row(R::YS, k, yk);
      JT_apply(yk, uy);
      ltl_solve<T>(t, uy);
      jbar_add(rb, uy);
      jbar_add(yk, urb);
      mbar_sub(urb, uy);
#pragma unroll 1
      for (int r = 0; r < NR; ++r) Db[r] += rb[r] * yk[r];
    }
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
      yb[r] = wb[r] + ay[r];
      bb[r] += rb[r];
    }
  }
  // ---- b = J u - aref
  {
    float jtb[NV];
    JT_apply(bb, jtb);
#pragma unroll
    for (int d = 0; d < NV; ++d) ub[d] += jtb[d];
    jbar_add(bb, u);
  }
  // ---- u = M^-1 qf
  float qfb[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) qfb[d] = ub[d];
  ltl_solve<T>(t, qfb);
  mbar_sub(qfb, u);
  // ---- actuators: qf[da] += clip(gain clip(act) + bias, force) gear
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
#pragma unroll 1
  for (int r = 0; r < NR; ++r) rec[R::DB + r] = Db[r], rec[R::AB + r] = -bb[r];
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) rec[R::GQ1 + k] = gq1[k];
#pragma unroll
  for (int d = 0; d < NV; ++d) rec[R::GQDT + d] = gqdt[d];
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
