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
//  A1 (one thread per env and substep of a segment): the step's own substep
//    (substep_impl, a slot per candidate) recording its tape: tree, contact
//    slots, accelerations, every projected-gradient iterate and the
//    products M^-1 J^T y_k, M^-1 J^T r_k its matrix applications formed.
//  A2 (two lanes per env): the reverse sweep of the integrator, the solves
//    with M = L^T L and the unrolled FISTA iterations, by hand, from the
//    tape.  Output: cotangents of the position- and velocity-dependent
//    ingredients -- qf_smooth, the mass matrix (a symmetric matrix on the
//    dof-ancestor pattern), the contact rows of J, the diagonal regularizer
//    d and aref -- plus the action gradient and the velocity pass-through of
//    qd' = qd + dt qdd.  The cotangent of J is never formed: J enters only
//    through products a^T J c, and each such product is carried back to the
//    quantities J is built from (cdof, the slots' lever arms and pyramid
//    directions), a few floats per dof and per slot.  The row vectors of the
//    sweep live in LDS; the lane pair splits each iteration's products.
//  B (one thread per env and q coordinate): run the same formulas that
//    build those ingredients from (q, qd) in dual numbers along one input
//    coordinate and contract the tangents with the cotangents of stage A:
//    d/de <cotangents, ingredients(q + e dq, qd)>.  The qd coordinates take
//    one more thread per env: only qf_smooth and aref move with qd, and one
//    reverse pass through them (adjoint_b_qd_all) gives the gradient in all
//    of qd at once.
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
  // per slot: geometry in the reverse sweep's LDS (link, r, d0, d1, d2, diag;
  // two-link slots also linka, ra) and cotangents (d0, d1, d2, r; also ra)
  static constexpr int SLW = T::PAIRS ? 18 : 14;
  static constexpr int SBW = T::PAIRS ? 15 : 12;
  static constexpr int SLB = MB + TL::NMB;           // [KS][SBW] cotangents of each slot's
                                                     //   d0, d1, d2, lever arm r (and ra)
  static constexpr int CAB = SLB + SBW * TL::KS;     // [3 NV] cotangent of cdof (angular)
  static constexpr int CVB = CAB + 3 * TL::NV;       // [3 NV] cotangent of cdof (linear)
  static constexpr int DB = CVB + 3 * TL::NV;        // [NR]  cotangent of d
  static constexpr int AB = DB + TL::NR;             // [NR]  cotangent of aref
  static constexpr int GQ1 = AB + TL::NR;            // [NQ]  cotangent of q'
  static constexpr int GQDT = GQ1 + TL::NQ;          // [NV]  qd' cotangent incl. the q' update
  static constexpr int SIZE = GQDT + TL::NV;
};

// LDS of the reverse sweep.  Shared by the lanes of an env (element e of an
// env at base[e * envs per workgroup]): the row vectors over the NR
// constraint rows and each slot's geometry (link, r, d0, d1, d2, diag).
// Private to a lane (element e at base[e * lanes per workgroup]): each
// slot's cotangents (d0, d1, d2, r), the cdof cotangents, and the link
// velocities and forces of the products in flight, indexed by link.
template <class T>
struct RevLds {
  using R = AdjRec<T>;
  static constexpr int NR = R::NR, KS = R::KS, NV = T::NV, L = T::L;
  static constexpr int XB = 0, YB = XB + NR, BB = YB + NR, DB = BB + NR, WB = DB + NR,
                       GB = WB + NR, RB = GB + NR, TR = RB + NR, TR2 = TR + NR;
  static constexpr int SL = TR2 + NR;          // [KS][SLW]
  static constexpr int SHARED = SL + R::SLW * KS;
  static constexpr int SB = 0;                 // [KS][SBW]
  static constexpr int CAB = SB + R::SBW * KS;  // [3 NV]
  static constexpr int CVB = CAB + 3 * NV;  // [3 NV]
  static constexpr int LV = CVB + 3 * NV;   // [L][6] link velocity (ang, vel)
  static constexpr int LF = LV + 6 * L;     // [L][6] link force (torque, force)
  static constexpr int PRIVATE = LF + 6 * L;
};

// lanes per env of the reverse sweep: two lanes run the same instructions
// on different products of each FISTA iteration, then add their sums
template <class T>
constexpr int rev_pair() {
  return 2;
}

// LDS floats of a reverse-sweep workgroup of a given lane count
template <class T>
constexpr int rev_lds_floats(int lanes) {
  return RevLds<T>::SHARED * (lanes / rev_pair<T>()) + RevLds<T>::PRIVATE * lanes;
}

// lanes per workgroup of the reverse sweep: the most (a power of two up to
// WAVE) whose LDS fits in 80 KB, so two workgroups share a CU
template <class T>
constexpr int rev_lanes() {
  for (int b = 64; b > rev_pair<T>(); b /= 2)
    if (rev_lds_floats<T>(b) * 4 <= 81920) return b;
  return rev_pair<T>();
}

// a vector of LDS floats with a lane stride
template <int S>
struct LVec {
  float* p;
  BGX_D float& operator[](int r) const { return p[r * S]; }
};

// the value of the other lane of an env's lane pair
#ifdef BGX_HOST_CHECK
BGX_D float pair_other(float v) { return v; }  // the host build runs one lane per env
#else
BGX_D float pair_other(float v) { return __shfl_xor(v, 1); }
#endif

BGX_D V3 ld3(const RecPtr& rec, int f) { return v3(rec[f], rec[f + 1], rec[f + 2]); }

// ---------------------------------------------------------------------------
// stage A2: reverse sweep from the tape (stage A1 is substep_impl recording
// its tape with a slot per candidate).  Slot and row loops are rolled over
// the active rows only (contact rows 4 s + k of the n_c slots, then the
// static limit rows); the row vectors and slot data sit in LDS.

// P = 1: one lane per env.  P = 2: a pair of lanes per env (sub = 0, 1)
// computes the same sweep; the products of J's and M's cotangent that a
// FISTA iteration adds are split between the pair (sub 0: the ones that
// pair g-bar with M^-1 J^T r_k and r-bar with M^-1 J^T y_k; sub 1: r_k with
// M^-1 J^T g-bar and y_k with M^-1 J^T r-bar), each lane sums its own,
// and the pair adds its sums at the end.  Every lane of the pair returns
// the same result.
template <class T, int AL, int P = 1>
BGX_D void adjoint_rev(const float* __restrict__ M, const float* q, const float* qd,
                       const float* act, const float* gq1, const float* gqd1, RecPtr rec,
                       float* shared, float* priv, float* gact, int sub = 0) {
  using R = AdjRec<T>;
  using Y = RevLds<T>;
  constexpr int NV = T::NV, L = T::L, NR = R::NR, NRC = R::NRC;
  static_assert(P == 1 || P == 2, "one or two lanes per env");
  const int n_c = tape_int(rec[R::NC], 0, R::KS);
  const int iters = tape_int(rec[R::ITERS], 0, ADJ_MAXITER);
  // products both lanes form alike are counted by lane 0 only
  const float own = (P == 1 || sub == 0) ? 1.f : 0.f;
  constexpr int ES = AL / P;  // envs per workgroup: stride of the shared region
  using SV = LVec<ES>;
  using PV = LVec<AL>;
  auto vec = [&](int off) { return SV{shared + off * ES}; };
  auto pvec = [&](int off) { return PV{priv + off * AL}; };
  const SV xb = vec(Y::XB), yb = vec(Y::YB), bb = vec(Y::BB), Db = vec(Y::DB),
           wb = vec(Y::WB), gb = vec(Y::GB), rb = vec(Y::RB), tr = vec(Y::TR),
           tr2 = vec(Y::TR2), sl = vec(Y::SL);
  const PV sb = pvec(Y::SB), cab = pvec(Y::CAB), cvb = pvec(Y::CVB), lv = pvec(Y::LV),
           lf = pvec(Y::LF);
  const SV ay = gb;  // A r-bar; g-bar is dead by then

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
  constexpr int SW = R::SLW, BW = R::SBW;
  // slot geometry into LDS, slot cotangents cleared
#pragma unroll 1
  for (int s = 0; s < n_c; ++s) {
    float g[SW];
    g[0] = rec[R::SLINK + s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      g[1 + k] = rec[R::SR + 3 * s + k];
      g[4 + k] = rec[R::SD0 + 3 * s + k];
      g[7 + k] = rec[R::SD1 + 3 * s + k];
      g[10 + k] = rec[R::SD2 + 3 * s + k];
    }
    g[13] = rec[R::SDIAG + s];
    if constexpr (T::PAIRS) {
      g[14] = rec[R::SLINKA + s];
#pragma unroll
      for (int k = 0; k < 3; ++k) g[15 + k] = rec[R::SRA + 3 * s + k];
    }
#pragma unroll
    for (int k = 0; k < SW; ++k) sl[SW * s + k] = g[k];
#pragma unroll
    for (int k = 0; k < BW; ++k) sb[BW * s + k] = 0.f;
  }
#pragma unroll
  for (int k = 0; k < 3 * NV; ++k) cab[k] = 0.f, cvb[k] = 0.f;
  auto slot_v3 = [&](int s, int o) {
    return v3(sl[SW * s + o], sl[SW * s + o + 1], sl[SW * s + o + 2]);
  };
  // f(r) for every active row r
  auto for_rows = [&](auto&& f) {
#pragma unroll 2
    for (int r = 0; r < 4 * n_c; ++r) f(r);
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) f(NRC + i);
  };
  // tape row vector of iteration k into dst (4 loads in flight per slot)
  auto load_row = [&](const SV& dst, int base, int k) {
    const int b0 = base + k * NR;
#pragma unroll 2
    for (int s = 0; s < n_c; ++s) {
      const float v0 = rec[b0 + 4 * s], v1 = rec[b0 + 4 * s + 1];
      const float v2 = rec[b0 + 4 * s + 2], v3_ = rec[b0 + 4 * s + 3];
      dst[4 * s] = v0, dst[4 * s + 1] = v1, dst[4 * s + 2] = v2, dst[4 * s + 3] = v3_;
    }
    float lv[R::NLR];
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) lv[i] = rec[b0 + NRC + i];
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) dst[NRC + i] = lv[i];
  };

  // link velocities under z into lv (links that host a contact)
  auto store_link_vel = [&](const float* z) {
    V3 va[L], vv[L];
    link_vel<T>(t.ca, t.cv, z, va, vv);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (hosts_contact<T>(l)) {
        lv[6 * l] = va[l].x, lv[6 * l + 1] = va[l].y, lv[6 * l + 2] = va[l].z;
        lv[6 * l + 3] = vv[l].x, lv[6 * l + 4] = vv[l].y, lv[6 * l + 5] = vv[l].z;
      }
    }
  };
  auto clear_link_force = [&]() {
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (hosts_contact<T>(l))
#pragma unroll
        for (int k = 0; k < 6; ++k) lf[6 * l + k] = 0.f;
  };
  auto add_link_force = [&](int l, V3 ta, V3 f) {
    float* p = &lf[6 * l];
    p[0] += ta.x, p[AL] += ta.y, p[2 * AL] += ta.z;
    p[3 * AL] += f.x, p[4 * AL] += f.y, p[5 * AL] += f.z;
  };
  auto load_link_force = [&](V3* fa, V3* fv) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (hosts_contact<T>(l)) {
        fa[l] = v3(lf[6 * l], lf[6 * l + 1], lf[6 * l + 2]);
        fv[l] = v3(lf[6 * l + 3], lf[6 * l + 4], lf[6 * l + 5]);
      } else {
        fa[l] = v3zero(), fv[l] = v3zero();
      }
    }
  };
  auto slot_link = [&](int s) { return tape_int(sl[SW * s], 0, L - 1); };
  // second link of a two-link slot (-1: none)
  auto slot_linka = [&](int s) {
    if constexpr (T::PAIRS) return tape_int(sl[SW * s + 14], -1, L - 1);
    else return -1;
  };
  auto link_vel_of = [&](int l, V3& w, V3& v) {
    const float* p = &lv[6 * l];
    w = v3(p[0], p[AL], p[2 * AL]);
    v = v3(p[3 * AL], p[4 * AL], p[5 * AL]);
  };
  // velocity of slot s's contact point under the link velocities in lv
  // (on its link, minus on its second link); w, wa: the two links' angular
  // velocities (wa = 0 without a second link)
  auto slot_pvel = [&](int s, V3& w, V3& wa) {
    V3 v;
    link_vel_of(slot_link(s), w, v);
    V3 vp = v + cross(w, slot_v3(s, 1));
    wa = v3zero();
    const int la = slot_linka(s);
    if (la >= 0) {
      V3 va;
      link_vel_of(la, wa, va);
      vp = vp - (va + cross(wa, slot_v3(s, 15)));
    }
    return vp;
  };

  // J z on the active rows (a lane pair splits the slots: sub 0 the even
  // ones, sub 1 the odd ones, into the pair's shared row vector)
  auto J_apply = [&](const float* z, const SV& out) {
    store_link_vel(z);
#pragma unroll 2
    for (int s = P == 2 ? sub : 0; s < n_c; s += P) {
      V3 w, wa;
      const V3 vp = slot_pvel(s, w, wa);
      const float e0 = dot(slot_v3(s, 4), vp);
      const float e1 = dot(slot_v3(s, 7), vp);
      const float e2 = dot(slot_v3(s, 10), vp);
      out[4 * s] = e0 + e1;
      out[4 * s + 1] = e0 - e1;
      out[4 * s + 2] = e0 + e2;
      out[4 * s + 3] = e0 - e2;
    }
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) out[NRC + i] = lside[i] * z[T::LIM_D[i]];
  };
  // slot s's pyramid rows of x as one force at the contact point
  auto slot_force = [&](const SV& x, int s, float& al0, float& al1, float& al2) {
    const float x0 = x[4 * s], x1 = x[4 * s + 1], x2 = x[4 * s + 2], x3 = x[4 * s + 3];
    al0 = (x0 + x1) + (x2 + x3), al1 = x0 - x1, al2 = x2 - x3;
    return slot_v3(s, 4) * al0 + slot_v3(s, 7) * al1 + slot_v3(s, 10) * al2;
  };
  // J^T x: a lane pair splits the slots as J_apply does, each lane projects
  // the link forces of its own slots, and the pair adds the two (the same
  // sum on both lanes)
  auto JT_apply = [&](const SV& x, float* tau) {
    clear_link_force();
#pragma unroll 2
    for (int s = P == 2 ? sub : 0; s < n_c; s += P) {
      float a0, a1, a2;
      const V3 w = slot_force(x, s, a0, a1, a2);
      add_link_force(slot_link(s), cross(slot_v3(s, 1), w), w);
      const int la = slot_linka(s);
      if (la >= 0) add_link_force(la, cross(w, slot_v3(s, 15)), v3zero() - w);
    }
    V3 fa[L], fv[L];
    load_link_force(fa, fv);
    project_forces<T>(t, fa, fv, tau);
    if constexpr (P == 2) {
#pragma unroll
      for (int d = 0; d < NV; ++d) tau[d] += pair_other(tau[d]);
    }
#pragma unroll
    for (int i = 0; i < T::NLIM; ++i) tau[T::LIM_D[i]] += lside[i] * x[NRC + i];
  };
  auto rdiag = [&](int r) { return r < NRC ? sl[SW * (r >> 2) + 13] : ldiag[r - NRC]; };
  // out = A x, ux = M^-1 J^T x
  auto A_apply = [&](const SV& x, const SV& out, float* ux) {
    JT_apply(x, ux);
    ltl_solve<T>(t, ux);
    J_apply(ux, out);
    for_rows([&](int r) { out[r] = out[r] + rdiag(r) * x[r]; });
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
  // Jbar += a c^T on the contact rows, carried to J's ingredients: contact
  // row block of slot s is C E with E[i][d] = d_i . (cv_d + ca_d x r) over
  // the dofs d above the slot's link and C = [1 1 0; 1 -1 0; 1 0 1; 1 0 -1].
  // With alpha = C^T a_s, f = sum_i alpha_i d_i and (w, v) the velocity of
  // the slot's link under c:  d_i-bar += alpha_i (v + w x r),
  // r-bar += f x w,  ca_d-bar += c_d (subtree sum of r x f),
  // cv_d-bar += c_d (subtree sum of f).
  auto jbar_add = [&](const SV& a, const float* c) {
    store_link_vel(c);
    clear_link_force();
#pragma unroll 2
    for (int s = 0; s < n_c; ++s) {
      float a0, a1, a2;
      const V3 f = slot_force(a, s, a0, a1, a2);
      const V3 r = slot_v3(s, 1);
      V3 w, wa;
      const V3 vp = slot_pvel(s, w, wa);
      const V3 rf = cross(f, w);
      float* g = &sb[BW * s];
      g[0] += a0 * vp.x, g[AL] += a0 * vp.y, g[2 * AL] += a0 * vp.z;
      g[3 * AL] += a1 * vp.x, g[4 * AL] += a1 * vp.y, g[5 * AL] += a1 * vp.z;
      g[6 * AL] += a2 * vp.x, g[7 * AL] += a2 * vp.y, g[8 * AL] += a2 * vp.z;
      g[9 * AL] += rf.x, g[10 * AL] += rf.y, g[11 * AL] += rf.z;
      add_link_force(slot_link(s), cross(r, f), f);
      const int la = slot_linka(s);
      if (la >= 0) {
        // the second link: velocity -(va + wa x ra), force -f at lever ra
        const V3 ra = slot_v3(s, 15);
        const V3 rfa = cross(wa, f);
        g[12 * AL] += rfa.x, g[13 * AL] += rfa.y, g[14 * AL] += rfa.z;
        add_link_force(la, cross(f, ra), v3zero() - f);
      }
    }
    V3 fa[L], fv[L];
    load_link_force(fa, fv);
    V3 acc_a[L], acc_v[L];
#pragma unroll
    for (int l = 0; l < L; ++l) acc_a[l] = v3zero(), acc_v[l] = v3zero();
#pragma unroll
    for (int l = L - 1; l >= 0; --l) {
      const V3 sa = fa[l] + acc_a[l], sv = fv[l] + acc_v[l];
      const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (k < nd) {
          const int d = T::DADR[l] + k;
          cab[3 * d] += c[d] * sa.x, cab[3 * d + 1] += c[d] * sa.y, cab[3 * d + 2] += c[d] * sa.z;
          cvb[3 * d] += c[d] * sv.x, cvb[3 * d + 1] += c[d] * sv.y, cvb[3 * d + 2] += c[d] * sv.z;
        }
      }
      const int par = T::PARENT[l];
      if (par >= 0) acc_a[par] = acc_a[par] + sa, acc_v[par] = acc_v[par] + sv;
    }
  };
  for_rows([&](int r) { Db[r] = xb[r] = yb[r] = bb[r] = 0.f; });

  // ---- qdd = u + w, w = M^-1 J^T x
  float ub[NV], z[NV], u[NV], w[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) {
    ub[d] = z[d] = dt * gqdt[d];
    u[d] = rec[R::U + d];
    w[d] = rec[R::W + d];
  }
  if (has_rows && iters > 0) {
    float zo[NV];
    load_row(tr, R::XS, iters - 1);
    ltl_solve<T>(t, z);
#pragma unroll
    for (int d = 0; d < NV; ++d) zo[d] = own * z[d];
    mbar_sub(zo, w);
    J_apply(z, xb);
    jbar_add(tr, zo);
  }
  // ---- unrolled FISTA, last iteration first:
  //   r_k = A y_k + b, g_k = A r_k, x_{k+1} = relu(y_k - eta_k g_k),
  //   y_{k+1} = x_{k+1} + beta_k (x_{k+1} - x_k)
#pragma unroll 1
  for (int k = iters - 1; k >= 0; --k) {
    const float bk = rec[R::BETA + k], ek = rec[R::ETA + k];
    load_row(tr, R::XS, k);  // x_{k+1}
    for_rows([&](int r) {
      const float xt = xb[r] + (1.f + bk) * yb[r];
      xb[r] = -bk * yb[r];
      const float wr = tr[r] > 0.f ? xt : 0.f;
      wb[r] = wr;
      gb[r] = -ek * wr;
    });
    load_row(tr, R::RS, k);  // r_k
    if constexpr (P == 1) {
      float ug[NV], ur[NV], urb[NV];
      A_apply(gb, rb, ug);  // r-bar = A g-bar
#pragma unroll
      for (int d = 0; d < NV; ++d) ur[d] = rec[R::UR + k * NV + d];  // M^-1 J^T r_k
      jbar_add(gb, ur);
      jbar_add(tr, ug);
      mbar_sub(ug, ur);
      for_rows([&](int r) { Db[r] += gb[r] * tr[r]; });
      A_apply(rb, ay, urb);  // y-bar = w-bar + A r-bar
      if (k > 0) {  // y_0 = 0
        float uy[NV];
        load_row(tr, R::YS, k);  // y_k
#pragma unroll
        for (int d = 0; d < NV; ++d) uy[d] = rec[R::UY + k * NV + d];  // M^-1 J^T y_k
        jbar_add(rb, uy);
        jbar_add(tr, urb);
        mbar_sub(urb, uy);
        for_rows([&](int r) { Db[r] += rb[r] * tr[r]; });
      }
    } else {
      if (k > 0) load_row(tr2, R::YS, k);  // y_k
      else for_rows([&](int r) { tr2[r] = 0.f; });  // y_0 = 0
      float ug[NV], urb[NV], uo[NV], ux[NV], c[NV];
      A_apply(gb, rb, ug);  // r-bar = A g-bar
      // sub 0: ur = M^-1 J^T r_k;  sub 1: uy = M^-1 J^T y_k (both on the tape)
#pragma unroll
      for (int d = 0; d < NV; ++d) uo[d] = rec[(sub ? R::UY : R::UR) + k * NV + d];
#pragma unroll
      for (int d = 0; d < NV; ++d) c[d] = sub ? ug[d] : uo[d];
      jbar_add(sub ? tr : gb, c);  // (g-bar, ur) | (r_k, ug)
      for_rows([&](int r) { Db[r] += gb[r] * tr[r]; });
      A_apply(rb, ay, urb);  // y-bar = w-bar + A r-bar
#pragma unroll
      for (int d = 0; d < NV; ++d) ux[d] = rec[(sub ? R::UR : R::UY) + k * NV + d];  // sub 0: uy
#pragma unroll
      for (int d = 0; d < NV; ++d) c[d] = sub ? urb[d] : ux[d];
      jbar_add(sub ? tr2 : rb, c);  // (r-bar, uy) | (y_k, urb)
#pragma unroll
      for (int d = 0; d < NV; ++d) c[d] = sub ? urb[d] : ug[d];
      mbar_sub(c, uo);  // (ug, ur) | (urb, uy)
      for_rows([&](int r) { Db[r] += rb[r] * tr2[r]; });
    }
    for_rows([&](int r) {
      yb[r] = wb[r] + ay[r];
      bb[r] += rb[r];
    });
  }
  // ---- b = J u - aref
  if (has_rows) {
    float jtb[NV], uo[NV];
    JT_apply(bb, jtb);
#pragma unroll
    for (int d = 0; d < NV; ++d) ub[d] += jtb[d], uo[d] = own * u[d];
    jbar_add(bb, uo);
  }
  // ---- u = M^-1 qf
  float qfb[NV], qfo[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) qfb[d] = ub[d];
  ltl_solve<T>(t, qfb);
#pragma unroll
  for (int d = 0; d < NV; ++d) qfo[d] = own * qfb[d];
  mbar_sub(qfo, u);
  if constexpr (P == 2) {  // the pair's sums
#pragma unroll
    for (int k = 0; k < R::NMB; ++k) Mb[k] += pair_other(Mb[k]);
#pragma unroll
    for (int k = 0; k < 3 * NV; ++k) {
      cab[k] += pair_other(cab[k]);
      cvb[k] += pair_other(cvb[k]);
    }
#pragma unroll 1
    for (int s = 0; s < n_c; ++s)
#pragma unroll
      for (int k = 0; k < BW; ++k) sb[BW * s + k] += pair_other(sb[BW * s + k]);
  }
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

  if (P == 2 && sub != 0) return;  // lane 0 of the pair writes
#pragma unroll
  for (int d = 0; d < NV; ++d) rec[R::QFB + d] = qfb[d];
#pragma unroll
  for (int k = 0; k < R::NMB; ++k) rec[R::MB + k] = Mb[k];
  for_rows([&](int r) {
    const float d_ = Db[r], a_ = -bb[r];
    rec[R::DB + r] = d_, rec[R::AB + r] = a_;
  });
#pragma unroll 1
  for (int s = 0; s < n_c; ++s) {
    float g[BW];
#pragma unroll
    for (int k = 0; k < BW; ++k) g[k] = sb[BW * s + k];
#pragma unroll
    for (int k = 0; k < BW; ++k) rec[R::SLB + BW * s + k] = g[k];
  }
#pragma unroll
  for (int k = 0; k < 3 * NV; ++k) rec[R::CAB + k] = cab[k], rec[R::CVB + k] = cvb[k];
#pragma unroll
  for (int k = 0; k < T::NQ; ++k) rec[R::GQ1 + k] = gq1[k];
#pragma unroll
  for (int d = 0; d < NV; ++d) rec[R::GQDT + d] = gqdt[d];
}

// ---------------------------------------------------------------------------
// stage B: d/de of <cotangents, ingredients> along input coordinate dir
// (dir < NQ: q[dir]; else qd[dir - NQ]); returns the input cotangent

// along q[dir]: every ingredient moves
template <class T>
BGX_D float adjoint_b_q(const float* __restrict__ M, const float* q_in, const float* qd_in,
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
  V3T<S> rc[L];
  root_coms<T>(M, xp, xr, rc);
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
  // crba emits each entry of the dof-ancestor pattern once (diagonal, then
  // the row's ancestors in ascending order): contract it as it comes instead
  // of holding the dual matrix
  crba<T>(M, cc, ci, ca, cv,
          [&](int i, int j, S v) { phi += rec[R::MB + mpat_idx<T>(i, j)] * v.d; });

  // contact rows of the active candidates
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    const int s = tape_int(rec[R::CSLOT + c], -1, R::KS - 1);
    if (s >= 0) {
      const int l = T::CAND_LINK[c], la = T::CAND_LINKA[c];
      const float* cd = cand_rec<T>(M, c);
      const int fb = tape_int(rec[R::CFB + c], 0, 1);
      V3T<S> ctr, axis, pos, n, t1, t2;
      S dist;
      if (T::PAIRS && T::CAND_KIND[c] >= KIND_SS) {
        pair_contact<T>(M, c, xp, xr, fb, dist, pos, n, t1, t2);
        if (T::CAND_FLIP[c]) n = n * -1.f, t1 = t1 * -1.f, t2 = t2 * -1.f;
      } else {
        candidate_center<T>(M, c, xp, xr, ctr, axis);
        dist = candidate_dist<T>(M, c, ctr);
        contact_frame<T>(M, c, ctr, axis, dist, fb, pos, t1, t2);
        n = lift3<S>(PZ ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN));
      }
      const float mu = cd[CD_MU];
      const V3T<S> r = pos - rc[l];
      const V3T<S> ra = la >= 0 ? pos - rc[la >= 0 ? la : 0] : r;
      const float* im = cd + CD_IMP;
      const S imp = impedance(im, dist);
      const float tw = cd[CD_TW];
      const S diag = (tw + mu * mu * tw) * (2.f * mu * mu * imp_ratio(imp));
      const S kt = im[IM_K] * imp * dist;
      const float bq = im[IM_B];
      V3T<S> vq = cdv[l] + cross(cda[l], r);
      if (la >= 0) {
        const int lc = la >= 0 ? la : 0;
        vq = vq - (cdv[lc] + cross(cda[lc], ra));
      }
      const S e0 = dot(n, vq), e1 = mu * dot(t1, vq), e2 = mu * dot(t2, vq);
      const int r0 = 4 * s;
      phi += rec[R::AB + r0] * -(bq * (e0.d + e1.d) + kt.d);
      phi += rec[R::AB + r0 + 1] * -(bq * (e0.d - e1.d) + kt.d);
      phi += rec[R::AB + r0 + 2] * -(bq * (e0.d + e2.d) + kt.d);
      phi += rec[R::AB + r0 + 3] * -(bq * (e0.d - e2.d) + kt.d);
      phi += ((rec[R::DB + r0] + rec[R::DB + r0 + 1]) +
              (rec[R::DB + r0 + 2] + rec[R::DB + r0 + 3])) * diag.d;
      // J's contact rows through their ingredients (d0 = n, d1 = mu t1,
      // d2 = mu t2, r); the cdof part is contracted once below
      const int g = R::SLB + R::SBW * s;
      const V3T<S> d1 = t1 * mu, d2 = t2 * mu;
      if constexpr (!PZ) phi += rec[g] * n.x.d + rec[g + 1] * n.y.d + rec[g + 2] * n.z.d;
      phi += rec[g + 3] * d1.x.d + rec[g + 4] * d1.y.d + rec[g + 5] * d1.z.d;
      phi += rec[g + 6] * d2.x.d + rec[g + 7] * d2.y.d + rec[g + 8] * d2.z.d;
      phi += rec[g + 9] * r.x.d + rec[g + 10] * r.y.d + rec[g + 11] * r.z.d;
      if constexpr (T::PAIRS)
        if (la >= 0) phi += rec[g + 12] * ra.x.d + rec[g + 13] * ra.y.d + rec[g + 14] * ra.z.d;
    }
  }
#pragma unroll
  for (int d = 0; d < NV; ++d) {
    phi += rec[R::CAB + 3 * d] * ca[d].x.d + rec[R::CAB + 3 * d + 1] * ca[d].y.d +
           rec[R::CAB + 3 * d + 2] * ca[d].z.d;
    phi += rec[R::CVB + 3 * d] * cv[d].x.d + rec[R::CVB + 3 * d + 1] * cv[d].y.d +
           rec[R::CVB + 3 * d + 2] * cv[d].z.d;
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
  }
  return phi;
}

BGX_D Sym3T<Dual> lift_sym(const Sym3& a) {
  return Sym3T<Dual>{lift<Dual>(a.xx), lift<Dual>(a.yy), lift<Dual>(a.zz),
                     lift<Dual>(a.xy), lift<Dual>(a.xz), lift<Dual>(a.yz)};
}

// along qd[j]: the geometry (frames, cdof, M, J, the contact frames) stays
// put; only qf_smooth, aref's velocity terms and the pass-through move
template <class T>
BGX_D float adjoint_b_qd(const float* __restrict__ M, const float* q_in, const float* qd_in,
                         const float* act, RecPtr rec, int j) {
  using R = AdjRec<T>;
  using S = Dual;
  constexpr int NV = T::NV, NQ = T::NQ, L = T::L, NRC = R::NRC;
  constexpr bool PZ = T::PLANE_Z != 0;
  V3 xp[L];
  Q4 xr[L];
  forward_kinematics<T, false, float>(M, q_in, qd_in, xp, xr, nullptr, nullptr);
  V3 rc[L];
  root_coms<T>(M, xp, xr, rc);
  S q[NQ], qd[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = lift<S>(q_in[k]);
#pragma unroll
  for (int k = 0; k < NV; ++k) qd[k] = dual(qd_in[k], k == j ? 1.f : 0.f);
  V3T<S> ca[NV], cv[NV], cc[L], cda[L], cdv[L];
  Sym3T<S> ci[L];
  {
    V3 ccf[L], caf[NV], cvf[NV];
    Sym3 cif[L];
    com_inertia<T>(M, xp, xr, rc, ccf, cif);
    dof_motions<T>(M, q_in, xp, xr, rc, caf, cvf);
#pragma unroll
    for (int l = 0; l < L; ++l) cc[l] = lift3<S>(ccf[l]), ci[l] = lift_sym(cif[l]);
#pragma unroll
    for (int d = 0; d < NV; ++d) ca[d] = lift3<S>(caf[d]), cv[d] = lift3<S>(cvf[d]);
  }
  link_vel<T>(ca, cv, qd, cda, cdv);
  S qf[NV];
  rnea_bias<T>(M, ca, cv, cda, cdv, cc, ci, qd, qf);
  smooth_forces<T>(M, q, qd, act, qf);
  float phi = 0.f;
#pragma unroll
  for (int d = 0; d < NV; ++d) phi += rec[R::QFB + d] * qf[d].d;

  // contact rows: aref = -b (J qd) - k imp dist, its velocity term
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    const int s = tape_int(rec[R::CSLOT + c], -1, R::KS - 1);
    if (s >= 0) {
      const int l = T::CAND_LINK[c], la = T::CAND_LINKA[c];
      const float* cd = cand_rec<T>(M, c);
      const int fb = tape_int(rec[R::CFB + c], 0, 1);
      V3 ctr, axis, pos, n, t1, t2;
      float dist;
      if (T::PAIRS && T::CAND_KIND[c] >= KIND_SS) {
        pair_contact<T>(M, c, xp, xr, fb, dist, pos, n, t1, t2);
        if (T::CAND_FLIP[c]) n = n * -1.f, t1 = t1 * -1.f, t2 = t2 * -1.f;
      } else {
        candidate_center<T>(M, c, xp, xr, ctr, axis);
        dist = candidate_dist<T>(M, c, ctr);
        contact_frame<T>(M, c, ctr, axis, dist, fb, pos, t1, t2);
        n = PZ ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN);
      }
      const float mu = cd[CD_MU];
      const V3T<S> r = lift3<S>(pos - rc[l]);
      const float bq = cd[CD_IMP + IM_B];
      V3T<S> vq = cdv[l] + cross(cda[l], r);
      if (la >= 0) {
        const int lc = la >= 0 ? la : 0;
        vq = vq - (cdv[lc] + cross(cda[lc], lift3<S>(pos - rc[lc])));
      }
      const float e0 = dot(lift3<S>(n), vq).d, e1 = mu * dot(lift3<S>(t1), vq).d;
      const float e2 = mu * dot(lift3<S>(t2), vq).d;
      const int r0 = 4 * s;
      phi -= bq * (rec[R::AB + r0] * (e0 + e1) + rec[R::AB + r0 + 1] * (e0 - e1) +
                   rec[R::AB + r0 + 2] * (e0 + e2) + rec[R::AB + r0 + 3] * (e0 - e2));
    }
  }
  // limit rows: aref = -b side qd[d] - k imp pos
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i) {
    const int d = T::LIM_D[i], qi = T::LIM_Q[i];
    if (d == j) {
      const float* df = dof_rec<T>(M, d);
      const float pmin = q_in[qi] - df[DF_LO];
      const float pmax = df[DF_HI] - q_in[qi];
      const float pos = fminf(fminf(pmin, pmax), 0.f);
      if (pos < 0.f) {
        const float side = pmin < pmax ? 1.f : -1.f;
        phi += rec[R::AB + NRC + i] * (-df[DF_IMP + IM_B] * side);
      }
    }
  }
  return phi + rec[R::GQDT + j];
}

// Along every qd coordinate at once: the gradient in qd of <cotangents,
// ingredients(q, qd)> by one reverse pass, where adjoint_b_qd takes one
// coordinate at a time in dual numbers.  Only qf_smooth and aref move with
// qd: qf_smooth through the link velocities (link_vel) and RNEA's velocity
// terms (rnea_bias: the cdofd cross terms and cd x* (I cd)), the damping and
// the actuators' velocity bias (smooth_forces); aref through the contact
// points' and limited dofs' velocities.  Transposes used: inertia_mul is a
// symmetric map; cross_motion(s, .)^T = -cross_force(s, .); cross_force
// taken as linear in s has the transpose y -> (f_a x y_a + f_v x y_v,
// f_v x y_a).  Writes d phi / d qd_j (with the pass-through) to gqd[j].
template <class T>
BGX_D void adjoint_b_qd_all(const float* __restrict__ M, const float* q_in,
                            const float* qd_in, const float* act, RecPtr rec, float* gqd) {
  using R = AdjRec<T>;
  constexpr int NV = T::NV, L = T::L, NRC = R::NRC;
  constexpr bool PZ = T::PLANE_Z != 0;
  V3 xp[L];
  Q4 xr[L];
  forward_kinematics<T, false, float>(M, q_in, qd_in, xp, xr, nullptr, nullptr);
  V3 rc[L];
  root_coms<T>(M, xp, xr, rc);
  V3 cc[L], ca[NV], cv[NV];
  Sym3 ci[L];
  com_inertia<T>(M, xp, xr, rc, cc, ci);
  dof_motions<T>(M, q_in, xp, xr, rc, ca, cv);
  V3 cda[L], cdv[L];
  link_vel<T>(ca, cv, qd_in, cda, cdv);
  float qfb[NV], g[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) {
    qfb[d] = rec[R::QFB + d];
    g[d] = rec[R::GQDT + d] - dof_rec<T>(M, d)[DF_DAMPING] * qfb[d];
  }
  // actuators' velocity bias, where the force is not clipped
#pragma unroll
  for (int a = 0; a < T::NU; ++a) {
    const float* ac = M + T::OFF_ACT + a * ACT_STRIDE;
    const int qa = T::ACT_Q[a], da = T::ACT_D[a];
    const float ctl = clipf(act[a], ac[AC_CTRL_LO], ac[AC_CTRL_HI]);
    const float f = ac[AC_GAIN] * ctl +
                    ac[AC_GEAR] * (q_in[qa] * ac[AC_BIAS_Q] + qd_in[da] * ac[AC_BIAS_QD]);
    if (f >= ac[AC_FORCE_LO] && f <= ac[AC_FORCE_HI])
      g[da] += qfb[da] * ac[AC_GEAR] * ac[AC_GEAR] * ac[AC_BIAS_QD];
  }
  // cotangent of each link's RNEA force: qf_smooth = ... - bias, bias_d =
  // cdof_d . (subtree sum of forces), so the force of link l carries
  // -sum qfb_d cdof_d over the dofs of l and its ancestors (root to leaf)
  V3 fba[L], fbv[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int par = T::PARENT[l];
    V3 a = par >= 0 ? fba[par] : v3zero(), v = par >= 0 ? fbv[par] : v3zero();
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = T::DADR[l] + k;
        a = a - ca[d] * qfb[d];
        v = v - cv[d] * qfb[d];
      }
    }
    fba[l] = a, fbv[l] = v;
  }
  // cotangents of the link velocities (cba, cbv) and of RNEA's
  // accelerations (xba, xbv)
  V3 cba[L], cbv[L], xba[L], xbv[L];
#pragma unroll
  for (int l = 0; l < L; ++l) cba[l] = cbv[l] = xba[l] = xbv[l] = v3zero();
  // aref's velocity terms of the active contacts: aref = -b (d . v_contact)
  // - k imp dist over the pyramid rows
#pragma unroll
  for (int c = 0; c < T::NCAND; ++c) {
    const int s = tape_int(rec[R::CSLOT + c], -1, R::KS - 1);
    if (s >= 0) {
      const int l = T::CAND_LINK[c], la = T::CAND_LINKA[c];
      const float* cd = cand_rec<T>(M, c);
      const int fb = tape_int(rec[R::CFB + c], 0, 1);
      V3 ctr, axis, pos, n, t1, t2;
      float dist;
      if (T::PAIRS && T::CAND_KIND[c] >= KIND_SS) {
        pair_contact<T>(M, c, xp, xr, fb, dist, pos, n, t1, t2);
        if (T::CAND_FLIP[c]) n = n * -1.f, t1 = t1 * -1.f, t2 = t2 * -1.f;
      } else {
        candidate_center<T>(M, c, xp, xr, ctr, axis);
        dist = candidate_dist<T>(M, c, ctr);
        contact_frame<T>(M, c, ctr, axis, dist, fb, pos, t1, t2);
        n = PZ ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN);
      }
      const float mu = cd[CD_MU];
      const float bq = cd[CD_IMP + IM_B];
      const int r0 = 4 * s;
      const float a0 = rec[R::AB + r0], a1 = rec[R::AB + r0 + 1], a2 = rec[R::AB + r0 + 2],
                  a3 = rec[R::AB + r0 + 3];
      const V3 vb = (n * ((a0 + a1) + (a2 + a3)) + t1 * (mu * (a0 - a1)) +
                     t2 * (mu * (a2 - a3))) * -bq;
      cba[l] = cba[l] + cross(pos - rc[l], vb);
      cbv[l] = cbv[l] + vb;
      if (la >= 0) {
        const int lc = la >= 0 ? la : 0;
        cba[lc] = cba[lc] - cross(pos - rc[lc], vb);
        cbv[lc] = cbv[lc] - vb;
      }
    }
  }
  // leaf to root: each link's force, its accelerations' dof terms, then its
  // velocity's dof terms; a link's cotangents are final once its children
  // (higher indices) have added theirs
#pragma unroll
  for (int l = L - 1; l >= 0; --l) {
    const int par = T::PARENT[l], da = T::DADR[l];
    const float m = link_rec<T>(M, l)[LK_MASS];
    // f_l = I x_l + cd_l x* (I cd_l)
    V3 ta, tv;
    inertia_mul(cc[l], ci[l], m, fba[l], fbv[l], ta, tv);
    xba[l] = xba[l] + ta, xbv[l] = xbv[l] + tv;
    V3 ha, hv;
    inertia_mul(cc[l], ci[l], m, cda[l], cdv[l], ha, hv);
    cba[l] = cba[l] + (cross(ha, fba[l]) + cross(hv, fbv[l]));
    cbv[l] = cbv[l] + cross(hv, fba[l]);
    {
      V3 ma, mv, na, nv;
      cross_motion(cda[l], cdv[l], fba[l], fbv[l], ma, mv);
      inertia_mul(cc[l], ci[l], m, ma, mv, na, nv);
      cba[l] = cba[l] - na, cbv[l] = cbv[l] - nv;
    }
    // x_l = x_parent + sum_k cross_motion(s_k, cdof_d) qd_d
    if (T::LTYPE[l] == 0) {
      V3 sa = v3zero(), sv = v3zero();
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sa = sa + ca[da + k] * qd_in[da + k];
        sv = sv + cv[da + k] * qd_in[da + k];
      }
      V3 sba = v3zero(), sbv = v3zero();
#pragma unroll
      for (int k = 3; k < 6; ++k) {
        const int d = da + k;
        V3 oa, ov, pa, pv;
        cross_motion(sa, sv, ca[d], cv[d], oa, ov);
        g[d] += dot(xba[l], oa) + dot(xbv[l], ov);
        cross_force(ca[d], cv[d], xba[l], xbv[l], pa, pv);
        sba = sba + pa * qd_in[d], sbv = sbv + pv * qd_in[d];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) g[da + k] += dot(sba, ca[da + k]) + dot(sbv, cv[da + k]);
    } else {
      constexpr int NJ = 3;
      V3 sa = par >= 0 ? cda[par >= 0 ? par : 0] : v3zero();
      V3 sv = par >= 0 ? cdv[par >= 0 ? par : 0] : v3zero();
      V3 sba[NJ], sbv[NJ];
#pragma unroll
      for (int k = 0; k < NJ; ++k) {
        sba[k] = sbv[k] = v3zero();
        if (k < T::LTYPE[l]) {
          const int d = da + k;
          if (k > 0) sa = sa + ca[d - 1] * qd_in[d - 1], sv = sv + cv[d - 1] * qd_in[d - 1];
          V3 oa, ov, pa, pv;
          cross_motion(sa, sv, ca[d], cv[d], oa, ov);
          g[d] += dot(xba[l], oa) + dot(xbv[l], ov);
          cross_force(ca[d], cv[d], xba[l], xbv[l], pa, pv);
          sba[k] = pa * qd_in[d], sbv[k] = pv * qd_in[d];
        }
      }
      // s_k = cd_parent + sum_{j<k} cdof_j qd_j
      V3 aa = v3zero(), av = v3zero();
#pragma unroll
      for (int k = NJ - 1; k >= 0; --k) {
        if (k < T::LTYPE[l]) {
          if (k + 1 < T::LTYPE[l])
            g[da + k] += dot(aa, ca[da + k]) + dot(av, cv[da + k]);
          aa = aa + sba[k], av = av + sbv[k];
        }
      }
      if (par >= 0) {
        const int p = par >= 0 ? par : 0;
        cba[p] = cba[p] + aa, cbv[p] = cbv[p] + av;
      }
    }
    // cd_l = cd_parent + sum cdof_d qd_d
    const int nd = T::LTYPE[l] == 0 ? 6 : T::LTYPE[l];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nd) {
        const int d = da + k;
        g[d] += dot(cba[l], ca[d]) + dot(cbv[l], cv[d]);
      }
    }
    if (par >= 0) {
      const int p = par >= 0 ? par : 0;
      cba[p] = cba[p] + cba[l], cbv[p] = cbv[p] + cbv[l];
      xba[p] = xba[p] + xba[l], xbv[p] = xbv[p] + xbv[l];
    }
  }
  // limit rows: aref = -b side qd_d - k imp pos
#pragma unroll
  for (int i = 0; i < T::NLIM; ++i) {
    const int d = T::LIM_D[i], qi = T::LIM_Q[i];
    const float* df = dof_rec<T>(M, d);
    const float pmin = q_in[qi] - df[DF_LO];
    const float pmax = df[DF_HI] - q_in[qi];
    const float pos = fminf(fminf(pmin, pmax), 0.f);
    if (pos < 0.f) {
      const float side = pmin < pmax ? 1.f : -1.f;
      g[d] += rec[R::AB + NRC + i] * (-df[DF_IMP + IM_B] * side);
    }
  }
#pragma unroll
  for (int d = 0; d < NV; ++d) gqd[d] = g[d];
}

// one tree on a free root over a z-up plane: moving the root along x or y
// moves everything rigidly, which no ingredient sees (frames about the root
// COM, plane distances along z); only the position update passes through
template <class T>
constexpr bool xy_invariant() {
  int roots = 0;
  for (int l = 0; l < T::L; ++l) roots += T::PARENT[l] < 0 ? 1 : 0;
  return T::PLANE_Z != 0 && roots == 1 && T::PARENT[0] < 0 && T::LTYPE[0] == 0 &&
         T::QADR[0] == 0;
}

template <class T>
BGX_D float adjoint_b(const float* __restrict__ M, const float* q_in, const float* qd_in,
                      const float* act, RecPtr rec, int dir) {
  if constexpr (xy_invariant<T>())
    if (dir < 2) return rec[AdjRec<T>::GQ1 + dir];
  return dir < T::NQ ? adjoint_b_q<T>(M, q_in, qd_in, act, rec, dir)
                     : adjoint_b_qd<T>(M, q_in, qd_in, act, rec, dir - T::NQ);
}

}  // namespace bgx
