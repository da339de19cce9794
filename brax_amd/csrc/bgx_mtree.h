// Trees, several lanes per environment.
//
// One thread per env leaves a large tree's batch short of wavefronts (16,384
// humanoids are 256 wavefronts of 64 for the chip's 1,024 SIMDs) and spills
// its registers.  Here an env is a group of G lanes (G = the next power of
// two >= max(dofs, contact candidates); 64 / G envs per wavefront).  Lane d
// is dof d; the lane of a link's first dof also holds the link; lane c holds
// contact candidate c.  The step is the reference's (pipeline.py:64-94), the
// same algorithm as the one-thread kernel (bgx_step.h) laid out across
// lanes:
//  * root-to-leaf passes (kinematics, link velocities, RNEA accelerations)
//    go level by level, a link reading its parent's result from LDS;
//  * leaf-to-root passes (subtree forces, composite inertias) go level by
//    level the other way, children added in the order the serial pass adds
//    them;
//  * the mass matrix is built a row per lane (CRBA, mass.py:27-83) and
//    inverted in place by Gauss-Jordan elimination, one pivot per step with
//    the pivot row broadcast through LDS, so that M^-1 v is one broadcast
//    matrix-vector product with no serial chain;
//  * a candidate's four contact rows and a dof's limit row sit on their
//    lanes (an inactive candidate's rows are zero, which leaves every
//    projected-gradient iterate as if they were dropped, constraint.py:174),
//    and the solver's sums are butterfly reductions over the group.
// The topology must be one tree whose root link (link 0) has the free joint
// or a hinge stack, without sphere / capsule pairs, with at most one
// actuator per dof acting on that dof's own coordinate (codegen checks).
#pragma once

#include <hip/hip_runtime.h>

#include "bgx_chain.h"
#include "bgx_launch.h"
#include "bgx_step.h"

namespace bgx {

template <class T>
struct MTab {
  static constexpr int L = T::L, NV = T::NV, NQ = T::NQ, NC = T::NCAND;
  static constexpr int W = NV > NC ? NV : NC;
  static constexpr int G = W <= 2 ? 2 : W <= 4 ? 4 : W <= 8 ? 8 : W <= 16 ? 16 : 32;
  static constexpr int E = WAVE / G;  // envs per wavefront
  struct Tab {
    int depth[L];
    int nch[L];
    int ch[L][L];      // children, highest index first (the serial passes' order)
    int lane_link[G];  // the link whose first dof is this lane, or -1
    int lim_of_dof[NV];
    int act_of_dof[NV];
    int maxd;
    // for the solver's products: the candidates on each dof's subtree, and
    // the dofs on each candidate's path from the root (root first)
    int ncsub[NV];
    int csub[NV][NC > 0 ? NC : 1];
    int npath[NC > 0 ? NC : 1];
    int path[NC > 0 ? NC : 1][NV];
  };
  static constexpr Tab make() {
    Tab t{};
    t.maxd = 0;
    for (int l = 0; l < L; ++l) {
      int dd = 0;
      for (int x = l; T::PARENT[x] >= 0; x = T::PARENT[x]) ++dd;
      t.depth[l] = dd;
      if (dd > t.maxd) t.maxd = dd;
      t.nch[l] = 0;
    }
    for (int l = L - 1; l >= 0; --l) {
      const int p = T::PARENT[l];
      if (p >= 0) t.ch[p][t.nch[p]++] = l;
    }
    for (int g = 0; g < G; ++g) {
      t.lane_link[g] = -1;
      for (int l = 0; l < L; ++l)
        if (T::DADR[l] == g) t.lane_link[g] = l;
    }
    for (int d = 0; d < NV; ++d) t.lim_of_dof[d] = -1, t.act_of_dof[d] = -1;
    for (int i = 0; i < T::NLIM; ++i) t.lim_of_dof[T::LIM_D[i]] = i;
    for (int u = 0; u < T::NU; ++u) t.act_of_dof[T::ACT_D[u]] = u;
    for (int d = 0; d < NV; ++d) {
      t.ncsub[d] = 0;
      for (int c = 0; c < NC; ++c) {
        bool in = false;  // the candidate's link is the dof's link or below it
        for (int x = T::CAND_LINK[c]; x >= 0; x = T::PARENT[x]) in = in || x == T::DOF_LINK[d];
        if (in) t.csub[d][t.ncsub[d]++] = c;
      }
    }
    for (int c = 0; c < NC; ++c) {
      int chain[L > 0 ? L : 1], n = 0;
      for (int x = T::CAND_LINK[c]; x >= 0; x = T::PARENT[x]) chain[n++] = x;
      t.npath[c] = 0;
      for (int i = n - 1; i >= 0; --i) {
        const int x = chain[i];
        const int nd = T::LTYPE[x] == 0 ? 6 : T::LTYPE[x];
        for (int k = 0; k < nd; ++k) t.path[c][t.npath[c]++] = T::DADR[x] + k;
      }
    }
    return t;
  }
  static constexpr Tab tab = make();
  static constexpr int MAXD = tab.maxd;
  static constexpr int MAXCH = [] {
    int m = 0;
    for (int l = 0; l < L; ++l) m = tab.nch[l] > m ? tab.nch[l] : m;
    return m;
  }();
  static constexpr int MAXCSUB = [] {
    int m = 0;
    for (int d = 0; d < NV; ++d) m = tab.ncsub[d] > m ? tab.ncsub[d] : m;
    return m;
  }();
  static constexpr int MAXPATH = [] {
    int m = 0;
    for (int c = 0; c < NC; ++c) m = tab.npath[c] > m ? tab.npath[c] : m;
    return m;
  }();
  // LDS of one env (floats)
  static constexpr int LJ = 0;            // [L][7] joint transforms (pos, rot)
  static constexpr int LX = LJ + 7 * L;   // [L][7] world frames
  static constexpr int CDF = LX + 7 * L;  // [NV][6] cdof (ang, vel)
  static constexpr int ZV = CDF + 6 * NV;  // [NV] a per-dof vector
  static constexpr int LA = ZV + NV;      // [L][6] link velocities / accelerations
  static constexpr int LB = LA + 6 * L;   // [L][6] link velocities at qd
  static constexpr int LF = LB + 6 * L;   // [L][6] subtree sums
  static constexpr int CF = LF + 6 * L;   // [NC][6] candidate forces (torque, force)
  static constexpr int PV = CF + 6 * (NC > 0 ? NC : 1);  // [2][NV] pivot rows
  static constexpr int CI = PV + 2 * NV;  // [L][10] composite inertias (cc, ci, mass)
  static constexpr int FS = CI + 10 * L;  // [NV][6] I_c cdof
  static constexpr int QS = FS + 6 * NV;  // [NQ + NV] the state, for the frames
  static constexpr int ZW = QS + NQ + NV;  // [NV] a second per-dof vector
  // M, then M^-1, row d at MI + d MIS (an odd stride: lanes reading their
  // rows hit distinct banks)
  static constexpr int MIS = NV % 2 ? NV + 2 : NV + 1;
  static constexpr int MI = ZW + NV;
  static constexpr int SIZE = MI + NV * MIS;
};

// whether topology T can run the multi-lane tree kernel
template <class T>
constexpr bool mtree_ok() {
  if (T::PER_ENV || T::PAIRS || T::NV > 32 || T::NCAND > 32) return false;
  for (int l = 1; l < T::L; ++l)
    if (T::PARENT[l] < 0 || T::LTYPE[l] == 0) return false;  // one tree, free joint at the root
  if (T::L < 1 || T::PARENT[0] >= 0) return false;
  if (T::LTYPE[0] == 0 && T::QADR[0] != 0) return false;
  for (int u = 0; u < T::NU; ++u) {
    for (int v = 0; v < u; ++v)
      if (T::ACT_D[u] == T::ACT_D[v]) return false;
    const int d = T::ACT_D[u], l = T::DOF_LINK[d];
    if (T::LTYPE[l] == 0 || T::ACT_Q[u] != T::QADR[l] + (d - T::DADR[l])) return false;
  }
  return true;
}

template <int G>
BGX_D float gsum(float x) {
#pragma unroll
  for (int m = 1; m < G; m *= 2) x = x + sh_xor<G>(x, m);
  return x;
}
BGX_D bool wave_any(bool b) { return __ballot(b) != 0ull; }

BGX_D V3 ld3l(const float* p) { return v3(p[0], p[1], p[2]); }
BGX_D Q4 ld4l(const float* p) { return Q4{p[0], p[1], p[2], p[3]}; }
BGX_D void st3l(float* p, V3 v) { p[0] = v.x, p[1] = v.y, p[2] = v.z; }
BGX_D void st4l(float* p, Q4 v) { p[0] = v.w, p[1] = v.x, p[2] = v.y, p[3] = v.z; }

// per-lane roles and per-env state of the multi-lane step
template <class T>
struct MLane {
  int g, gb;       // lane in the group, the group's first lane in the wavefront
  int d, dl, dk;   // dof (or -1), its link, its index in the link's stack
  int ll, lld;     // the link this lane holds (or -1), its depth
  int c;           // contact candidate (or -1)
  int li, ai;      // limit row and actuator of the dof (or -1)
  unsigned cmask;  // the candidates on the dof's subtree (bit c)
  unsigned pmask;  // the dofs on the candidate's root path (bit d)
  float rq[7];     // the root's free-joint coordinates (every lane)
  float qh, qd;    // this dof's hinge coordinate and velocity
};

template <class T>
BGX_D MLane<T> mlane_roles(int g, int gb) {
  using MT = MTab<T>;
  MLane<T> m;
  m.g = g, m.gb = gb;
  m.d = g < T::NV ? g : -1;
  const int dc = m.d >= 0 ? m.d : 0;
  m.dl = m.d >= 0 ? T::DOF_LINK[dc] : -1;
  m.dk = m.d >= 0 ? m.d - T::DADR[T::DOF_LINK[dc]] : 0;
  m.ll = MT::tab.lane_link[g];
  m.lld = m.ll >= 0 ? MT::tab.depth[m.ll] : -1;
  m.c = g < T::NCAND ? g : -1;
  m.li = m.d >= 0 ? MT::tab.lim_of_dof[dc] : -1;
  m.ai = m.d >= 0 ? MT::tab.act_of_dof[dc] : -1;
  m.cmask = 0u, m.pmask = 0u;
  if (m.d >= 0)
    for (int k = 0; k < MT::tab.ncsub[dc]; ++k) m.cmask |= 1u << MT::tab.csub[dc][k];
  if (m.c >= 0)
    for (int k = 0; k < MT::tab.npath[m.c]; ++k) m.pmask |= 1u << MT::tab.path[m.c][k];
  return m;
}

// one substep of one env on its group's lanes.  act: the env's action row
// (null: zero); lds: the env's MTab<T>::SIZE floats.  probe (non-null):
// qf_smooth and qdd of this dof, and (group lane 0) the active contacts,
// solver iterations and line-search trials.
template <class T>
BGX_D void mtree_substep(const float* __restrict__ M, MLane<T>& s, const float* act,
                         float* lds, float* probe) {
  using MT = MTab<T>;
  constexpr int G = MT::G, NV = T::NV, L = T::L, NC = T::NCAND, MAXD = MT::MAXD;
  const auto& tb = MT::tab;
  const int d = s.d, ll = s.ll, c = s.c;
  const int dlc = s.dl >= 0 ? s.dl : 0, llc = ll >= 0 ? ll : 0, cc_ = c >= 0 ? c : 0;

  // ---- kinematics.forward (kinematics.py:31-108): joint transforms, then
  // world frames root to leaves (rotations normalized at the end, :106)
  if (d >= 0) lds[MT::ZV + d] = s.qh;
  __syncthreads();
  V3 xp = v3zero();
  Q4 xr = qident();
  if (ll >= 0) {
    const float* lk = link_rec<T>(M, ll);
    V3 p = v3zero();
    Q4 r = qident();
    if (T::LTYPE[ll] == 0) {
      p = v3(s.rq[0], s.rq[1], s.rq[2]);
      r = Q4{s.rq[3], s.rq[4], s.rq[5], s.rq[6]};
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k < T::LTYPE[ll]) {
          const int dd = T::DADR[ll] + k;
          const V3 ang = dof_ang<T>(M, dd), vel = dof_vel<T>(M, dd);
          const float qk = lds[MT::ZV + dd];
          const Q4 rk = qnormalize(qrot_axis(ang, qk));
          const V3 pk = vel * qk;
          if (k == 0) {
            p = pk, r = rk;
          } else {
            p = p + rotate(pk, r);
            r = qmul(r, rk);
          }
        }
      }
    }
    const V3 jpos = ldv3(lk + LK_JPOS);
    p = (p + jpos) - rotate(jpos, r);
    p = ldv3(lk + LK_TPOS) + rotate(p, ldq4(lk + LK_TROT));
    r = qmul(ldq4(lk + LK_TROT), r);
    xp = p, xr = r;
  }
#pragma unroll
  for (int lev = 0; lev <= MAXD; ++lev) {
    if (ll >= 0 && s.lld == lev) {
      const int par = T::PARENT[ll];
      if (par >= 0) {
        const float* pf = lds + MT::LX + 7 * par;
        const V3 pp = ld3l(pf);
        const Q4 pr = ld4l(pf + 3);
        xp = pp + rotate(xp, pr);
        xr = qmul(pr, xr);
      }
      st3l(lds + MT::LX + 7 * ll, xp);
      st4l(lds + MT::LX + 7 * ll + 3, xr);
    }
    __syncthreads();
  }
  if (ll >= 0) {
    xr = qnormalize(xr);
    st4l(lds + MT::LX + 7 * ll + 3, xr);
  }
  // ---- root COM (dynamics.py:37-43)
  float m_l = 0.f;
  V3 xi = v3zero();
  if (ll >= 0) {
    const float* lk = link_rec<T>(M, ll);
    m_l = lk[LK_MASS];
    xi = xp + rotate(ldv3(lk + LK_IPOS), xr);
  }
  V3 sm = xi * m_l;
  sm = v3(gsum<G>(sm.x), gsum<G>(sm.y), gsum<G>(sm.z));
  const float ims = rcp(gsum<G>(m_l));
  const V3 rc{sm.x * ims, sm.y * ims, sm.z * ims};
  __syncthreads();  // world frames normalized in LDS

  // ---- cinr of the link (dynamics.py:44-55)
  V3 cc = v3zero();
  Sym3 ci{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (ll >= 0) {
    const float* lk = link_rec<T>(M, ll);
    const Q4 xir = qmul(xr, ldq4(lk + LK_IROT));
    const V3 p = xi - rc;
    float R[3][3];
    quat_to_3x3(xir, R);
    const Sym3 I0 = ldsym(lk + LK_I);
    float RI[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const V3 cr = symmul(I0, V3{R[i][0], R[i][1], R[i][2]});
      RI[i][0] = cr.x, RI[i][1] = cr.y, RI[i][2] = cr.z;
    }
    auto rir = [&](int i, int j) {
      return RI[i][0] * R[j][0] + RI[i][1] * R[j][1] + RI[i][2] * R[j][2];
    };
    const float pp = dot(p, p);
    ci = Sym3{rir(0, 0) + (pp - p.x * p.x) * m_l, rir(1, 1) + (pp - p.y * p.y) * m_l,
              rir(2, 2) + (pp - p.z * p.z) * m_l, rir(0, 1) - p.x * p.y * m_l,
              rir(0, 2) - p.x * p.z * m_l, rir(1, 2) - p.y * p.z * m_l};
    cc = p * m_l;
  }
  // ---- cdof of the dof (dynamics.py:56-89)
  V3 ca = v3zero(), cv = v3zero();
  if (d >= 0) {
    const int l = dlc;
    const float* lk = link_rec<T>(M, l);
    const int par = T::PARENT[l], da = T::DADR[l];
    V3 Pp = v3zero();
    Q4 Pr = qident();
    if (T::LTYPE[l] == 0 || par >= 0) {
      const float* pf = lds + MT::LX + 7 * (T::LTYPE[l] == 0 ? l : (par >= 0 ? par : 0));
      Pp = ld3l(pf), Pr = ld4l(pf + 3);
    }
    const V3 t1p = Pp + rotate(ldv3(lk + LK_TPOS), Pr);
    const Q4 jr = qmul(Pr, ldq4(lk + LK_TROT));
    const V3 jp = t1p + rotate(ldv3(lk + LK_JPOS), jr);
    V3 aa = dof_ang<T>(M, d), vv = dof_vel<T>(M, d);
    if (T::LTYPE[l] != 0) {  // stacked joints (dynamics.py:59-83)
      V3 accp = v3zero();
      Q4 accr = qident();
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (k < s.dk) {
          const int dd = da + k;
          const V3 ang = dof_ang<T>(M, dd), vel = dof_vel<T>(M, dd);
          const float qk = lds[MT::ZV + dd];
          const Q4 rk = qnormalize(qrot_axis(ang, qk));
          accp = accp + rotate(vel * qk, accr);
          accr = qmul(accr, rk);
        }
      }
      const V3 a0 = rotate(aa, accr);
      vv = rotate(vv, accr) + cross(accp, a0);
      aa = a0;
    }
    aa = rotate(aa, jr);
    ca = aa;
    cv = vv - cross(rc - jp, aa);
    float* cf = lds + MT::CDF + 6 * d;
    st3l(cf, ca);
    st3l(cf + 3, cv);
  }
  __syncthreads();  // the hinge coordinates in ZV are read; qd replaces them
  if (d >= 0) lds[MT::ZV + d] = s.qd;
  __syncthreads();

  // link velocities under the per-dof vector in ZV, root to leaves
  // (dynamics.py:93-103), into LDS at base
  auto down_vel = [&](int base, V3& va, V3& vv) {
#pragma unroll
    for (int lev = 0; lev <= MAXD; ++lev) {
      if (ll >= 0 && s.lld == lev) {
        const int par = T::PARENT[ll];
        V3 a = v3zero(), v = v3zero();
        if (par >= 0) {
          const float* pf = lds + base + 6 * par;
          a = ld3l(pf), v = ld3l(pf + 3);
        }
        const int nd = T::LTYPE[ll] == 0 ? 6 : T::LTYPE[ll];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          if (k < nd) {
            const int dd = T::DADR[ll] + k;
            const float* cf = lds + MT::CDF + 6 * dd;
            const float z = lds[MT::ZV + dd];
            a = a + ld3l(cf) * z;
            v = v + ld3l(cf + 3) * z;
          }
        }
        va = a, vv = v;
        st3l(lds + base + 6 * ll, a);
        st3l(lds + base + 6 * ll + 3, v);
      }
      __syncthreads();
    }
  };
  // subtree sums leaf to root of the per-link pairs (fa, fv) this lane's
  // link contributes: S = own + children's S, children highest index first
  // (own first: project_forces_on's a = f + acc with acc = 0 + ...; own
  // last: crba's in-place cc[par] += cc[l]) into LDS at LF
  auto up_sum = [&](V3 fa, V3 fv, V3& sa, V3& sv) {
#pragma unroll
    for (int lev = MAXD; lev >= 0; --lev) {
      if (ll >= 0 && s.lld == lev) {
        V3 aa = v3zero(), av = v3zero();
#pragma unroll
        for (int k = 0; k < MT::MAXCH; ++k) {
          if (k < tb.nch[ll]) {
            const float* cf = lds + MT::LF + 6 * tb.ch[ll][k];
            aa = aa + ld3l(cf);
            av = av + ld3l(cf + 3);
          }
        }
        sa = fa + aa, sv = fv + av;
        st3l(lds + MT::LF + 6 * ll, sa);
        st3l(lds + MT::LF + 6 * ll + 3, sv);
      }
      __syncthreads();
    }
  };
  // the dof's generalized force from its link's subtree sum in LF
  auto project = [&]() {
    const float* sf = lds + MT::LF + 6 * dlc;
    return dot(cv, ld3l(sf + 3)) + dot(ca, ld3l(sf));
  };

  // ---- cd: link velocities at qd, kept in LB
  V3 cda = v3zero(), cdv = v3zero();
  down_vel(MT::LB, cda, cdv);

  // ---- RNEA bias (dynamics.py:105-184): accelerations root to leaves with
  // the root at -g and cdofd formed on the fly, then subtree forces
  const V3 grav = ldv3(M + H_GRAVITY);
  V3 acc_a = v3zero(), acc_v = v3zero();
#pragma unroll
  for (int lev = 0; lev <= MAXD; ++lev) {
    if (ll >= 0 && s.lld == lev) {
      const int par = T::PARENT[ll], da = T::DADR[ll];
      V3 a = v3zero(), v = v3zero() - grav;
      if (par >= 0) {
        const float* pf = lds + MT::LA + 6 * par;
        a = ld3l(pf), v = ld3l(pf + 3);
      }
      auto cdof_of = [&](int dd, V3& oa, V3& ov) {
        const float* cf = lds + MT::CDF + 6 * dd;
        oa = ld3l(cf), ov = ld3l(cf + 3);
      };
      auto qd_of = [&](int dd) { return lds[MT::ZV + dd]; };
      if (T::LTYPE[ll] == 0) {
        V3 sa = v3zero(), sv = v3zero();
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          V3 oa, ov;
          cdof_of(da + k, oa, ov);
          sa = sa + oa * qd_of(da + k);
          sv = sv + ov * qd_of(da + k);
        }
#pragma unroll
        for (int k = 3; k < 6; ++k) {
          V3 ma, mv, oa, ov;
          cdof_of(da + k, ma, mv);
          cross_motion(sa, sv, ma, mv, oa, ov);
          a = a + oa * qd_of(da + k);
          v = v + ov * qd_of(da + k);
        }
      } else {
        V3 sa = v3zero(), sv = v3zero();
        if (par >= 0) {
          const float* pf = lds + MT::LB + 6 * par;
          sa = ld3l(pf), sv = ld3l(pf + 3);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k < T::LTYPE[ll]) {
            const int dd = da + k;
            if (k > 0) {
              V3 pa, pv;
              cdof_of(dd - 1, pa, pv);
              sa = sa + pa * qd_of(dd - 1);
              sv = sv + pv * qd_of(dd - 1);
            }
            V3 ma, mv, oa, ov;
            cdof_of(dd, ma, mv);
            cross_motion(sa, sv, ma, mv, oa, ov);
            a = a + oa * qd_of(dd);
            v = v + ov * qd_of(dd);
          }
        }
      }
      acc_a = a, acc_v = v;
      st3l(lds + MT::LA + 6 * ll, a);
      st3l(lds + MT::LA + 6 * ll + 3, v);
    }
    __syncthreads();
  }
  V3 fa = v3zero(), fv = v3zero();
  if (ll >= 0) {
    V3 f1a, f1v, f2a, f2v, f3a, f3v;
    inertia_mul(cc, ci, m_l, acc_a, acc_v, f1a, f1v);
    inertia_mul(cc, ci, m_l, cda, cdv, f2a, f2v);
    cross_force(cda, cdv, f2a, f2v, f3a, f3v);
    fa = f1a + f3a, fv = f1v + f3v;
  }
  {
    V3 sa, sv;
    up_sum(fa, fv, sa, sv);
  }
  // ---- qf_smooth = passive - bias + tau (dynamics.py:187-236, actuator.py:23-57)
  float qf = 0.f;
  if (d >= 0) {
    const float bias = project();
    const float* df = dof_rec<T>(M, d);
    float p = -df[DF_DAMPING] * s.qd;
    if (T::LTYPE[dlc] != 0) p = -s.qh * df[DF_STIFFNESS] + p;
    qf = p - bias;
    if (s.ai >= 0) {
      const float* ac = M + T::OFF_ACT + s.ai * ACT_STRIDE;
      const float ctl = clipf(act ? act[s.ai] : 0.f, ac[AC_CTRL_LO], ac[AC_CTRL_HI]);
      const float b = ac[AC_GEAR] * (s.qh * ac[AC_BIAS_Q] + s.qd * ac[AC_BIAS_QD]);
      float f = ac[AC_GAIN] * ctl + b;
      if (f < ac[AC_FORCE_LO]) f = ac[AC_FORCE_LO];
      if (f > ac[AC_FORCE_HI]) f = ac[AC_FORCE_HI];
      qf += f * ac[AC_GEAR];
    }
  }
  if (probe && d >= 0) probe[d] = qf;

  // ---- mass.matrix (mass.py:27-83): composite inertias leaf to root, then
  // row d of M on lane d, M_dj = cdof_j . (I_c(link d) cdof_d) on the
  // dof-ancestor pattern (its transpose above the diagonal)
#pragma unroll
  for (int lev = MAXD; lev >= 0; --lev) {
    if (ll >= 0 && s.lld == lev) {
#pragma unroll
      for (int k = 0; k < MT::MAXCH; ++k) {
        if (k < tb.nch[ll]) {
          const float* cf = lds + MT::CI + 10 * tb.ch[ll][k];
          cc = cc + ld3l(cf);
          ci = symadd(ci, Sym3{cf[3], cf[4], cf[5], cf[6], cf[7], cf[8]});
          m_l = m_l + cf[9];
        }
      }
      float* cf = lds + MT::CI + 10 * ll;
      st3l(cf, cc);
      cf[3] = ci.xx, cf[4] = ci.yy, cf[5] = ci.zz, cf[6] = ci.xy, cf[7] = ci.xz, cf[8] = ci.yz;
      cf[9] = m_l;
    }
    __syncthreads();
  }
  V3 Fa = v3zero(), Fv = v3zero();
  if (d >= 0) {
    const float* cf = lds + MT::CI + 10 * dlc;
    inertia_mul(ld3l(cf), Sym3{cf[3], cf[4], cf[5], cf[6], cf[7], cf[8]}, cf[9], ca, cv, Fa,
                Fv);
    float* fs = lds + MT::FS + 6 * d;
    st3l(fs, Fa);
    st3l(fs + 3, Fv);
  }
  __syncthreads();
  float* mrow = lds + MT::MI + (d >= 0 ? d : 0) * MT::MIS;  // row d of M, then of M^-1
  {
    const unsigned danc = d >= 0 ? T::DANC[d] : 0u;
    const float arm = d >= 0 ? dof_rec<T>(M, d)[DF_ARMATURE] : 1.f;
#pragma unroll 2
    for (int j = 0; j < NV; ++j) {
      const float* cf = lds + MT::CDF + 6 * j;
      const float* fs = lds + MT::FS + 6 * j;
      const float lower = dot(ld3l(cf + 3), Fv) + dot(ld3l(cf), Fa);
      const float upper = dot(cv, ld3l(fs + 3)) + dot(ca, ld3l(fs));
      const bool in_lower = j < d && ((danc >> j) & 1u);
      const bool in_upper = j > d && ((T::DANC[j] >> (d >= 0 ? d : 0)) & 1u);
      const float mj = j == d ? (dot(cv, Fv) + dot(ca, Fa)) + arm
                              : (in_lower ? lower : (in_upper ? upper : 0.f));
      if (d >= 0) mrow[j] = mj;
    }
  }
  // ---- M^-1 in place by Gauss-Jordan elimination (M is positive definite:
  // no pivoting), row d on lane d, the pivot row read by every lane
#pragma unroll 1
  for (int k = 0; k < NV; ++k) {
    __syncthreads();
    const float* prow = lds + MT::MI + k * MT::MIS;
    const float pk = rcp(prow[k]);
    __syncthreads();  // every lane has the pivot before row k changes
    if (d == k) {
#pragma unroll 4
      for (int j = 0; j < NV; ++j) mrow[j] = j == k ? pk : mrow[j] * pk;
    }
    __syncthreads();
    if (d >= 0 && d != k) {
      const float f = mrow[k];
#pragma unroll 4
      for (int j = 0; j < NV; ++j) mrow[j] = j == k ? -f * pk : mrow[j] - f * prow[j];
    }
  }
  __syncthreads();
  // M^-1 times the per-dof vector in ZV, on this dof
  auto minv_mul = [&]() {
    float z = 0.f;
#pragma unroll
    for (int j = 0; j < NV; ++j) z += mrow[j] * lds[MT::ZV + j];
    return z;
  };
  // the velocity (angular, linear) of candidate c's link under the per-dof
  // vector at base: its root path's cdof times the vector, root first (the
  // order of dynamics.py:93-103's scan)
  auto path_vel = [&](int base, V3& wa, V3& wv) {
    wa = v3zero(), wv = v3zero();
#pragma unroll
    for (int dd = 0; dd < NV; ++dd) {  // ancestors first: dof indices grow down the tree
      if ((s.pmask >> dd) & 1u) {
        const float* cf = lds + MT::CDF + 6 * dd;
        const float z = lds[base + dd];
        wa = wa + ld3l(cf) * z;
        wv = wv + ld3l(cf + 3) * z;
      }
    }
  };
  // ---- u = M^-1 qf_smooth
  if (d >= 0) lds[MT::ZV + d] = qf;
  __syncthreads();
  const float u = d >= 0 ? minv_mul() : 0.f;
  if (d >= 0) lds[MT::ZW + d] = u;
  __syncthreads();

  // ---- contacts (contact.py + MJX plane collisions): candidate c's rows
  bool active = false;
  V3 r = v3zero(), d0 = v3zero(), d1 = v3zero(), d2 = v3zero();
  float sdiag = 0.f, sb[4] = {0.f, 0.f, 0.f, 0.f};
  if (c >= 0) {
    const int cl = T::CAND_LINK[cc_];
    const float* lf = lds + MT::LX + 7 * cl;
    const V3 xpl = ld3l(lf);
    const Q4 xrl = ld4l(lf + 3);
    const float* cd = cand_rec<T>(M, cc_);
    const V3 gp = xpl + rotate(ldv3(cd + CD_GPOS), xrl);
    V3 ctr = gp, axis = v3zero();
    if (T::CAND_KIND[cc_] == KIND_CAPSULE) {
      axis = quat_zaxis(qmul(xrl, ldq4(cd + CD_GQUAT)));
      ctr = gp + axis * (cd[CD_HALFLEN] * cd[CD_SIGN]);
    }
    const float dist = T::PLANE_Z ? (ctr.z - cd[CD_PPOS + 2]) - cd[CD_RADIUS]
                                  : dot(ctr - ldv3(cd + CD_PPOS), ldv3(cd + CD_PN)) -
                                        cd[CD_RADIUS];
    active = dist < 0.f;
    if (active) {
      V3 pos, t1, t2;
      contact_frame<T>(M, cc_, ctr, axis, dist, -1, pos, t1, t2);
      const V3 n = T::PLANE_Z ? v3(0.f, 0.f, 1.f) : ldv3(cd + CD_PN);
      const float mu = cd[CD_MU];
      r = pos - rc;
      d0 = n, d1 = t1 * mu, d2 = t2 * mu;
      const float* im = cd + CD_IMP;
      const float imp = impedance(im, dist);
      const float tw = cd[CD_TW];
      sdiag = (tw + mu * mu * tw) * (2.f * mu * mu * imp_ratio(imp));
      const float kt = im[IM_K] * imp * dist;
      const float bq = im[IM_B];
      // b = J u - aref, aref = -b J qd - k imp dist (constraint.py:58-65, :205-210)
      V3 ua, uv;
      path_vel(MT::ZW, ua, uv);
      const float* fq = lds + MT::LB + 6 * cl;
      const V3 vu = uv + cross(ua, r);
      const V3 vq = ld3l(fq + 3) + cross(ld3l(fq), r);
      const float e0u = dot(d0, vu), e1u = dot(d1, vu), e2u = dot(d2, vu);
      const float e0q = dot(d0, vq), e1q = dot(d1, vq), e2q = dot(d2, vq);
      sb[0] = (e0u + e1u) + (bq * (e0q + e1q) + kt);
      sb[1] = (e0u - e1u) + (bq * (e0q - e1q) + kt);
      sb[2] = (e0u + e2u) + (bq * (e0q + e2q) + kt);
      sb[3] = (e0u - e2u) + (bq * (e0q - e2q) + kt);
    }
  }
  // ---- the dof's limit row (constraint.py:101-132)
  float lside = 0.f, ldiag = 0.f, lb = 0.f;
  if (s.li >= 0) {
    const float* df = dof_rec<T>(M, d);
    const float pmin = s.qh - df[DF_LO];
    const float pmax = df[DF_HI] - s.qh;
    const float pos = fminf(fminf(pmin, pmax), 0.f);
    const bool act_l = pos < 0.f;
    lside = act_l ? (pmin < pmax ? 1.f : -1.f) : 0.f;
    const float* im = df + DF_IMP;
    const float imp = impedance(im, pos);
    const float aref = act_l ? (-im[IM_B] * (lside * s.qd) - im[IM_K] * imp * pos) : 0.f;
    ldiag = act_l ? df[DF_INVWEIGHT] * imp_ratio(imp) : 0.f;
    lb = lside * u - aref;
  }
  const float n_active = gsum<G>(active ? 1.f : 0.f);
  const bool has_rows = n_active > 0.f || gsum<G>(lside != 0.f ? 1.f : 0.f) > 0.f;

  // ---- products with A = J M^-1 J^T + diag through the tree: forces of
  // in's rows at the contacts, to the links, subtree sums, projections on
  // cdof, M^-1, link velocities, contact-point velocities; z: the dof's
  // component of M^-1 J^T in
  auto apply_jt = [&](const float* in, float in_l) {
    if (c >= 0) {
      const V3 w = d0 * ((in[0] + in[1]) + (in[2] + in[3])) + d1 * (in[0] - in[1]) +
                   d2 * (in[2] - in[3]);
      const V3 ta = cross(r, w);
      float* cf = lds + MT::CF + 6 * cc_;
      st3l(cf, active ? ta : v3zero());
      st3l(cf + 3, active ? w : v3zero());
    }
    __syncthreads();
    // the dof's generalized force: its cdof against the forces of the
    // candidates on its subtree (dynamics.py:174-182's sums, gathered)
    float tau = 0.f;
    if (d >= 0) {
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        if ((s.cmask >> k) & 1u) {
          const float* cf = lds + MT::CF + 6 * k;
          tau += dot(cv, ld3l(cf + 3)) + dot(ca, ld3l(cf));
        }
      }
      tau += lside * in_l;
      lds[MT::ZV + d] = tau;
    }
    __syncthreads();
    return d >= 0 ? minv_mul() : 0.f;
  };
  auto apply_a = [&](const float* in, float in_l, float* out, float& out_l) {
    const float z = apply_jt(in, in_l);
    if (d >= 0) lds[MT::ZW + d] = z;
    __syncthreads();
    if (c >= 0) {
      V3 wa, wv;
      path_vel(MT::ZW, wa, wv);
      const V3 vp = wv + cross(wa, r);
      const float e0 = dot(d0, vp), e1 = dot(d1, vp), e2 = dot(d2, vp);
      out[0] = active ? e0 + e1 + sdiag * in[0] : 0.f;
      out[1] = active ? e0 - e1 + sdiag * in[1] : 0.f;
      out[2] = active ? e0 + e2 + sdiag * in[2] : 0.f;
      out[3] = active ? e0 - e2 + sdiag * in[3] : 0.f;
    }
    out_l = lside * z + ldiag * in_l;
  };

  // ---- jaxopt.ProjectedGradient: FISTA + backtracking line search from 0
  // (the one-thread kernel's loop; an env's control is uniform over its
  // group, the wavefront runs until every env of it has finished)
  float X[4] = {0.f, 0.f, 0.f, 0.f}, Y[4] = {0.f, 0.f, 0.f, 0.f};
  float Gr[4] = {0.f, 0.f, 0.f, 0.f}, Rr[4] = {0.f, 0.f, 0.f, 0.f};
  float xl = 0.f, yl = 0.f, gl = 0.f, rl = 0.f;
  const int maxiter = (int)M[H_MAXITER], maxls = (int)M[H_MAXLS];
  const float tol = M[H_TOL], lseps = M[H_LSEPS];
  float tk = 1.f, eta = 1.f, err = INFINITY;
  bool running = has_rows;
  int iters = 0, trials = 0;
  const bool has_l = s.li >= 0;
  for (int it = 0; it < maxiter; ++it) {
    if (it > 0) running = running && (err > tol);
    if (!wave_any(running)) break;
    // gradient at y: r = A y + b, g = A^T r
    float o[4], ol;
    {
      apply_a(Y, yl, o, ol);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (running) Rr[k] = active ? o[k] + sb[k] : 0.f;
      if (running) rl = ol + lb;
    }
    float fy = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) fy += active ? 0.5f * Rr[k] * Rr[k] : 0.f;
    fy += has_l ? 0.5f * rl * rl : 0.f;
    fy = gsum<G>(fy);
    {
      apply_a(Rr, rl, o, ol);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (running) Gr[k] = o[k];
      if (running) gl = ol;
    }
    // line search: trial x+ = relu(y - eta g), residual into r
    bool cond = running;
    for (int ls = 0; ls <= maxls; ++ls) {
      if (!wave_any(cond)) break;
      if (ls > 0 && cond) eta *= 0.5f;
      const float e_ = eta;
      float in[4], in_l;
#pragma unroll
      for (int k = 0; k < 4; ++k) in[k] = fmaxf(Y[k] - e_ * Gr[k], 0.f);
      in_l = fmaxf(yl - e_ * gl, 0.f);
      apply_a(in, in_l, o, ol);
      float fx = 0.f, dg = 0.f, dd = 0.f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float rv = o[k] + sb[k];
        if (cond) Rr[k] = rv;
        const float df = in[k] - Y[k];
        fx += active ? 0.5f * rv * rv : 0.f;
        dg += active ? df * Gr[k] : 0.f;
        dd += active ? df * df : 0.f;
      }
      {
        const float rv = ol + lb;
        if (cond) rl = rv;
        const float df = in_l - yl;
        fx += has_l ? 0.5f * rv * rv : 0.f;
        dg += has_l ? df * gl : 0.f;
        dd += has_l ? df * df : 0.f;
      }
      fx = gsum<G>(fx), dg = gsum<G>(dg), dd = gsum<G>(dd);
      if (cond) {
        trials += 1;
        cond = e_ * (fx - fy) > e_ * dg + 0.5f * dd + lseps;
      }
    }
    // error at x+: |relu(x+ - grad f(x+)) - x+| with grad = A^T (A x+ + b)
    const float e_ = eta;
    apply_a(Rr, rl, o, ol);
    float e2 = 0.f;
    float xp[4], xpl;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      xp[k] = fmaxf(Y[k] - e_ * Gr[k], 0.f);
      const float dv = fmaxf(xp[k] - o[k], 0.f) - xp[k];
      e2 += active ? dv * dv : 0.f;
    }
    xpl = fmaxf(yl - e_ * gl, 0.f);
    {
      const float dv = fmaxf(xpl - ol, 0.f) - xpl;
      e2 += has_l ? dv * dv : 0.f;
    }
    e2 = gsum<G>(e2);
    if (running) {
      err = fsqrt(e2);
      const float tn = 0.5f * (1.f + fsqrt(1.f + 4.f * tk * tk));
      const float beta = (tk - 1.f) * rcp(tn);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float xk = active ? xp[k] : 0.f;
        Y[k] = active ? xk + beta * (xk - X[k]) : 0.f;
        X[k] = xk;
        // the next y's gradient; o holds A (A x+ + b)

      }
      if (has_l) {
        yl = xpl + beta * (xpl - xl);
        xl = xpl;

      }
      eta = eta <= 1e-6f ? 1.f : eta * 2.f;
      tk = tn;
      iters += 1;
    }
  }

  // ---- qdd = M^-1 (qf_smooth + J^T x) (integrator.py:58-64)
  float qdd = u;
  if (wave_any(has_rows)) {
    const float z = apply_jt(X, xl);
    if (has_rows) qdd = u + z;
  }
  __syncthreads();  // the per-dof vectors are read; the integrator reuses them
  if (probe) {
    if (d >= 0) probe[NV + d] = qdd;
    if (s.g == 0) {
      probe[2 * NV] = n_active;
      probe[2 * NV + 1] = (float)iters;
      probe[2 * NV + 2] = (float)trials;
    }
  }

  // ---- semi-implicit Euler (integrator.py:26-84)
  const float dt = M[H_DT];
  if (d >= 0) s.qd = s.qd + qdd * dt;
  if (T::LTYPE[0] == 0) {
    // every lane integrates the root's coordinates from its dofs' velocities
    if (d >= 0) lds[MT::ZV + d] = s.qd;
    __syncthreads();
    const Q4 rot = integrate_rot(Q4{s.rq[3], s.rq[4], s.rq[5], s.rq[6]},
                                 v3(lds[MT::ZV + 3], lds[MT::ZV + 4], lds[MT::ZV + 5]), dt);
    s.rq[3] = rot.w, s.rq[4] = rot.x, s.rq[5] = rot.y, s.rq[6] = rot.z;
#pragma unroll
    for (int k = 0; k < 3; ++k) s.rq[k] = s.rq[k] + lds[MT::ZV + k] * dt;
    __syncthreads();
  }
  if (d >= 0 && T::LTYPE[dlc] != 0) s.qh = s.qh + s.qd * dt;
}

// the state of an env's group into q (NQ) and qd (NV) in LDS
template <class T>
BGX_D void mtree_gather(const MLane<T>& s, float* q, float* qd) {
  if (s.d >= 0) {
    qd[s.d] = s.qd;
    const int l = s.dl;
    if (T::LTYPE[l] != 0) q[T::QADR[l] + s.dk] = s.qh;
  }
  if (T::LTYPE[0] == 0 && s.g == 0)
#pragma unroll
    for (int k = 0; k < 7; ++k) q[k] = s.rq[k];
}

// The multi-lane rollout: env blockIdx.x * E + e of checkpoint segment
// blockIdx.y on group e of the wavefront, n_steps x n_frames substeps
// (the outputs of tree_item: trajectories, substep states, probe, status;
// the frames of the final state are the frames kernel's, launched after).
// An env past the batch computes on the last env and writes nothing.
template <class T>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(3))) void k_mtree(Launch a) {
  using MT = MTab<T>;
  constexpr int G = MT::G, E = MT::E, NQ = T::NQ, NV = T::NV, NU = T::NU;
  __shared__ float smem[E * MT::SIZE];
  const int wl = threadIdx.x, g = wl % G, gb = wl - g, e = wl / G;
  const long long env = (long long)blockIdx.x * E + e;
  const bool live = env < a.n;
  const long long ec = live ? env : a.n - 1;
  const long long seg = blockIdx.y;
  float* lds = smem + e * MT::SIZE;
  const float* M = a.M;
  MLane<T> s = mlane_roles<T>(g, gb);
  const float* qp = a.q + (seg * a.n + ec) * NQ;
  const float* qdp = a.qd + (seg * a.n + ec) * NV;
#pragma unroll
  for (int k = 0; k < 7; ++k) s.rq[k] = T::LTYPE[0] == 0 ? qp[k] : 0.f;
  s.qh = (s.d >= 0 && T::LTYPE[s.dl] != 0) ? qp[T::QADR[s.dl] + s.dk] : 0.f;
  s.qd = s.d >= 0 ? qdp[s.d] : 0.f;
  const float* actp = a.act ? a.act + seg * a.seg_act_stride + ec * NU : nullptr;
  const long long n_sub = (long long)a.n_steps * a.n_frames;
  float* probe = (live && a.probe) ? a.probe + env * (2 * NV + 3) : nullptr;
  for (int st = 0; st < a.n_steps; ++st) {
    const float* act = actp ? actp + st * a.act_stride : nullptr;
    for (int f = 0; f < a.n_frames; ++f) {
      if (live && a.sub_q) {
        const long long o = (seg * n_sub + (long long)st * a.n_frames + f) * a.n + env;
        mtree_gather<T>(s, a.sub_q + o * NQ, a.sub_qd + o * NV);
      }
      const bool last = st == a.n_steps - 1 && f == a.n_frames - 1;
      mtree_substep<T>(M, s, act, lds, last ? probe : nullptr);
    }
    if (live && a.q_traj) mtree_gather<T>(s, a.q_traj + st * a.q_step + env * NQ, lds + MT::QS);
    if (live && a.qd_traj) mtree_gather<T>(s, lds + MT::QS, a.qd_traj + st * a.qd_step + env * NV);
  }
  // final state; non-finite status over the group
  float* qs = lds + MT::QS;
  __syncthreads();
  mtree_gather<T>(s, qs, qs + NQ);
  __syncthreads();
  bool bad = false;
  if (s.g == 0) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) bad = bad || !isfinite(qs[k]);
#pragma unroll
    for (int k = 0; k < NV; ++k) bad = bad || !isfinite(qs[NQ + k]);
  }
  if (live) {
    if (s.g < NQ) a.q[(seg * a.n + env) * NQ + s.g] = qs[s.g];
    if (s.g < NV) a.qd[(seg * a.n + env) * NV + s.g] = qs[NQ + s.g];
    if (NQ > G) {
      for (int k = s.g + G; k < NQ; k += G) a.q[(seg * a.n + env) * NQ + k] = qs[k];
    }
    if (s.g == 0) {
      const int stt = bad ? ST_NONFINITE : 0;
      if (a.status) {
        if (a.n_seg > 1) atomicOr(a.status + env, stt);
        else a.status[env] = a.status_or ? (a.status[env] | stt) : stt;
      }
    }
  }
}

}  // namespace bgx
