// Serial hinge chains, several lanes per environment.
//
// The heterogeneous workload's chains (hinge links in series from the world,
// up to 32 dofs, no contacts, no limits, one motor per hinge) are long and
// thin: one thread per env keeps a 32x32 mass matrix and its factor in
// registers and spills.  Here an env is a group of G lanes (G = the next
// power of two >= its link count, 64/G envs per wavefront) and lane l is
// link l and dof l:
//  * the tree passes of the step are scans over the group's lanes: world
//    frames are an inclusive prefix composition of the joint transforms
//    (kinematics.py:89-104), link velocities and RNEA accelerations prefix
//    sums (dynamics.py:90-161), subtree forces and composite inertias suffix
//    sums (dynamics.py:174-182, mass.py:36-52), the root COM a butterfly sum
//    (dynamics.py:37-43); Hillis-Steele scans of log2(G) shuffle steps;
//  * the mass matrix (mass.py:54-83; a chain's is dense) is built row per
//    lane from the cdofs broadcast through LDS, and factored M = L L^T by a
//    right-looking Cholesky whose columns go through LDS; the solves are
//    forward and back substitutions, one pivot per step;
//  * everything else (cinr, cdof, passive forces, actuators, integration) is
//    per lane.
// The step is the reference's (pipeline.py:64-94) restricted to this family;
// an env's results do not depend on G (lanes past the last link are exact
// identities in every scan) or on the envs sharing its wavefront.
#pragma once

#include <hip/hip_runtime.h>

#include "bgx_step.h"

namespace bgx {

constexpr int CHAIN_MAX = 32;        // links (dofs) of the longest chain
constexpr int CH_STRIDE = WAVE + 1;  // LDS stride of the factor's columns (bank spread)
// LDS of one wavefront: the cdof broadcast [6][WAVE], then the factor's
// columns [CHAIN_MAX][CH_STRIDE]
constexpr int CHAIN_LDS_CDOF = 0;
constexpr int CHAIN_LDS_COL = 6 * WAVE;
constexpr int CHAIN_LDS_FLOATS = CHAIN_LDS_COL + CHAIN_MAX * CH_STRIDE;

// float offsets of a chain's model table (codegen.pack_model of a chain
// topology with L links: header, L link records, L dof records, L actuator
// records, one contact-candidate record of padding)
constexpr int chain_off_link() { return H_ZERO + 1; }
constexpr int chain_off_dof(int L) { return chain_off_link() + LINK_STRIDE * L; }
constexpr int chain_off_act(int L) { return chain_off_dof(L) + DOF_STRIDE * L; }
constexpr int chain_table_size(int L) { return chain_off_act(L) + ACT_STRIDE * L + CAND_STRIDE; }

constexpr int chain_lanes(int L) {
  return L <= 2 ? 2 : L <= 4 ? 4 : L <= 8 ? 8 : L <= 16 ? 16 : 32;
}

// one link's constants (a lane's share of the model table), read once per
// launch; lanes past the last link hold the identity link: no mass, no
// motion, unit armature, no actuator
struct ChainLink {
  V3 tpos, jpos, ipos, ang;
  Q4 trot, irot;
  Sym3 inertia;
  float mass, armature, damping, stiffness;
  float ctrl_lo, ctrl_hi, force_lo, force_hi, gain, gear, bias_q, bias_qd;
};

BGX_D ChainLink chain_link(const float* __restrict__ M, int L, int l) {
  ChainLink c;
  if (l < L) {
    const float* lk = M + chain_off_link() + LINK_STRIDE * l;
    const float* df = M + chain_off_dof(L) + DOF_STRIDE * l;
    const float* ac = M + chain_off_act(L) + ACT_STRIDE * l;
    c.tpos = ldv3(lk + LK_TPOS), c.trot = ldq4(lk + LK_TROT), c.jpos = ldv3(lk + LK_JPOS);
    c.ipos = ldv3(lk + LK_IPOS), c.irot = ldq4(lk + LK_IROT);
    c.inertia = ldsym(lk + LK_I), c.mass = lk[LK_MASS];
    c.ang = ldv3(df + DF_ANG), c.armature = df[DF_ARMATURE];
    c.damping = df[DF_DAMPING], c.stiffness = df[DF_STIFFNESS];
    c.ctrl_lo = ac[AC_CTRL_LO], c.ctrl_hi = ac[AC_CTRL_HI];
    c.force_lo = ac[AC_FORCE_LO], c.force_hi = ac[AC_FORCE_HI];
    c.gain = ac[AC_GAIN], c.gear = ac[AC_GEAR];
    c.bias_q = ac[AC_BIAS_Q], c.bias_qd = ac[AC_BIAS_QD];
  } else {
    c.tpos = c.jpos = c.ipos = c.ang = v3zero();
    c.trot = c.irot = qident();
    c.inertia = Sym3{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    c.mass = 0.f, c.armature = 1.f, c.damping = 0.f, c.stiffness = 0.f;
    c.ctrl_lo = c.ctrl_hi = c.force_lo = c.force_hi = 0.f;
    c.gain = c.gear = c.bias_q = c.bias_qd = 0.f;
  }
  return c;
}

// ---------------------------------------------------------------------------
// group shuffles (width G: lanes [gb, gb + G) of the wavefront)

template <int G>
BGX_D float sh_up(float x, int d) {
  return __shfl_up(x, d, G);
}
template <int G>
BGX_D float sh_down(float x, int d) {
  return __shfl_down(x, d, G);
}
template <int G>
BGX_D float sh_xor(float x, int m) {
  return __shfl_xor(x, m, G);
}
template <int G>
BGX_D float sh_at(float x, int k) {
  return __shfl(x, k, G);
}
template <int G>
BGX_D V3 up3(V3 a, int d) {
  return V3{sh_up<G>(a.x, d), sh_up<G>(a.y, d), sh_up<G>(a.z, d)};
}
template <int G>
BGX_D Q4 up4(Q4 a, int d) {
  return Q4{sh_up<G>(a.w, d), sh_up<G>(a.x, d), sh_up<G>(a.y, d), sh_up<G>(a.z, d)};
}
template <int G>
BGX_D V3 down3(V3 a, int d) {
  return V3{sh_down<G>(a.x, d), sh_down<G>(a.y, d), sh_down<G>(a.z, d)};
}

// inclusive prefix sum of per-lane 3-vector pairs over the group
template <int G>
BGX_D void prefix_sum2(int l, V3& a, V3& b) {
#pragma unroll
  for (int d = 1; d < G; d *= 2) {
    const V3 pa = up3<G>(a, d), pb = up3<G>(b, d);
    if (l >= d) a = pa + a, b = pb + b;
  }
}

// inclusive suffix sum of per-lane 3-vector pairs over the group
template <int G>
BGX_D void suffix_sum2(int l, V3& a, V3& b) {
#pragma unroll
  for (int d = 1; d < G; d *= 2) {
    const V3 na = down3<G>(a, d), nb = down3<G>(b, d);
    if (l + d < G) a = a + na, b = b + nb;
  }
}

// world frames of the links from their joint transforms: x_l = x_(l-1) o j_l
// as an inclusive prefix composition (kinematics.py:89-104); the rotation is
// returned unnormalized (the caller normalizes, kinematics.py:106)
template <int G>
BGX_D void chain_frames(int l, const ChainLink& c, float q, V3& xp, Q4& xr) {
  // jcalc of a hinge and the joint anchor / link transform (kinematics.py:52-87)
  const Q4 rj = qnormalize(qrot_axis(c.ang, q));
  V3 p = c.jpos - rotate(c.jpos, rj);
  p = c.tpos + rotate(p, c.trot);
  xp = p;
  xr = qmul(c.trot, rj);
#pragma unroll
  for (int d = 1; d < G; d *= 2) {
    const V3 pp = up3<G>(xp, d);
    const Q4 pr = up4<G>(xr, d);
    if (l >= d) {
      xp = pp + rotate(xp, pr);
      xr = qmul(pr, xr);
    }
  }
}

// ---------------------------------------------------------------------------
// one substep of one chain on a group of G lanes

// lane l of the group (wl: lane in the wavefront, gb: the group's first
// lane); lds: the wavefront's CHAIN_LDS_FLOATS.  q, qd: this lane's dof;
// act: its motor's control.  probe (non-null): qf_smooth, qdd of this dof.
template <int G>
BGX_D void chain_substep(const ChainLink& c, V3 grav, float dt, int l, int wl, int gb,
                         float& q, float& qd, float act, float* lds, float* probe_qf,
                         float* probe_qdd) {
  // ---- kinematics.forward
  V3 xp;
  Q4 xr;
  chain_frames<G>(l, c, q, xp, xr);
  xr = qnormalize(xr);

  // ---- root COM (dynamics.py:37-43): butterfly sums over the group
  const V3 xip = xp + rotate(c.ipos, xr);
  V3 sm = xip * c.mass;
  float ms = c.mass;
#pragma unroll
  for (int d = 1; d < G; d *= 2) {
    sm = sm + V3{sh_xor<G>(sm.x, d), sh_xor<G>(sm.y, d), sh_xor<G>(sm.z, d)};
    ms = ms + sh_xor<G>(ms, d);
  }
  const float ims = rcp(ms);
  const V3 rc{sm.x * ims, sm.y * ims, sm.z * ims};

  // ---- cinr (dynamics.py:44-55): inertia about the root COM
  V3 cc;
  Sym3 ci;
  {
    const Q4 xir = qmul(xr, c.irot);
    const V3 p = xip - rc;
    float R[3][3];
    quat_to_3x3(xir, R);
    float RI[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const V3 cr = symmul(c.inertia, V3{R[i][0], R[i][1], R[i][2]});
      RI[i][0] = cr.x, RI[i][1] = cr.y, RI[i][2] = cr.z;
    }
    auto rir = [&](int i, int j) {
      return RI[i][0] * R[j][0] + RI[i][1] * R[j][1] + RI[i][2] * R[j][2];
    };
    const float m = c.mass, pp = dot(p, p);
    ci = Sym3{rir(0, 0) + (pp - p.x * p.x) * m, rir(1, 1) + (pp - p.y * p.y) * m,
              rir(2, 2) + (pp - p.z * p.z) * m, rir(0, 1) - p.x * p.y * m,
              rir(0, 2) - p.x * p.z * m, rir(1, 2) - p.y * p.z * m};
    cc = p * m;
  }

  // ---- cdof (dynamics.py:56-89): the hinge axis in the parent's frame
  // composed with the link transform, about the root COM
  V3 ca, cv;
  {
    V3 pp = up3<G>(xp, 1);
    Q4 pr = up4<G>(xr, 1);
    if (l == 0) pp = v3zero(), pr = qident();
    const V3 t1p = pp + rotate(c.tpos, pr);
    const Q4 jr = qmul(pr, c.trot);
    const V3 jp = t1p + rotate(c.jpos, jr);
    ca = rotate(c.ang, jr);
    cv = v3zero() - cross(rc - jp, ca);
  }

  // ---- cd: link velocities at qd (dynamics.py:90-103)
  V3 cda = ca * qd, cdv = cv * qd;
  prefix_sum2<G>(l, cda, cdv);

  // ---- RNEA bias (dynamics.py:105-184): accelerations from -g at the root,
  // cdofd = cd(parent) x cdof
  V3 aa, av;
  {
    V3 pa = up3<G>(cda, 1), pv = up3<G>(cdv, 1);
    if (l == 0) pa = pv = v3zero();
    V3 oa, ov;
    cross_motion(pa, pv, ca, cv, oa, ov);
    aa = oa * qd;
    av = l == 0 ? (v3zero() - grav) + ov * qd : ov * qd;
    prefix_sum2<G>(l, aa, av);
  }
  V3 fa, fv;
  {
    V3 f1a, f1v, f2a, f2v, f3a, f3v;
    inertia_mul(cc, ci, c.mass, aa, av, f1a, f1v);
    inertia_mul(cc, ci, c.mass, cda, cdv, f2a, f2v);
    cross_force(cda, cdv, f2a, f2v, f3a, f3v);
    fa = f1a + f3a, fv = f1v + f3v;
  }
  suffix_sum2<G>(l, fa, fv);
  const float bias = dot(cv, fv) + dot(ca, fa);

  // ---- qf_smooth = passive - bias + tau (dynamics.py:187-236, actuator.py:23-57)
  float qf = (-q * c.stiffness + (-c.damping * qd)) - bias;
  {
    const float ctl = clipf(act, c.ctrl_lo, c.ctrl_hi);
    const float b = c.gear * (q * c.bias_q + qd * c.bias_qd);
    float f = c.gain * ctl + b;
    if (f < c.force_lo) f = c.force_lo;
    if (f > c.force_hi) f = c.force_hi;
    qf += f * c.gear;
  }
  if (probe_qf) *probe_qf = qf;

  // ---- mass.matrix (mass.py:27-83): composite inertias are suffix sums;
  // lane i forms row i, M_ij = cdof_j . (I_c(i) cdof_i), j <= i
  float km = c.mass;
  {
    V3 ci_d{ci.xx, ci.yy, ci.zz}, ci_o{ci.xy, ci.xz, ci.yz};
    suffix_sum2<G>(l, cc, ci_d);
#pragma unroll
    for (int d = 1; d < G; d *= 2) {
      const V3 n = down3<G>(ci_o, d);
      const float nm = sh_down<G>(km, d);
      if (l + d < G) ci_o = ci_o + n, km = km + nm;
    }
    ci = Sym3{ci_d.x, ci_d.y, ci_d.z, ci_o.x, ci_o.y, ci_o.z};
  }
  V3 Fa, Fv;
  inertia_mul(cc, ci, km, ca, cv, Fa, Fv);
  const float diag = dot(cv, Fv) + dot(ca, Fa) + c.armature;
  float* cdl = lds + CHAIN_LDS_CDOF;
  cdl[0 * WAVE + wl] = ca.x, cdl[1 * WAVE + wl] = ca.y, cdl[2 * WAVE + wl] = ca.z;
  cdl[3 * WAVE + wl] = cv.x, cdl[4 * WAVE + wl] = cv.y, cdl[5 * WAVE + wl] = cv.z;
  __syncthreads();
  float mrow[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    const int s = gb + j;
    const V3 aj{cdl[0 * WAVE + s], cdl[1 * WAVE + s], cdl[2 * WAVE + s]};
    const V3 vj{cdl[3 * WAVE + s], cdl[4 * WAVE + s], cdl[5 * WAVE + s]};
    const float off = dot(vj, Fv) + dot(aj, Fa);
    mrow[j] = j < l ? off : (j == l ? diag : 0.f);
  }

  // ---- M = L L^T: right-looking Cholesky, lane i keeps row i of L; column
  // k of L goes to LDS for the rows below it and the back substitution
  float* col = lds + CHAIN_LDS_COL;
  float inv[G];
#pragma unroll
  for (int k = 0; k < G; ++k) {
    const float lkk = fsqrt(sh_at<G>(mrow[k], k));
    inv[k] = rcp(lkk);
    if (l == k) mrow[k] = lkk;
    if (l > k) mrow[k] *= inv[k];
    col[k * CH_STRIDE + wl] = mrow[k];
    __syncthreads();
#pragma unroll
    for (int j = k + 1; j < G; ++j)
      if (j <= l) mrow[j] -= mrow[k] * col[k * CH_STRIDE + gb + j];
  }

  // ---- qdd = M^-1 qf_smooth (no constraint rows in this family):
  // L y = qf, then L^T qdd = y
  float y = qf;
#pragma unroll
  for (int k = 0; k < G; ++k) {
    const float yk = sh_at<G>(y, k) * inv[k];
    if (l == k) y = yk;
    if (l > k) y -= mrow[k] * yk;
  }
#pragma unroll
  for (int k = G - 1; k >= 0; --k) {
    const float xk = sh_at<G>(y, k) * inv[k];
    if (l == k) y = xk;
    if (l < k) y -= col[l * CH_STRIDE + gb + k] * xk;
  }
  __syncthreads();  // the LDS is rewritten by the next substep
  if (probe_qdd) *probe_qdd = y;

  // ---- semi-implicit Euler (integrator.py:26-84)
  qd = qd + y * dt;
  q = q + qd * dt;
}

// world frames and velocities of the links (pipeline.init, kinematics.py:
// 31-108 with xd): xa_l = sum over the path of rotate(ang qd, x_rot),
// xv_l = xv_(l-1) + xa_(l-1) x (xp_l - xp_(l-1))
template <int G>
BGX_D void chain_frames_xd(int l, const ChainLink& c, float q, float qd, V3& xp, Q4& xr,
                           V3& xa, V3& xv) {
  chain_frames<G>(l, c, q, xp, xr);
  V3 w = rotate(c.ang * qd, xr), z = v3zero();
  prefix_sum2<G>(l, w, z);
  xa = w;
  V3 pa = up3<G>(xa, 1), pp = up3<G>(xp, 1);
  V3 t = l == 0 ? v3zero() : cross(pa, xp - pp), z2 = v3zero();
  prefix_sum2<G>(l, t, z2);
  xv = t;
  xr = qnormalize(xr);
}

}  // namespace bgx
