// Record layout of the float32 model table (written by brax_amd/codegen.py).
#pragma once

namespace bgx {

// header
constexpr int H_GRAVITY = 0;  // [3]
constexpr int H_DT = 3;
constexpr int H_MAXITER = 4;  // solver_iterations (stored as float)
constexpr int H_MAXLS = 5;    // solver_maxls
constexpr int H_TOL = 6;      // projected-gradient stopping tolerance
constexpr int H_LSEPS = 7;    // line-search slack
constexpr int H_ZERO = 8;     // always 0 (an offset the compiler cannot fold)

// impedance record: constants of constraint.py:29-65 derived on the host
// from (timeconst, dampratio, dmin, dmax, width, mid, power)
constexpr int IMP_SIZE = 9;
constexpr int IM_DMIN = 0, IM_DMAX = 1, IM_INVWIDTH = 2, IM_MID = 3,
              IM_POWER = 4, IM_ACOEF = 5 /* 1 / mid^(power-1) */,
              IM_BCOEF = 6 /* 1 / (1-mid)^(power-1) */, IM_B = 7, IM_K = 8;

// per link
constexpr int LINK_STRIDE = 24;
constexpr int LK_TPOS = 0, LK_TROT = 3, LK_JPOS = 7, LK_IPOS = 10, LK_IROT = 13,
              LK_I = 17 /* xx yy zz xy xz yz */, LK_MASS = 23;

// per dof
constexpr int DOF_STRIDE = 24;
constexpr int DF_ANG = 0, DF_VEL = 3, DF_ARMATURE = 6, DF_DAMPING = 7,
              DF_STIFFNESS = 8, DF_LO = 9, DF_HI = 10, DF_INVWEIGHT = 11,
              DF_IMP = 12 /* [IMP_SIZE] */;

// per actuator
constexpr int ACT_STRIDE = 8;
constexpr int AC_CTRL_LO = 0, AC_CTRL_HI = 1, AC_FORCE_LO = 2, AC_FORCE_HI = 3,
              AC_GAIN = 4, AC_GEAR = 5, AC_BIAS_Q = 6, AC_BIAS_QD = 7;

// per contact candidate: plane vs sphere, plane vs one capsule end, or one
// sphere / capsule pair.  Geom 2's pose in its link's frame (the world pose
// for a world geom), radius and half length; for a pair, geom 1's alike in
// the space a plane candidate gives its plane
constexpr int CAND_STRIDE = 36;
constexpr int CD_GPOS = 0, CD_GQUAT = 3, CD_RADIUS = 7, CD_HALFLEN = 8,
              CD_SIGN = 9, CD_MU = 10, CD_IMP = 11 /* [IMP_SIZE] */,
              CD_TW = 20, CD_PPOS = 21, CD_PN = 24, CD_T1 = 27, CD_T2 = 30;
constexpr int CD_G1POS = 21, CD_G1QUAT = 24, CD_R1 = 28, CD_HL1 = 29;

constexpr int KIND_SPHERE = 0;   // plane - sphere
constexpr int KIND_CAPSULE = 1;  // plane - one capsule end
constexpr int KIND_SS = 2;       // sphere - sphere
constexpr int KIND_SC = 3;       // sphere (geom 1) - capsule (geom 2)
constexpr int KIND_CC = 4;       // capsule - capsule

}  // namespace bgx
