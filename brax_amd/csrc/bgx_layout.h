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

// per link
constexpr int LINK_STRIDE = 24;
constexpr int LK_TPOS = 0, LK_TROT = 3, LK_JPOS = 7, LK_IPOS = 10, LK_IROT = 13,
              LK_I = 17 /* xx yy zz xy xz yz */, LK_MASS = 23;

// per dof
constexpr int DOF_STRIDE = 20;
constexpr int DF_ANG = 0, DF_VEL = 3, DF_ARMATURE = 6, DF_DAMPING = 7,
              DF_STIFFNESS = 8, DF_LO = 9, DF_HI = 10, DF_INVWEIGHT = 11,
              DF_SOLREF = 12 /* [2] */, DF_SOLIMP = 14 /* [5] */;

// per actuator
constexpr int ACT_STRIDE = 8;
constexpr int AC_CTRL_LO = 0, AC_CTRL_HI = 1, AC_FORCE_LO = 2, AC_FORCE_HI = 3,
              AC_GAIN = 4, AC_GEAR = 5, AC_BIAS_Q = 6, AC_BIAS_QD = 7;

// per contact candidate (plane vs sphere, or plane vs one capsule end)
constexpr int CAND_STRIDE = 32;
constexpr int CD_GPOS = 0, CD_GQUAT = 3, CD_RADIUS = 7, CD_HALFLEN = 8,
              CD_SIGN = 9, CD_MU = 10, CD_SOLREF = 11 /* [2] */,
              CD_SOLIMP = 13 /* [5] */, CD_TW = 18, CD_PPOS = 19, CD_PN = 22,
              CD_T1 = 25, CD_T2 = 28;

constexpr int KIND_SPHERE = 0;
constexpr int KIND_CAPSULE = 1;

}  // namespace bgx
