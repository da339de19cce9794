// Host build of stage B's qd gradient two ways: one reverse pass over all qd
// coordinates (adjoint_b_qd_all, what the kernel runs) and one dual-number
// pass per coordinate (adjoint_b along NQ + j).  Compiled by
// tests/test_adjoint_qd_reverse.py with g++ and -DBGX_TOPO=<topology struct>.
//
//   adj_qd_check <inputs.bin>
// inputs.bin: float32 model table (T::SIZE), q (NQ), qd (NV), act (NU).
//
// It tapes one substep (stage A1), runs the reverse sweep (A2) for its
// cotangents and compares the two on that record; then again with every
// cotangent field (QFB .. GQDT) replaced by pseudo-random values in
// [-1, 1], which the gradient is linear in, so that each term is exercised
// whatever A2 wrote.  Prints per record "<case>: max |reverse - dual| D,
// max |dual| G, active contacts C"; exits 0 when both ran.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../brax_amd/csrc/bgx_adjoint.h"
#include "../../brax_amd/csrc/topologies.inc"

#ifndef BGX_TOPO
#error "compile with -DBGX_TOPO=<topology struct>"
#endif

using namespace bgx;
using T = BGX_TOPO;
using R = AdjRec<T>;

static std::vector<float> read_all(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) std::exit(2);
  std::vector<float> v;
  float x;
  while (std::fread(&x, sizeof x, 1, f) == 1) v.push_back(x);
  std::fclose(f);
  return v;
}

static void compare(const char* name, const float* M, const std::vector<float>& q,
                    const std::vector<float>& qd, const std::vector<float>& act,
                    std::vector<float>& rec) {
  std::vector<float> rev(T::NV);
  adjoint_b_qd_all<T>(M, q.data(), qd.data(), act.data(), RecPtr{rec.data(), 1}, rev.data());
  double diff = 0.0, big = 0.0;
  for (int j = 0; j < T::NV; ++j) {
    const float dual = adjoint_b<T>(M, q.data(), qd.data(), act.data(), RecPtr{rec.data(), 1},
                                    T::NQ + j);
    diff = std::fmax(diff, std::fabs((double)rev[j] - (double)dual));
    big = std::fmax(big, std::fabs((double)dual));
  }
  std::printf("%s: max |reverse - dual| %.6e, max |dual| %.6e, active contacts %d\n", name, diff,
              big, (int)rec[R::NC]);
}

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::vector<float> in = read_all(argv[1]);
  if ((int)in.size() != T::SIZE + T::NQ + T::NV + T::NU) return 2;
  const std::vector<float> model(in.begin(), in.begin() + T::SIZE);
  const float* M = model.data();
  std::vector<float> q(in.begin() + T::SIZE, in.begin() + T::SIZE + T::NQ);
  std::vector<float> qd(in.begin() + T::SIZE + T::NQ, in.begin() + T::SIZE + T::NQ + T::NV);
  std::vector<float> act(in.begin() + T::SIZE + T::NQ + T::NV, in.end());
  if (act.empty()) act.push_back(0.f);

  constexpr int KS = TapeLayout<T>::KS;
  std::vector<float> lds(lds_floats<T>() * WAVE);
  std::vector<float> rec(R::SIZE, 0.f);
  {
    std::vector<float> q1 = q, qd1 = qd;
    PrivRows<KS> rows;
    substep_ks<T, KS>(M, q1.data(), qd1.data(), act.data(), lds.data(), rows, nullptr,
                      RecPtr{rec.data(), 1});
  }
  {
    std::vector<float> shared(RevLds<T>::SHARED), priv(RevLds<T>::PRIVATE);
    std::vector<float> gq1(T::NQ), gqd1(T::NV), gact(T::NU > 0 ? T::NU : 1, 0.f);
    for (int k = 0; k < T::NQ; ++k) gq1[k] = std::sin(1.0f + k);
    for (int k = 0; k < T::NV; ++k) gqd1[k] = std::cos(0.5f + 2.0f * k);
    adjoint_rev<T, 1, 1>(M, q.data(), qd.data(), act.data(), gq1.data(), gqd1.data(),
                         RecPtr{rec.data(), 1}, shared.data(), priv.data(), gact.data(), 0);
  }
  compare("reverse-sweep cotangents", M, q, qd, act, rec);
  unsigned s = 12345u;
  for (int f = R::QFB; f < R::SIZE; ++f) {
    s = s * 1664525u + 1013904223u;
    rec[f] = (float)((s >> 8) * (1.0 / 8388608.0) - 1.0);
  }
  compare("random cotangents", M, q, qd, act, rec);
  return 0;
}
