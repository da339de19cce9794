// Host build of the adjoint's reverse sweep (stage A2) and tangent pass
// (stage B) for memory-safety checks on the CPU.  Compiled by
// tests/test_adjoint_tape_clamps.py with g++ -fsanitize=address,undefined
// and -DBGX_TOPO=<topology struct>; the same device code as the library,
// one lane per env (P = 1), with every array the kernels address on the
// heap at its exact size so that an out-of-bounds access aborts.
//
//   adj_tape_check <inputs.bin>
// inputs.bin: float32 model table (T::SIZE), q (NQ), qd (NV), act (NU).
//
// It tapes one substep (stage A1, substep_ks with a slot per candidate),
// runs A2 and B (each q coordinate in dual numbers, each qd coordinate in
// dual numbers and all of qd by the reverse pass) on that tape, then on
// copies of it whose integer fields --
// active-contact count, solver iterations, each slot's link, each
// candidate's slot and frame flag -- hold NaN, infinities, huge, negative
// and just-out-of-range values.  Prints one line per case; exits 0 when
// every case ran (the sanitizers abort on a bad access).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
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

// A2 then B (every input coordinate) on one tape; returns how many outputs
// are finite
static int run_reverse(const float* M, const std::vector<float>& q,
                       const std::vector<float>& qd, const std::vector<float>& act,
                       std::vector<float>& rec) {
  std::vector<float> shared(RevLds<T>::SHARED), priv(RevLds<T>::PRIVATE);
  std::vector<float> gq1(T::NQ, 1.f), gqd1(T::NV, 1.f), gact(T::NU > 0 ? T::NU : 1, 0.f);
  adjoint_rev<T, 1, 1>(M, q.data(), qd.data(), act.data(), gq1.data(), gqd1.data(),
                       RecPtr{rec.data(), 1}, shared.data(), priv.data(), gact.data(), 0);
  int finite = 0;
  for (int d = 0; d < T::NQ + T::NV; ++d)
    finite += std::isfinite(adjoint_b<T>(M, q.data(), qd.data(), act.data(),
                                         RecPtr{rec.data(), 1}, d)) ? 1 : 0;
  for (int k = 0; k < T::NU; ++k) finite += std::isfinite(gact[k]) ? 1 : 0;
  // the qd gradient by one reverse pass (what the kernel runs for qd)
  std::vector<float> gqd(T::NV);
  adjoint_b_qd_all<T>(M, q.data(), qd.data(), act.data(), RecPtr{rec.data(), 1}, gqd.data());
  for (int j = 0; j < T::NV; ++j) finite += std::isfinite(gqd[j]) ? 1 : 0;
  return finite;
}

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::vector<float> in = read_all(argv[1]);
  if ((int)in.size() != T::SIZE + T::NQ + T::NV + T::NU) {
    std::printf("input size %zu, expected %d\n", in.size(), T::SIZE + T::NQ + T::NV + T::NU);
    return 2;
  }
  const std::vector<float> model(in.begin(), in.begin() + T::SIZE);
  const float* M = model.data();
  std::vector<float> q(in.begin() + T::SIZE, in.begin() + T::SIZE + T::NQ);
  std::vector<float> qd(in.begin() + T::SIZE + T::NQ, in.begin() + T::SIZE + T::NQ + T::NV);
  std::vector<float> act(in.begin() + T::SIZE + T::NQ + T::NV, in.end());
  if (act.empty()) act.push_back(0.f);

  // stage A1: one substep with a slot per candidate, recording its tape
  constexpr int KS = TapeLayout<T>::KS;
  std::vector<float> lds(lds_floats<T>() * WAVE);
  std::vector<float> rec(R::SIZE, 0.f);
  {
    std::vector<float> q1 = q, qd1 = qd;
    PrivRows<KS> rows;
    substep_ks<T, KS>(M, q1.data(), qd1.data(), act.data(), lds.data(), rows, nullptr,
                      RecPtr{rec.data(), 1});
  }
  const int n_c = (int)rec[R::NC], iters = (int)rec[R::ITERS];
  std::vector<float> ok = rec;
  const int f0 = run_reverse(M, q, qd, act, ok);
  constexpr int n_out = T::NQ + 2 * T::NV + T::NU;
  std::printf("valid tape: %d active contacts, %d solver iterations, %d of %d outputs finite\n",
              n_c, iters, f0, n_out);
  if (f0 != n_out) return 3;

  const float inf = std::numeric_limits<float>::infinity();
  const float bad[] = {std::numeric_limits<float>::quiet_NaN(), inf, -inf, 1e30f, -1e30f,
                       3e9f, -3e9f, -1.f, -7.f, (float)(KS + 1), (float)(T::L + 1), 64.f,
                       (float)(ADJ_MAXITER + 1), 1e6f};
  std::vector<int> fields = {R::NC, R::ITERS};
  for (int s = 0; s < KS; ++s) fields.push_back(R::SLINK + s);
  for (int c = 0; c < T::NCAND; ++c) fields.push_back(R::CSLOT + c), fields.push_back(R::CFB + c);
  int cases = 0;
  for (int f : fields)
    for (float v : bad) {
      std::vector<float> r = rec;
      r[f] = v;
      run_reverse(M, q, qd, act, r);
      ++cases;
    }
  // every integer field damaged at once
  for (float v : bad) {
    std::vector<float> r = rec;
    for (int f : fields) r[f] = v;
    run_reverse(M, q, qd, act, r);
    ++cases;
  }
  std::printf("corrupted tapes: %d cases over %zu fields, no out-of-bounds access\n", cases,
              fields.size());
  return 0;
}
