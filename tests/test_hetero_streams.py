"""Host logic of the heterogeneous batch: dealing groups to streams."""
import itertools

from brax_amd import systems as S
from brax_amd.generalized import hetero


def _makespan(costs, out, k):
  return max(sum(c for c, j in zip(costs, out) if j == s) for s in range(k))


def test_longest_first_spreads_the_long_groups():
  # per-wave costs of chains 2..32 and a quadruped, as the kernel traces rank them
  costs = [float(n ** 3) * (4.0 if n > 24 else 1.0) for n in range(2, 33)] + [4.0 * 14 ** 3]
  out = hetero.assign_streams(costs, 4)
  assert sorted(set(out)) == [0, 1, 2, 3]
  # the four longest groups start on four different streams
  longest = sorted(range(len(costs)), key=lambda g: -costs[g])[:4]
  assert len({out[g] for g in longest}) == 4
  # within 4/3 of a lower bound on the best makespan (LPT is within 4/3 of the best)
  bound = max(max(costs), sum(costs) / 4)
  assert _makespan(costs, out, 4) <= 4 / 3 * bound


def test_small_cases_are_optimal_or_close():
  for costs in ([5.0], [3.0, 3.0, 2.0, 2.0, 2.0], [7.0, 1.0, 1.0, 1.0]):
    for k in (1, 2, 3):
      out = hetero.assign_streams(costs, k)
      best = min(_makespan(costs, list(p), k)
                 for p in itertools.product(range(k), repeat=len(costs)))
      assert _makespan(costs, out, k) <= 4 / 3 * best + 1e-9


def test_group_cost_ranks_rolled_and_constrained_trees():
  class G:
    def __init__(self, sys):
      self.sys = sys
  c = {n: hetero._group_cost(G(S.chain(n))) for n in (8, 24, 25, 32)}
  assert c[8] < c[24] < c[25] < c[32]
  assert c[25] / c[24] > 3  # rolled loops past the unroll limit
  assert hetero._group_cost(G(S.quadruped(0))) == 4.0 * 14 ** 3
