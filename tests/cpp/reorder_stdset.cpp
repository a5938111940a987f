// Farthest-point sphere visiting order computed with libstdc++'s own
// std::unordered_set<size_t>, driven through the same container operations as
// reorder_spheres_for_informative_collision_check
// (cpp/plainmp/constraints/primitive_sphere_collision.cpp:53-96): reserve(k),
// insert 0..k-1 in increasing order, seed from *begin(), erase the seed, then
// per step scan the set in iteration order keeping the first strict maximum of
// the distance to the visited set and erase it by iterator.
//
// Test infrastructure: pins the oracle's vector model of that iteration order.
// Input on stdin: a count of groups, then per group k followed by k lines of
// x y z.  Output: one line per group with the visiting order.
#include <cmath>
#include <cstdio>
#include <limits>
#include <unordered_set>
#include <vector>

int main() {
  int groups = 0;
  if (std::scanf("%d", &groups) != 1) return 1;
  for (int g = 0; g < groups; g++) {
    size_t k = 0;
    if (std::scanf("%zu", &k) != 1) return 1;
    std::vector<double> x(k), y(k), z(k);
    for (size_t i = 0; i < k; i++)
      if (std::scanf("%lf %lf %lf", &x[i], &y[i], &z[i]) != 3) return 1;
    std::unordered_set<size_t> open;
    open.reserve(k);
    for (size_t i = 0; i < k; i++) open.insert(i);
    std::vector<size_t> order;
    size_t seed = *open.begin();
    order.push_back(seed);
    open.erase(seed);
    while (!open.empty()) {
      double far = 0.0;
      auto pick = open.begin();
      for (auto it = open.begin(); it != open.end(); ++it) {
        double near = std::numeric_limits<double>::max();
        for (size_t v : order) {
          double dx = x[*it] - x[v], dy = y[*it] - y[v], dz = z[*it] - z[v];
          double d = std::sqrt(dx * dx + dy * dy + dz * dz);
          if (d < near) near = d;
        }
        if (near > far) {
          far = near;
          pick = it;
        }
      }
      order.push_back(*pick);
      open.erase(pick);
    }
    for (size_t i = 0; i < order.size(); i++) std::printf(i ? " %zu" : "%zu", order[i]);
    std::printf("\n");
  }
  return 0;
}
