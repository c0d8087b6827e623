// Runs the drop-in's triangle orientation (tmc::detail::orient_triangles,
// include/tmc_dmc.hpp) on a triangle list read from stdin ("nv nt" then nt
// lines "a b c") and prints the flip count and the oriented list.
#include <cstdio>
#include <vector>

#include "tmc_dmc.hpp"

int main() {
    long long nv = 0, nt = 0;
    if (std::scanf("%lld %lld", &nv, &nt) != 2) return 1;
    std::vector<std::array<int, 3>> t((size_t)nt);
    for (auto& q : t)
        if (std::scanf("%d %d %d", &q[0], &q[1], &q[2]) != 3) return 1;
    const long flips = tmc::detail::orient_triangles(nv, t);
    std::printf("%ld\n", flips);
    for (const auto& q : t) std::printf("%d %d %d\n", q[0], q[1], q[2]);
    return 0;
}
