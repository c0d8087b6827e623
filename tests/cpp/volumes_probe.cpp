// Evaluates dmc::Volumes::surface of include/tmc_dmc.hpp at the points given
// as "surface x y z" lines on stdin (surface: index of Volumes::Surface) and
// prints the values with 17 significant digits.
#include <cstdio>
#include <iostream>

#include "tmc_dmc.hpp"

int main() {
    int s;
    double x, y, z;
    while (std::cin >> s >> x >> y >> z)
        std::printf("%.17g\n", dmc::Volumes::surface(static_cast<dmc::Volumes::Surface>(s), x, y, z));
    return 0;
}
