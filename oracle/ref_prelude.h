// ref_prelude.h -- force-included into every translation unit of the
// reference build (oracle/_ref).  The standard headers are pulled in first,
// then `private` is opened up so that ref_harness.cpp can call the
// reference's internal stages.  Test infrastructure only.
#include <algorithm>
#include <array>
#include <bitset>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <map>
#include <random>
#include <sstream>
#include <stack>
#include <string>
#include <tuple>
#include <utility>
#include <vector>
#define private public
