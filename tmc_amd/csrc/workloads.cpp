// workloads.cpp -- synthetic scalar fields for tests and benchmarks
// (libtmc_workloads.so, host only).
//
// Fields are evaluated in double and rounded once to float32, so every
// consumer (engine, oracle, reference build) sees the same buffer.  Random
// numbers come from a counter-based splitmix64 (value k of stream `seed` is
// mix(seed + (k+1)*golden)), uniform in (0,1) from the top 53 bits, normals
// by Box-Muller.  Work is split over z-planes across threads; results do not
// depend on the thread count.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform(uint64_t seed, uint64_t k) {
    return ((double)(mix(seed + (k + 1) * kGolden) >> 11) + 0.5) * 0x1.0p-53;
}
inline double normal(uint64_t seed, uint64_t idx) {
    const double u1 = uniform(seed, 2 * idx), u2 = uniform(seed, 2 * idx + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

template <class F>
void parallel_planes(int nz, int threads, F fn) {
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (threads <= 0) threads = 1;
    if (threads > 64) threads = 64;
    if (threads > nz) threads = nz;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([=] {
            for (int k = t; k < nz; k += threads) fn(k);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

// Splitmix64 uniform / normal streams (exposed for tests).
double wl_uniform(uint64_t seed, uint64_t k) { return uniform(seed, k); }
double wl_normal(uint64_t seed, uint64_t idx) { return normal(seed, idx); }

// Genus-2 implicit f = 2y(y^2-3x^2)(1-z^2) + (x^2+y^2)^2 - (9z^2-1)(1-z^2)
// sampled at cell centres of [lo,hi]^3: x_i = lo + (i+1/2)(hi-lo)/n.
void wl_genus2(int nx, int ny, int nz, double lo, double hi, float* out, int threads) {
    const double hx = (hi - lo) / nx, hy = (hi - lo) / ny, hz = (hi - lo) / nz;
    parallel_planes(nz, threads, [=](int k) {
        const double z = lo + (k + 0.5) * hz;
        for (int j = 0; j < ny; j++) {
            const double y = lo + (j + 0.5) * hy;
            float* row = out + ((int64_t)k * ny + j) * nx;
            for (int i = 0; i < nx; i++) {
                const double x = lo + (i + 0.5) * hx;
                const double f = 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) * (x * x + y * y) -
                                 (9 * z * z - 1) * (1 - z * z);
                row[i] = (float)f;
            }
        }
    });
}

// Triply periodic Schwarz-P cos x + cos y + cos z on [lo,hi]^3, node-aligned
// (x_i = lo + i(hi-lo)/(n-1)), plus N(0, sigma^2) noise per point.
void wl_schwarz(int nx, int ny, int nz, double lo, double hi, double sigma, uint64_t seed, float* out,
                int threads) {
    std::vector<double> cx(nx), cy(ny), cz(nz);
    for (int i = 0; i < nx; i++) cx[i] = std::cos(lo + i * (hi - lo) / (nx - 1));
    for (int j = 0; j < ny; j++) cy[j] = std::cos(lo + j * (hi - lo) / (ny - 1));
    for (int k = 0; k < nz; k++) cz[k] = std::cos(lo + k * (hi - lo) / (nz - 1));
    const double* px = cx.data();
    const double* py = cy.data();
    const double* pz = cz.data();
    parallel_planes(nz, threads, [=](int k) {
        for (int j = 0; j < ny; j++) {
            const int64_t base = ((int64_t)k * ny + j) * nx;
            for (int i = 0; i < nx; i++) {
                double f = (px[i] + py[j]) + pz[k];
                if (sigma != 0.0) f += sigma * normal(seed, (uint64_t)(base + i));
                out[base + i] = (float)f;
            }
        }
    });
}

// Sum of nblobs Gaussians a*exp(-|p-c|^2/(2 s^2)) on [0,1]^3 at cell centres
// (i+1/2)/n, c ~ U[0,1]^3, s ~ U[smin,smax], a ~ U[amin,amax] (parameters
// drawn in order c.x c.y c.z s a per blob from stream seed), evaluated as
// the product of per-axis factors, plus N(0, sigma^2) noise from stream
// seed+1.
void wl_blobs(int nx, int ny, int nz, int nblobs, double smin, double smax, double amin, double amax, double sigma,
              uint64_t seed, float* out, int threads) {
    std::vector<double> ex((size_t)nblobs * nx), ey((size_t)nblobs * ny), ez((size_t)nblobs * nz), amp(nblobs);
    for (int b = 0; b < nblobs; b++) {
        const double c0 = uniform(seed, 5 * b), c1 = uniform(seed, 5 * b + 1), c2 = uniform(seed, 5 * b + 2);
        const double s = smin + (smax - smin) * uniform(seed, 5 * b + 3);
        amp[b] = amin + (amax - amin) * uniform(seed, 5 * b + 4);
        const double inv = 1.0 / (2 * s * s);
        for (int i = 0; i < nx; i++) {
            const double t = (i + 0.5) / nx - c0;
            ex[(size_t)b * nx + i] = std::exp(-t * t * inv);
        }
        for (int j = 0; j < ny; j++) {
            const double t = (j + 0.5) / ny - c1;
            ey[(size_t)b * ny + j] = std::exp(-t * t * inv);
        }
        for (int k = 0; k < nz; k++) {
            const double t = (k + 0.5) / nz - c2;
            ez[(size_t)b * nz + k] = std::exp(-t * t * inv);
        }
    }
    const double* pex = ex.data();
    const double* pey = ey.data();
    const double* pez = ez.data();
    const double* pa = amp.data();
    parallel_planes(nz, threads, [=](int k) {
        std::vector<double> w(nblobs);
        for (int j = 0; j < ny; j++) {
            for (int b = 0; b < nblobs; b++) w[b] = pa[b] * pey[(size_t)b * ny + j] * pez[(size_t)b * nz + k];
            const int64_t base = ((int64_t)k * ny + j) * nx;
            for (int i = 0; i < nx; i++) {
                double f = 0.0;
                for (int b = 0; b < nblobs; b++) f += w[b] * pex[(size_t)b * nx + i];
                if (sigma != 0.0) f += sigma * normal(seed + 1, (uint64_t)(base + i));
                out[base + i] = (float)f;
            }
        }
    });
}

}  // extern "C"
