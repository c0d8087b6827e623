// tmc_dmc.hpp -- the reference's C++ call surfaces on top of the C ABI.
//
// Header-only; link with libtmc_dmc.so.  Two surfaces are provided:
//
//  dmc::UniformGrid, dmc::Volumes, dmc::DualMarchingCubes
//      as cpp_dmc declares them (cpp_dmc/UniformGrid.h:18-354,
//      cpp_dmc/Volumes.h:19-336, cpp_dmc/DualMarchingCubes.h:122-182):
//      dualMC(double i0, UGrid&, vector<Vertex>& v, vector<Normal>& n,
//             vector<int>& tris, vector<int>& quads, bool sFlag)
//      Results are FP64 and bit-identical to cpp_dmc on the same grid.
//
//  p_mc::DualMarchingCubes
//      as the CUDA reference declares it (dmc/DualMarchingCubes.h:34-90):
//      init<T>(filename), init(dim, SurfaceCase),
//      dualMC(float i0, v, n, t, q, he, hef, hev, config) with float
//      triples, normals negated and halfedges as int4 {origin, face, next,
//      twin}, driven by the config keys of dmc/kernel.cu:25-32.
//
// Errors raise dmc::Error instead of exiting the process.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmc_dmc.h"

namespace tmc_detail {

inline void check(dmc_context* ctx, dmc_status s, const char* what) {
    if (s != DMC_OK) {
        std::string msg = std::string(what) + ": " + dmc_status_str(s);
        if (ctx) msg += std::string(" (") + dmc_last_error(ctx) + ")";
        throw std::runtime_error(msg);
    }
}

// One engine context per thread of use (contexts must not be shared).
class Engine {
public:
    explicit Engine(int device = 0) { check(nullptr, dmc_create(&ctx_, device), "dmc_create"); }
    ~Engine() { if (ctx_) dmc_destroy(ctx_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    dmc_context* get() const { return ctx_; }

private:
    dmc_context* ctx_ = nullptr;
};

}  // namespace tmc_detail

namespace dmc {

using Error = std::runtime_error;

// cpp_dmc/Vector.h:9-110 (the parts callers use)
struct Vector {
    double p[3]{0, 0, 0};
    double& operator[](size_t i) { return p[i]; }
    const double& operator[](size_t i) const { return p[i]; }
    double* data() { return p; }
    void normalize() {
        const double s = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        p[0] /= s;
        p[1] /= s;
        p[2] /= s;
    }
};

// cpp_dmc/UniformGrid.h:18-354: scalars (double) and gradients (double3)
// on nx*ny*nz points, x fastest.
class UniformGrid {
public:
    using Vertex = Vector;
    using Point = Vector;
    using Normal = Vector;
    using Index = std::array<int, 3>;
    using BBox = std::array<Point, 8>;

    // cpp_dmc/UniformGrid.cpp:3-46: 3 int dims, 3 double spacings, float data;
    // gradients estimated and flipped.
    void init(const std::string& filename) {
        std::ifstream f(filename, std::ios::binary);
        if (!f.is_open()) throw Error("UniformGrid::init: cannot open " + filename);
        int n[3];
        double d[3];
        f.read(reinterpret_cast<char*>(n), sizeof(n));
        f.read(reinterpret_cast<char*>(d), sizeof(d));
        if (!f || n[0] < 2 || n[1] < 2 || n[2] < 2) throw Error("UniformGrid::init: bad header in " + filename);
        BBox bb;
        const double mx[3] = {d[0] * (n[0] - 1.), d[1] * (n[1] - 1.), d[2] * (n[2] - 1.)};
        for (int c = 0; c < 8; c++)
            for (int a = 0; a < 3; a++) bb[c][a] = ((c >> a) & 1) ? mx[a] : 0.0;
        init(n[0], n[1], n[2], bb);
        m_dx = d[0];
        m_dy = d[1];
        m_dz = d[2];
        std::vector<float> buf(m_scalars.size());
        f.read(reinterpret_cast<char*>(buf.data()), buf.size() * sizeof(float));
        if (!f) throw Error("UniformGrid::init: short file " + filename);
        for (size_t q = 0; q < buf.size(); q++) m_scalars[q] = buf[q];
        estimateGradient();   // gradients now match the scalars
        flip_gradient();
    }
    // cpp_dmc/UniformGrid.cpp:48-79: unit cube
    void init(const int nx, const int ny, const int nz) {
        BBox bb;
        for (int c = 0; c < 8; c++)
            for (int a = 0; a < 3; a++) bb[c][a] = (c >> a) & 1;
        init(nx, ny, nz, bb);
    }
    // cpp_dmc/UniformGrid.cpp:81-116
    void init(const int nx, const int ny, const int nz, BBox& bb) { init(nx, ny, nz, bb, 0.0); }
    void init(const int nx, const int ny, const int nz, BBox& bb, const double val) {
        m_nx = nx;
        m_ny = ny;
        m_nz = nz;
        const size_t n = (size_t)nx * ny * nz;
        m_scalars.assign(n, val);
        m_gradient.assign(n, Normal{});
        m_bbox = bb;
        m_dx = (m_bbox[7][0] - m_bbox[0][0]) / (static_cast<double>(m_nx) - 1.0);
        m_dy = (m_bbox[7][1] - m_bbox[0][1]) / (static_cast<double>(m_ny) - 1.0);
        m_dz = (m_bbox[7][2] - m_bbox[0][2]) / (static_cast<double>(m_nz) - 1.0);
        grad_estimated_ = false;
    }
    double dx() const { return m_dx; }
    double dy() const { return m_dy; }
    double dz() const { return m_dz; }
    void set_dx(const double d) { m_dx = d; }
    void set_dy(const double d) { m_dy = d; }
    void set_dz(const double d) { m_dz = d; }
    int size() const { return m_nx * m_ny * m_nz; }
    int x_size() const { return m_nx; }
    int y_size() const { return m_ny; }
    int z_size() const { return m_nz; }
    BBox bbox() const { return m_bbox; }
    double minX() const { return m_bbox[0][0]; }
    double minY() const { return m_bbox[0][1]; }
    double minZ() const { return m_bbox[0][2]; }
    double maxX() const { return m_bbox[7][0]; }
    double maxY() const { return m_bbox[7][1]; }
    double maxZ() const { return m_bbox[7][2]; }
    int global_index(const int i, const int j, const int k) const { return (k * m_ny * m_nx + j * m_nx + i); }
    Point point(const int i, const int j, const int k) const {
        return {m_bbox[0][0] + i * m_dx, m_bbox[0][1] + j * m_dy, m_bbox[0][2] + k * m_dz};
    }
    // setting scalars leaves the stored gradients as they are (stale, like the
    // reference), so the engine must use the stored ones from then on
    void scalar(const int gindex, const double val) {
        m_scalars[gindex] = val;
        grad_estimated_ = false;
    }
    void scalar(const int i, const int j, const int k, const double val) {
        m_scalars[global_index(i, j, k)] = val;
        grad_estimated_ = false;
    }
    double scalar(const int gindex) const { return m_scalars[gindex]; }
    double scalar(const int i, const int j, const int k) const { return m_scalars[global_index(i, j, k)]; }
    void setScalars(const double val) {
        std::fill(m_scalars.begin(), m_scalars.end(), val);
        grad_estimated_ = false;
    }
    void gradient(const int i, const int j, const int k, const Normal& g) {
        m_gradient[global_index(i, j, k)] = g;
        grad_estimated_ = false;
    }
    Vector gradient(const int gindex) const { return m_gradient[gindex]; }
    Vector gradient(const int i, const int j, const int k) const { return m_gradient[global_index(i, j, k)]; }
    double max_scalar() const { return *std::max_element(m_scalars.begin(), m_scalars.end()); }
    double min_scalar() const { return *std::min_element(m_scalars.begin(), m_scalars.end()); }
    // cpp_dmc/UniformGrid.cpp:156-184 (with its boundary index rule)
    void estimateGradient() {
        auto index = [](const int i, const int max) { return (i < 0) ? 0 : (i >= max) ? max - i : i; };
        m_gradient.resize(m_scalars.size());
        for (int k = 0; k < m_nz; k++)
            for (int j = 0; j < m_ny; j++)
                for (int i = 0; i < m_nx; i++) {
                    Normal& g = m_gradient[global_index(i, j, k)];
                    g[0] = (scalar(index(i + 1, m_nx), j, k) - scalar(index(i - 1, m_nx), j, k)) / (2 * m_dx);
                    g[1] = (scalar(i, index(j + 1, m_ny), k) - scalar(i, index(j - 1, m_ny), k)) / (2 * m_dy);
                    g[2] = (scalar(i, j, index(k + 1, m_nz)) - scalar(i, j, index(k - 1, m_nz))) / (2 * m_dz);
                }
        grad_estimated_ = true;
        flipped_ = false;
    }
    // cpp_dmc/UniformGrid.cpp:186-197
    void flip_gradient() {
        for (auto& g : m_gradient) {
            g[0] *= -1.;
            g[1] *= -1.;
            g[2] *= -1.;
        }
        flipped_ = !flipped_;
    }

    // engine hooks
    const double* scalars_data() const { return m_scalars.data(); }
    const double* gradients_data() const { return m_gradient.empty() ? nullptr : m_gradient[0].p; }
    // True when the gradients are exactly estimateGradient() + flip_gradient()
    // of the current scalars: the engine then recomputes them on the device.
    bool gradients_from_field() const { return grad_estimated_ && flipped_; }

private:
    int m_nx{0}, m_ny{0}, m_nz{0};
    double m_dx{0}, m_dy{0}, m_dz{0};
    BBox m_bbox{};
    std::vector<double> m_scalars;
    std::vector<Normal> m_gradient;
    bool grad_estimated_{false};
    bool flipped_{false};
};

// cpp_dmc/Volumes.h:19-336 (the implicit surfaces on [-1,1]^3 with their
// analytic central-difference gradients, stored unflipped).
class Volumes {
public:
    enum class Surface {
        Sphere,
        Torus,
        TwoHoledTorus,
        MonkeySaddle,
        GenusTwo,
        iWP,
        Neovius,
        SteinerRoman,
        Kummer,
        Tetrahedron
    };

    // cpp_dmc/Volumes.h:45-106: 3 int dims, 3 float spacings, then nx*ny*nz
    // values of type T (x fastest).  The box runs from 0 to
    // double(float spacing * (n-1)) on each axis, the spacings are the file's
    // floats, and the gradients are estimated and flipped.  Errors throw
    // instead of exiting.
    template <typename T>
    void readFromFile(const std::string& i_file, UniformGrid& ugrid) {
        std::ifstream f(i_file, std::ios::binary);
        if (!f.is_open()) throw Error("Volumes::readFromFile: cannot open " + i_file);
        int n[3];
        float d[3];
        f.read(reinterpret_cast<char*>(n), sizeof(n));
        f.read(reinterpret_cast<char*>(d), sizeof(d));
        if (!f || n[0] < 2 || n[1] < 2 || n[2] < 2) throw Error("Volumes::readFromFile: bad header in " + i_file);
        const double mx[3] = {static_cast<double>(d[0] * (n[0] - 1)), static_cast<double>(d[1] * (n[1] - 1)),
                              static_cast<double>(d[2] * (n[2] - 1))};
        UniformGrid::BBox bb;
        for (int c = 0; c < 8; c++)
            for (int a = 0; a < 3; a++) bb[c][a] = ((c >> a) & 1) ? mx[a] : 0.0;
        ugrid.init(n[0], n[1], n[2], bb, 0);
        ugrid.set_dx(d[0]);
        ugrid.set_dy(d[1]);
        ugrid.set_dz(d[2]);
        std::vector<T> buf((size_t)n[0] * n[1] * n[2]);
        f.read(reinterpret_cast<char*>(buf.data()), buf.size() * sizeof(T));
        if (!f) throw Error("Volumes::readFromFile: short file " + i_file);
        for (size_t q = 0; q < buf.size(); q++) ugrid.scalar((int)q, static_cast<double>(buf[q]));
        ugrid.estimateGradient();
        ugrid.flip_gradient();
    }

    template <Surface T>
    void scalar(UniformGrid& ugrid, const int nx, const int ny, const int nz) {
        UniformGrid::BBox bb;
        for (int c = 0; c < 8; c++)
            for (int a = 0; a < 3; a++) bb[c][a] = ((c >> a) & 1) ? 1.0 : -1.0;
        ugrid.init(nx, ny, nz, bb);
        // the reference reads minX for all three axes (Volumes.h:115-117)
        const double minX = ugrid.minX(), minY = ugrid.minX(), minZ = ugrid.minX();
        const double dx = ugrid.dx(), dy = ugrid.dy(), dz = ugrid.dz();
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    const double x = minX + i * dx, y = minY + j * dy, z = minZ + k * dz;
                    ugrid.scalar(i, j, k, surface(T, x, y, z));
                    Vector g;
                    g[0] = (surface(T, x + dx, y, z) - surface(T, x - dx, y, z)) / (2 * dx);
                    g[1] = (surface(T, x, y + dy, z) - surface(T, x, y - dy, z)) / (2 * dy);
                    g[2] = (surface(T, x, y, z + dz) - surface(T, x, y, z - dz)) / (2 * dz);
                    ugrid.gradient(i, j, k, g);
                }
    }

    static double surface(Surface s, const double x_, const double y_, const double z_) {
        auto sq = [](double v) { return v * v; };
        const double pi = 3.14159265358979323846;
        switch (s) {
        case Surface::Sphere: return x_ * x_ + y_ * y_ + z_ * z_;
        case Surface::Torus: {
            const double R = 0.6 * 0.6, r = 0.3 * 0.3;
            double val = (x_ * x_ + y_ * y_ + z_ * z_ + R - r);
            val = val * val;
            return val - 4 * R * (x_ * x_ + y_ * y_);
        }
        case Surface::TwoHoledTorus: {
            const double R = sq(0.4), r = sq(0.2);
            const double x1 = x_ + 0.4, x2 = x_ - 0.4;
            double v1 = sq((sq(x1) + sq(y_) + sq(z_) + R - r));
            v1 = v1 - 4 * R * (sq(x1) + sq(y_));
            double v2 = sq((sq(x2) + sq(y_) + sq(z_) + R - r));
            v2 = v2 - 4 * R * (sq(x2) + sq(y_));
            return std::min(v1, v2);
        }
        case Surface::MonkeySaddle: {
            const double a = 0.5, x = a * x_, y = a * y_, z = a * z_;
            return z - x * x * x - 3 * x * y * y;
        }
        case Surface::GenusTwo: {
            double x = (x_ + 1.0) / 2.0, y = (y_ + 1.0) / 2.0, z = (z_ + 1.0) / 2.0;
            x = 1.0 * (4 * x - 2);
            y = 1.0 * (4 * y - 2);
            z = 1.0 * (4 * z - 2);
            return 2 * y * (y * y - 3 * x * x) * (1 - z * z) + (x * x + y * y) * (x * x + y * y) -
                   (9 * z * z - 1) * (1 - z * z);
        }
        case Surface::iWP: {
            const double a = 5.01, x = a * (x_ + 1) * pi, y = a * (y_ + 1) * pi, z = a * (z_ + 1) * pi;
            return std::cos(x) * std::cos(y) + std::cos(y) * std::cos(z) + std::cos(z) * std::cos(x) -
                   std::cos(x) * std::cos(y) * std::cos(z);
        }
        case Surface::Neovius: {
            const double a = 2, x = a * (x_ + 1) * pi, y = a * (y_ + 1) * pi, z = a * (z_ + 1) * pi;
            return 3 * (std::cos(x) + std::cos(y) + std::cos(z)) + 4 * std::cos(x) * std::cos(y) * std::cos(z);
        }
        case Surface::SteinerRoman: {
            const double x = x_, y = y_, z = z_;
            return sq(x * x + y * y + z * z - 1.0f) - (sq(z - 1) - 2.0f * x * x) * (sq(z + 1) - 2 * y * y);
        }
        case Surface::Kummer: {
            const double x = 2 * x_, y = 2 * y_, z = 2 * z_, mu = 1.3;
            const double lambda = (3 * mu * mu - 1) / (3 - mu * mu), w2 = std::sqrt(2);
            const double p = 1 - z - w2 * x, q = 1 - z + w2 * x, r = 1 + z + w2 * y, t = 1 + z - w2 * y;
            double v = (x * x + y * y + z * z - mu * mu);
            v = v * v;
            return v - lambda * p * q * r * t;
        }
        case Surface::Tetrahedron: return tetrahedron(x_, y_, z_);
        }
        return 0;
    }

private:
    // cpp_dmc/Vector.h:165-184 and Volumes.h:248-335: signed distance to the
    // nearest face of the tetrahedron (-0.8,-0.8,-0.8), (0.8,-0.8,-0.8),
    // (0.8,0.8,-0.8), (0,0,0.8), the same operations in the same order.
    struct V3 {
        double x, y, z;
    };
    static V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
    static V3 add(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
    static V3 scale(double s, const V3& a) { return {a.x * s, a.y * s, a.z * s}; }
    static double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    static V3 cross(const V3& a, const V3& b) {
        return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    }
    static double dist(const V3& a, const V3& b) {
        const double d = (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y) + (a.z - b.z) * (a.z - b.z);
        return std::sqrt(d);
    }
    static V3 unit(V3 a) {
        const double s = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
        a.x /= s;
        a.y /= s;
        a.z /= s;
        return a;
    }
    static double point_edge(const V3& v0, const V3& v1, const V3& p) {
        const V3 n = unit(sub(v1, v0));
        const double l = dot(sub(p, v0), n);
        if (l <= 0 || l >= dist(v0, v1)) {
            const double l0 = dist(p, v0), l1 = dist(p, v1);
            return l0 < l1 ? l0 : l1;
        }
        return dist(p, add(v0, scale(l, n)));
    }
    static double point_triangle(const V3& v0, const V3& v1, const V3& v2, const V3& p) {
        const V3 n = unit(cross(sub(v1, v0), sub(v2, v0)));
        const double l = -dot(sub(p, v0), n);
        const V3 q = add(p, scale(l, n));
        const double d0 = dot(n, cross(sub(v0, q), sub(v1, q)));
        const double d1 = dot(n, cross(sub(v1, q), sub(v2, q)));
        const double d2 = dot(n, cross(sub(v2, q), sub(v0, q)));
        if (d0 > 0 && d1 > 0 && d2 > 0) return dot(n, sub(p, v0));
        const double e0 = point_edge(v0, v1, p), e1 = point_edge(v1, v2, p), e2 = point_edge(v2, v0, p);
        return std::min(std::min(e0, e1), e2);
    }
    static double tetrahedron(const double x_, const double y_, const double z_) {
        const V3 v0{-0.8, -0.8, -0.8}, v1{0.8, -0.8, -0.8}, v2{0.8, 0.8, -0.8}, v3{0.0, 0.0, 0.8};
        const V3 p{x_, y_, z_};
        const double d0 = point_triangle(v0, v2, v1, p);
        const double d1 = point_triangle(v0, v1, v3, p);
        const double d2 = point_triangle(v1, v2, v3, p);
        const double d3 = point_triangle(v0, v3, v2, p);
        double d = std::min(std::fabs(d0), std::fabs(d1));
        d = std::min(d, std::fabs(d2));
        d = std::min(d, std::fabs(d3));
        if (d == std::fabs(d0)) return d0;
        if (d == std::fabs(d1)) return d1;
        if (d == std::fabs(d2)) return d2;
        if (d == std::fabs(d3)) return d3;
        return d;
    }

public:
};

// cpp_dmc/DualMarchingCubes.h:122-182
class DualMarchingCubes {
public:
    using Vertex = Vector;
    using Point = Vector;
    using Normal = Vector;
    using UGrid = UniformGrid;

    explicit DualMarchingCubes(int device = 0) : engine_(device) {}

    // cpp_dmc/DualMarchingCubes.h:163-165; sFlag runs 3X3Y (colour, then
    // isolated) and 3333 as cpp_dmc/DualMarchingCubes.cpp:62-76 does.
    void dualMC(const double i0, UGrid& ugrid, std::vector<Vertex>& v, std::vector<Normal>& n,
                std::vector<int>& tris, std::vector<int>& quads, const bool sFlag) {
        dmc_grid g{ugrid.x_size(), ugrid.y_size(), ugrid.z_size(), ugrid.minX(), ugrid.minY(), ugrid.minZ(),
                   ugrid.dx(), ugrid.dy(), ugrid.dz()};
        uint32_t flags = DMC_TRIANGLES | (sFlag ? DMC_SIMPLIFY_ALL : 0u);
        const double* grads = nullptr;
        if (!ugrid.gradients_from_field()) {
            flags |= DMC_GRADIENTS_SUPPLIED;
            grads = ugrid.gradients_data();
        }
        dmc_mesh m{};
        tmc_detail::check(engine_.get(), dmc_extract(engine_.get(), &g, ugrid.scalars_data(), DMC_FIELD_F64, i0,
                                                     flags, grads, &m), "dmc_extract");
        std::vector<double> pos(3 * m.nv), nrm(3 * m.nv);
        quads.assign(4 * m.nq, 0);
        tris.assign(3 * m.nt, 0);
        bnd_.assign(m.nbnd, 0);
        tmc_detail::check(engine_.get(),
                          dmc_mesh_download(engine_.get(), pos.data(), nrm.data(), quads.data(), tris.data(), nullptr,
                                            nullptr, nullptr, bnd_.data()),
                          "dmc_mesh_download");
        v.resize(m.nv);
        n.resize(m.nv);
        for (int64_t a = 0; a < m.nv; a++)
            for (int c = 0; c < 3; c++) {
                v[a][c] = pos[3 * a + c];
                n[a][c] = nrm[3 * a + c];
            }
        last_ = m;
    }
    // cpp_dmc/DualMarchingCubes.h:171-182
    bool getBndVertices(std::vector<int>& bndV) {
        if (bnd_.empty()) return false;
        bndV.insert(bndV.end(), bnd_.begin(), bnd_.end());
        return true;
    }
    const dmc_mesh& lastMesh() const { return last_; }
    dmc_context* context() const { return engine_.get(); }

private:
    tmc_detail::Engine engine_;
    std::vector<int> bnd_;
    dmc_mesh last_{};
};

}  // namespace dmc

namespace p_mc {

// dmc/DualMarchingCubes.h:34-90 (CUDA reference surface)
class DualMarchingCubes {
public:
    using Vertex = std::array<float, 3>;
    using Point = std::array<float, 3>;
    using Normal = std::array<float, 3>;
    using Halfedge = std::array<int, 4>;
    using HalfedgeFace = int;
    using HalfedgeVertex = int;
    using Triangle = std::array<int, 3>;
    using Quadrilateral = std::array<int, 4>;
    // dmc/UniformGrid.h:410-418
    enum class SurfaceCase : int {
        Sphere = 0, Torus = 1, TwoHoledTorus = 2, FourHoledTorus = 3, GenusTwo = 4,
        iWP = 5, pwHybrid = 6, neovius = 7, Goursat = 8, SteinerRoman = 9
    };

    explicit DualMarchingCubes(int device = 0) : engine_(device) {}

    // dmc/UniformGrid.h:430-462: 3 int dims, 3 float spacings, T data
    template <typename T>
    void init(const std::string& filename) {
        std::ifstream f(filename, std::ios::binary);
        if (!f.is_open()) throw std::runtime_error("init: cannot open " + filename);
        int n[3];
        float d[3];
        f.read(reinterpret_cast<char*>(n), sizeof(n));
        f.read(reinterpret_cast<char*>(d), sizeof(d));
        if (!f || n[0] < 2 || n[1] < 2 || n[2] < 2) throw std::runtime_error("init: bad header in " + filename);
        grid_ = dmc_grid{n[0], n[1], n[2], 0, 0, 0, d[0], d[1], d[2]};
        std::vector<T> buf((size_t)n[0] * n[1] * n[2]);
        f.read(reinterpret_cast<char*>(buf.data()), buf.size() * sizeof(T));
        if (!f) throw std::runtime_error("init: short file " + filename);
        field_.resize(buf.size());
        for (size_t q = 0; q < buf.size(); q++) field_[q] = static_cast<float>(buf[q]);
    }

    // dmc/UniformGrid.cu:177-223 with the surfaces of :21-174, on [-1,1]^3
    void init(const std::array<int, 3>& dim, SurfaceCase sc) {
        const float dx = 2.0f / (dim[0] - 1.0f), dy = 2.0f / (dim[1] - 1.0f), dz = 2.0f / (dim[2] - 1.0f);
        grid_ = dmc_grid{dim[0], dim[1], dim[2], -1.0, -1.0, -1.0, dx, dy, dz};
        field_.resize((size_t)dim[0] * dim[1] * dim[2]);
        for (int k = 0; k < dim[2]; k++)
            for (int j = 0; j < dim[1]; j++)
                for (int i = 0; i < dim[0]; i++) {
                    const float x = -1.0f + i * dx, y = -1.0f + j * dy, z = -1.0f + k * dz;
                    field_[((size_t)k * dim[1] + j) * dim[0] + i] = surface(sc, x, y, z);
                }
    }

    // dmc/DualMarchingCubes.h:74-77.  config keys (dmc/kernel.cu:25-32;
    // missing keys are 0): p3X3YColor, p3X3YOld, p3333,
    // halfedge-datastructure; valence fills the valenceDist* members with
    // the valence distributions of the mesh after each enabled stage and of
    // the triangles (dmc/DualMarchingCubes.cu:966-1043), element-quality
    // fills elementQualityQuads / elementQualityTris (:1050-1060), and
    // non-manifold sets nrNonManifoldEdges.  The distribution after a stage
    // is that of the mesh extracted with the stages up to it.
    void dualMC(const float i0, std::vector<Vertex>& v, std::vector<Normal>& n, std::vector<Triangle>& t,
                std::vector<Quadrilateral>& q, std::vector<Halfedge>& o_he, std::vector<HalfedgeFace>& o_hef,
                std::vector<HalfedgeVertex>& o_hev, std::map<std::string, int>& config) {
        if (field_.empty()) throw std::runtime_error("dualMC: call init first");
        auto key = [&](const char* k) { auto it = config.find(k); return it != config.end() && it->second; };
        clearInfo();
        uint32_t flags = DMC_TRIANGLES;
        const bool valence = key("valence");
        auto dist = [&](uint32_t f, bool tris, std::vector<unsigned int>& out) {
            dmc_mesh mm{};
            tmc_detail::check(engine_.get(),
                              dmc_extract(engine_.get(), &grid_, field_.data(), DMC_FIELD_F32, (double)i0, f, nullptr,
                                          &mm),
                              "dmc_extract");
            histogram(tris, out);
        };
        if (valence) dist(flags, false, valenceDistDMC);
        if (key("p3X3YColor")) {
            flags |= DMC_SIMPLIFY_3X3Y_COLOR;
            if (valence) dist(flags, false, valenceDistP3X3YColor);
        }
        if (key("p3X3YOld")) {
            flags |= DMC_SIMPLIFY_3X3Y_ISOLATED;
            if (valence) dist(flags, false, valenceDistP3X3YOld);
        }
        if (key("p3333")) {
            flags |= DMC_SIMPLIFY_3333;
            if (valence) dist(flags, false, valenceDistP3333);
        }
        dmc_mesh m{};
        tmc_detail::check(engine_.get(),
                          dmc_extract(engine_.get(), &grid_, field_.data(), DMC_FIELD_F32, (double)i0, flags, nullptr,
                                      &m),
                          "dmc_extract");
        v.resize(m.nv);
        n.resize(m.nv);
        q.resize(m.nq);
        t.resize(m.nt);
        tmc_detail::check(engine_.get(),
                          dmc_mesh_download_f32(engine_.get(), v.empty() ? nullptr : v[0].data(),
                                                n.empty() ? nullptr : n[0].data(), 1),
                          "dmc_mesh_download_f32");
        const bool he = key("halfedge-datastructure");
        std::vector<int32_t> hes(he ? 5 * 4 * m.nq : 0);
        o_hev.resize(he ? m.nv : 0);
        o_hef.resize(he ? m.nq : 0);
        tmc_detail::check(engine_.get(),
                          dmc_mesh_download(engine_.get(), nullptr, nullptr, q.empty() ? nullptr : q[0].data(),
                                            t.empty() ? nullptr : t[0].data(), he ? hes.data() : nullptr,
                                            he ? o_hev.data() : nullptr, he ? o_hef.data() : nullptr, nullptr),
                          "dmc_mesh_download");
        o_he.resize(he ? 4 * m.nq : 0);
        for (size_t h = 0; h < o_he.size(); h++)
            for (int c = 0; c < 4; c++) o_he[h][c] = hes[5 * h + c];
        if (valence) histogram(true, valenceDistTris);
        if (key("element-quality")) {
            elementQualityQuads.resize(m.nq);
            elementQualityTris.resize(m.nt);
            if (m.nq)
                tmc_detail::check(engine_.get(), dmc_element_quality(engine_.get(), 0, elementQualityQuads.data()),
                                  "dmc_element_quality");
            if (m.nt)
                tmc_detail::check(engine_.get(), dmc_element_quality(engine_.get(), 1, elementQualityTris.data()),
                                  "dmc_element_quality");
        }
        if (key("non-manifold")) nrNonManifoldEdges = (int)m.nonmanifold_edges;
        nrFailedProjections1 = (int)m.failed_projections_l1;
        nrFailedProjections2 = (int)m.failed_projections_l2;
    }
    // dmc/HalfedgeMesh.cu:430-450: unique edges of the last mesh (rendering)
    int edges(std::vector<std::array<int, 2>>& e) {
        int64_t ne = 0;
        tmc_detail::check(engine_.get(), dmc_mesh_edges(engine_.get(), nullptr, &ne), "dmc_mesh_edges");
        e.resize((size_t)ne);
        if (ne)
            tmc_detail::check(engine_.get(), dmc_mesh_edges(engine_.get(), e[0].data(), &ne), "dmc_mesh_edges");
        return (int)ne;
    }
    // dmc/DualMarchingCubes.h:94-127
    void clearInfo() {
        valenceDistDMC.clear();
        valenceDistP3X3YColor.clear();
        valenceDistP3X3YOld.clear();
        valenceDistP3333.clear();
        valenceDistTris.clear();
        elementQualityTris.clear();
        elementQualityQuads.clear();
        nrNonManifoldEdges = nrFailedProjections1 = nrFailedProjections2 = 0;
    }
    std::vector<unsigned int> valenceDistDMC;
    std::vector<unsigned int> valenceDistP3X3YColor;
    std::vector<unsigned int> valenceDistP3X3YOld;
    std::vector<unsigned int> valenceDistP3333;
    std::vector<unsigned int> valenceDistTris;
    std::vector<float> elementQualityTris;
    std::vector<float> elementQualityQuads;
    int nrNonManifoldEdges{0};
    int nrFailedProjections1{0};
    int nrFailedProjections2{0};

    const std::vector<float>& scalars() const { return field_; }
    const dmc_grid& grid() const { return grid_; }

    // valence distribution of the engine's last mesh, trimmed as the
    // reference trims it (dmc/VertexValence.cu:114-120)
    void histogram(bool tris, std::vector<unsigned int>& out) {
        uint32_t h[100];
        int32_t len = 0;
        tmc_detail::check(engine_.get(), dmc_valence_histogram(engine_.get(), tris ? 1 : 0, h, &len, nullptr),
                          "dmc_valence_histogram");
        out.assign(h, h + len);
    }

    static float surface(SurfaceCase sc, const float x, const float y, const float z) {
        auto sq = [](const float a) { return a * a; };
        auto qu = [](const float a) { return a * a * a * a; };
        auto torus_h = [&](float px, float py, float pz, float cx, float cy, float cz, float r1, float r2) {
            const float c = sq(r1), a = sq(r2);
            const float X = sq(px - cx), Y = sq(py - cy), Z = sq(pz - cz);
            return sq(X + Y + Z + c - a) - 4 * c * (X + Y);
        };
        auto torus_v = [&](float px, float py, float pz, float cx, float cy, float cz, float r1, float r2) {
            const float c = sq(r1), a = sq(r2);
            const float X = sq(px - cx), Y = sq(py - cy), Z = sq(pz - cz);
            return sq(X + Y + Z + c - a) - 4 * c * (X + Z);
        };
        const float pi = 3.14159265358979323846f;
        switch (sc) {
        case SurfaceCase::Sphere: return x * x + y * y + z * z - 0.16f;
        case SurfaceCase::Torus: return torus_h(x, y, z, 0, 0, 0, 0.3f, 0.15f);
        case SurfaceCase::TwoHoledTorus:
            return std::fmin(torus_h(x, y, z, -0.38f, 0, 0, 0.3f, 0.15f), torus_h(x, y, z, 0.2f, 0.38f, 0, 0.3f, 0.15f));
        case SurfaceCase::FourHoledTorus: {
            const float v1 = torus_h(x, y, z, -0.38f, 0, 0, 0.3f, 0.15f), v2 = torus_h(x, y, z, 0.38f, 0, 0, 0.3f, 0.15f);
            const float v3 = torus_v(x, y, z, 0, 0, -0.38f, 0.3f, 0.15f), v4 = torus_v(x, y, z, 0, 0, 0.38f, 0.3f, 0.15f);
            return std::fmin(std::fmin(v1, v2), std::fmin(v3, v4));
        }
        case SurfaceCase::GenusTwo: {
            const float X = 4.0f * ((x + 1.0f) / 2.0f) - 2.0f, Y = 4.0f * ((y + 1.0f) / 2.0f) - 2.0f,
                        Z = 4.0f * ((z + 1.0f) / 2.0f) - 2.0f;
            float r = 2 * Y * (Y * Y - 3 * X * X) * (1 - Z * Z);
            r += (X * X + Y * Y) * (X * X + Y * Y);
            r -= (9 * Z * Z - 1) * (1 - Z * Z);
            return r;
        }
        case SurfaceCase::iWP: {
            const float a = 5.01f, X = a * (x + 1) * pi, Y = a * (y + 1) * pi, Z = a * (z + 1) * pi;
            return std::cos(X) * std::cos(Y) + std::cos(Y) * std::cos(Z) + std::cos(Z) * std::cos(X) -
                   std::cos(X) * std::cos(Y) * std::cos(Z);
        }
        case SurfaceCase::pwHybrid: {
            const float a = 1.01f, X = a * (x + 1) * pi, Y = a * (y + 1) * pi, Z = a * (z + 1) * pi;
            return 4.0f * (std::cos(X) * std::cos(Y) + std::cos(Y) * std::cos(Z) + std::cos(Z) * std::cos(X)) -
                   3 * std::cos(X) * std::cos(Y) * std::cos(Z) + 0.8f;
        }
        case SurfaceCase::neovius: {
            const float X = (x + 1) * pi, Y = (y + 1) * pi, Z = (z + 1) * pi;
            return 3 * (std::cos(X) + std::cos(Y) + std::cos(Z)) + 4 * std::cos(X) * std::cos(Y) * std::cos(Z);
        }
        case SurfaceCase::Goursat:
            return qu(x) + qu(y) + qu(z) - (sq(x) + sq(y) + sq(z)) + 0.5f;
        case SurfaceCase::SteinerRoman: {
            const float X = 1.5f * x, Y = 1.5f * y, Z = 1.5f * z;
            return sq(X * X + Y * Y + Z * Z - 1.0f) - (sq(Z - 1) - 2.0f * X * X) * (sq(Z + 1) - 2 * Y * Y);
        }
        }
        return 0.0f;
    }

private:
    tmc_detail::Engine engine_;
    dmc_grid grid_{};
    std::vector<float> field_;
};

}  // namespace p_mc

// ---------------------------------------------------------------------------
// tmc::MarchingCubes -- the reference's topologically correct Marching Cubes
// (mc/MarchingCubes.h:113-420, mc/MarchingCubes.cpp).  operator() reads the
// volume file, extracts (tmc_extract: t_mc + p_slice, bit-identical to the
// reference), optionally orients the triangles (connectivity(), c_flag) and
// writes the OBJ / OFF files in the reference's formats.  The mesh is also
// available through points(), normals() and triangles().
// ---------------------------------------------------------------------------
namespace tmc {

using Error = std::runtime_error;

namespace detail {

// connectivity() of mc/MarchingCubes.cpp:1769-2226, the part that changes
// the mesh: each triangle's neighbour across its edges (v1,v2), (v2,v0),
// (v0,v1) is the first other triangle holding both vertices in the vertex's
// triangle list (triangle order); then, triangle by triangle, a depth-first
// walk over unvisited neighbours flips (v[1] <-> v[2], with its neighbours
// across edges 1 and 2) every neighbour that does not run the shared edge
// the other way -- before the neighbour itself is visited, so a triangle can
// be flipped more than once.  Returns the number of flips.
inline long orient_triangles(int64_t nv, std::vector<std::array<int, 3>>& T) {
    const int64_t nt = (int64_t)T.size();
    std::vector<std::vector<int>> v_tris((size_t)nv);
    for (int64_t t = 0; t < nt; t++)
        for (int x = 0; x < 3; x++) v_tris[(size_t)T[t][x]].push_back((int)t);
    auto has_edge = [](int a, int b, const std::array<int, 3>& q) {
        return (a == q[1] && b == q[2]) || (a == q[2] && b == q[1]) || (a == q[0] && b == q[2]) ||
               (a == q[2] && b == q[0]) || (a == q[1] && b == q[0]) || (a == q[0] && b == q[1]);
    };
    std::vector<std::array<int, 3>> nb((size_t)nt);
    for (int64_t t = 0; t < nt; t++) {
        const int v0 = T[t][0], v1 = T[t][1], v2 = T[t][2];
        nb[t] = {-1, -1, -1};
        const int e[3][2] = {{v1, v2}, {v2, v0}, {v0, v1}};
        for (int x = 0; x < 3; x++)
            for (int tt : v_tris[(size_t)e[x][0]])
                if (tt != t && has_edge(e[x][0], e[x][1], T[tt])) {
                    nb[t][x] = tt;
                    break;
                }
    }
    std::vector<char> oriented((size_t)nt, 0);
    std::vector<int> stack;
    long flips = 0;
    // the neighbour must run the shared edge a -> b as b -> a; else flip it
    auto match = [&](int a, int b, int id) {
        std::array<int, 3>& q = T[id];
        const int t0 = q[0], t1 = q[1], t2 = q[2];
        if ((a == t2 && b == t1) || (a == t0 && b == t2) || (a == t1 && b == t0)) return;
        flips++;
        q[1] = t2;
        q[2] = t1;
        std::swap(nb[id][1], nb[id][2]);
    };
    auto visit = [&](int c) {
        const std::array<int, 3> q = T[c];
        const std::array<int, 3> n = nb[c];
        const int e[3][2] = {{q[1], q[2]}, {q[2], q[0]}, {q[0], q[1]}};
        for (int x = 0; x < 3; x++)
            if (n[x] != -1 && !oriented[n[x]]) {
                match(e[x][0], e[x][1], n[x]);
                stack.push_back(n[x]);
            }
    };
    for (int64_t t = 0; t < nt; t++) {
        if (oriented[t]) continue;
        oriented[t] = 1;
        visit((int)t);
        while (!stack.empty()) {
            const int c = stack.back();
            stack.pop_back();
            if (oriented[c]) continue;
            oriented[c] = 1;
            visit(c);
        }
    }
    return flips;
}

}  // namespace detail

class MarchingCubes {
public:
    using Point = std::array<double, 3>;
    using Normal = std::array<double, 3>;
    using Triangle = std::array<int, 3>;

    explicit MarchingCubes(int device = 0) : engine_(device) {}

    // mc/MarchingCubes.cpp:11-147.  The file holds 3 ushort dims, 3 float
    // spacings and the ushort values, x fastest; the grid's origin is 0 and
    // its spacing (d * (n - 1.)) / (n - 1), as the reference's bounding box
    // gives it.
    void operator()(const double i0, const std::string& i_file, const bool c_flag, const bool obj_flag,
                    const std::string& o_objF, const bool off_flag, const std::string& o_offF) {
        std::ifstream f(i_file, std::ios::binary);
        if (!f) throw Error("can't open input file " + i_file);
        uint16_t dims[3];
        float sp[3];
        f.read(reinterpret_cast<char*>(dims), sizeof(dims));
        f.read(reinterpret_cast<char*>(sp), sizeof(sp));
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];
        std::vector<uint16_t> raw(n);
        f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(n * sizeof(uint16_t)));
        if (!f) throw Error("short volume file " + i_file);
        std::vector<float> field(raw.begin(), raw.end());
        dmc_grid g{};
        g.nx = dims[0];
        g.ny = dims[1];
        g.nz = dims[2];
        const int nn[3] = {g.nx, g.ny, g.nz};
        double d[3];
        for (int x = 0; x < 3; x++) d[x] = ((double)sp[x] * (nn[x] - 1.)) / (nn[x] - 1);
        g.dx = d[0];
        g.dy = d[1];
        g.dz = d[2];
        t_mc(g, field.data(), i0);
        if (c_flag) connectivity();
        if (obj_flag) write_obj(o_objF);
        if (off_flag) write_off(o_offF);
    }

    // t_mc (mc/MarchingCubes.cpp:150-318) on a float field of g's shape
    void t_mc(const dmc_grid& g, const float* field, double i0) {
        tmc_mesh m{};
        tmc_detail::check(engine_.get(), tmc_extract(engine_.get(), &g, field, DMC_FIELD_F32, i0, 0, &m),
                          "tmc_extract");
        points_.resize((size_t)m.nv);
        normals_.resize((size_t)m.nv);
        triangles_.resize((size_t)m.nt);
        tmc_detail::check(engine_.get(),
                          tmc_download(engine_.get(), points_.empty() ? nullptr : points_[0].data(),
                                       normals_.empty() ? nullptr : normals_[0].data(),
                                       triangles_.empty() ? nullptr : triangles_[0].data()),
                          "tmc_download");
    }

    // mc/MarchingCubes.cpp:1769-2226: consistent triangle orientation
    void connectivity() { detail::orient_triangles((int64_t)points_.size(), triangles_); }

    const std::vector<Point>& points() const { return points_; }
    const std::vector<Normal>& normals() const { return normals_; }
    const std::vector<Triangle>& triangles() const { return triangles_; }

    // mc/MarchingCubes.cpp:95-118 (the face line repeats the second normal
    // index for the third vertex, as there)
    void write_obj(const std::string& path) const {
        std::ofstream o(path);
        if (!o) throw Error("can't open output file " + path);
        o << "# Topologically correct and manifold isosurface\n";
        for (const Point& p : points_)
            o << "v " << std::setprecision(7) << std::fixed << p[0] << " " << p[1] << " " << p[2] << std::endl;
        for (const Normal& q : normals_)
            o << "vn " << std::setprecision(7) << std::fixed << q[0] << " " << q[1] << " " << q[2] << std::endl;
        for (const Triangle& t : triangles_)
            o << "f " << (t[0] + 1) << "//" << (t[0] + 1) << " " << (t[1] + 1) << "//" << (t[1] + 1) << " "
              << (t[2] + 1) << "//" << (t[1] + 1) << std::endl;
    }

    // mc/MarchingCubes.cpp:122-143
    void write_off(const std::string& path) const {
        std::ofstream o(path);
        if (!o) throw Error("can't open output file " + path);
        o << "OFF\n";
        o << points_.size() << " " << triangles_.size() << " " << 0 << std::endl;
        for (const Point& p : points_) o << p[0] << " " << p[1] << " " << p[2] << std::endl;
        for (const Triangle& t : triangles_) o << "3 " << t[0] << " " << t[1] << " " << t[2] << std::endl;
    }

private:
    tmc_detail::Engine engine_;
    std::vector<Point> points_;
    std::vector<Normal> normals_;
    std::vector<Triangle> triangles_;
};

}  // namespace tmc
