// vecrand.hpp -- drop-in for ParM's src/vecrand.hpp: flt, Vec, Matrix, NDIM
// and the seeded random helpers, without Eigen or Boost.
//
// Build with -DVEC2D or -DVEC3D (default 3D), as the reference's makefile
// does (src/vecrand.hpp:3-16).  Vec keeps the subset of Eigen's API that
// ParM's callers use: operator[] / (), Zero(), +, -, * and / by scalars,
// dot, squaredNorm, norm, normalized, cwiseProduct, cwiseQuotient.
// squaredNorm sums (x*x + y*y) + z*z, the order the device path uses.
#ifndef VECRAND_H
#define VECRAND_H

#ifdef VEC2D
#define NDIM 2
#else
#ifndef VEC3D
#define VEC3D
#endif
#define NDIM 3
#endif

#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <vector>

// Callers of the C++ API get the reference's global names; a build that
// links several NDIM variants into one process (the Python modules) defines
// PARM_NS_BEGIN / PARM_NS_END to put each variant in its own namespace.
#ifndef PARM_NS_BEGIN
#define PARM_NS_BEGIN
#define PARM_NS_END
#endif

using namespace std;

PARM_NS_BEGIN
typedef unsigned int uint;
typedef double flt;

class Vec {
    flt c_[NDIM];

   public:
    Vec() {
        for (int k = 0; k < NDIM; k++) c_[k] = 0;
    }
#ifdef VEC2D
    Vec(flt x, flt y) {
        c_[0] = x;
        c_[1] = y;
    }
#else
    Vec(flt x, flt y, flt z) {
        c_[0] = x;
        c_[1] = y;
        c_[2] = z;
    }
#endif
    static Vec Zero() { return Vec(); }
    static int size() { return NDIM; }
    flt &operator[](int k) { return c_[k]; }
    flt operator[](int k) const { return c_[k]; }
    flt &operator()(int k) { return c_[k]; }
    flt operator()(int k) const { return c_[k]; }
    const flt *data() const { return c_; }
    flt *data() { return c_; }
    Vec &operator+=(const Vec &o) {
        for (int k = 0; k < NDIM; k++) c_[k] += o.c_[k];
        return *this;
    }
    Vec &operator-=(const Vec &o) {
        for (int k = 0; k < NDIM; k++) c_[k] -= o.c_[k];
        return *this;
    }
    Vec &operator*=(flt s) {
        for (int k = 0; k < NDIM; k++) c_[k] *= s;
        return *this;
    }
    Vec &operator/=(flt s) {
        for (int k = 0; k < NDIM; k++) c_[k] /= s;
        return *this;
    }
    Vec operator+(const Vec &o) const { Vec r(*this); return r += o; }
    Vec operator-(const Vec &o) const { Vec r(*this); return r -= o; }
    Vec operator-() const {
        Vec r;
        for (int k = 0; k < NDIM; k++) r.c_[k] = -c_[k];
        return r;
    }
    Vec operator*(flt s) const { Vec r(*this); return r *= s; }
    Vec operator/(flt s) const { Vec r(*this); return r /= s; }
    flt dot(const Vec &o) const {
        flt s = c_[0] * o.c_[0];
        for (int k = 1; k < NDIM; k++) s = s + c_[k] * o.c_[k];
        return s;
    }
    flt squaredNorm() const { return dot(*this); }
    flt norm() const { return std::sqrt(squaredNorm()); }
    Vec normalized() const { return *this / norm(); }
    Vec cwiseProduct(const Vec &o) const {
        Vec r;
        for (int k = 0; k < NDIM; k++) r.c_[k] = c_[k] * o.c_[k];
        return r;
    }
    Vec cwiseQuotient(const Vec &o) const {
        Vec r;
        for (int k = 0; k < NDIM; k++) r.c_[k] = c_[k] / o.c_[k];
        return r;
    }
    bool operator==(const Vec &o) const {
        for (int k = 0; k < NDIM; k++)
            if (c_[k] != o.c_[k]) return false;
        return true;
    }
};
inline Vec operator*(flt s, const Vec &v) { return v * s; }

class Matrix {
    flt m_[NDIM][NDIM];

   public:
    Matrix() {
        for (int i = 0; i < NDIM; i++)
            for (int j = 0; j < NDIM; j++) m_[i][j] = 0;
    }
    static Matrix Zero() { return Matrix(); }
    flt &operator()(int i, int j) { return m_[i][j]; }
    flt operator()(int i, int j) const { return m_[i][j]; }
};

const flt OVERNDIM = ((flt)1.0) / NDIM;

// Seeded generators for initial conditions (the reference uses Boost's
// mt19937; this build uses the standard library's, so streams differ).
inline std::mt19937 &randengine() {
    static std::mt19937 e(5489u);
    return e;
}
inline unsigned int seed(unsigned int n) {
    randengine().seed(n);
    return n;
}
inline flt rand01() { return std::uniform_real_distribution<flt>(0.0, 1.0)(randengine()); }
inline flt gauss01() { return std::normal_distribution<flt>(0.0, 1.0)(randengine()); }
inline Vec rand_vec() {
    Vec v;
    for (int k = 0; k < NDIM; k++) v[k] = gauss01();
    return v;
}
inline Vec rand_vec_boxed() {
    Vec v;
    for (int k = 0; k < NDIM; k++) v[k] = rand01();
    return v;
}

PARM_NS_END
#endif
