// Vector3D.h — the 3-vector type of the reference's EPP API
// (src/EPP/Vector3D.h:5-83): public X, Y, Z; + - on vectors, * as dot
// product, % as cross product, scaling, Length, Normalize, ordering.
#pragma once

#include <cmath>
#include <ostream>
#include <tuple>

template <typename T>
class Vector3D {
 public:
    T X{}, Y{}, Z{};

    Vector3D() = default;
    Vector3D(const T& v) : X(v), Y(v), Z(v) {}
    explicit Vector3D(const T& x, const T& y, const T& z) : X(x), Y(y), Z(z) {}
    template <typename O>
    Vector3D(const Vector3D<O>& o) : X(static_cast<T>(o.X)), Y(static_cast<T>(o.Y)), Z(static_cast<T>(o.Z)) {}

    const Vector3D& operator+() const { return *this; }
    Vector3D operator-() const { return Vector3D(-X, -Y, -Z); }
    template <typename O>
    Vector3D operator+(const Vector3D<O>& o) const { return Vector3D(X + o.X, Y + o.Y, Z + o.Z); }
    template <typename O>
    Vector3D operator-(const Vector3D<O>& o) const { return Vector3D(X - o.X, Y - o.Y, Z - o.Z); }
    template <typename O>
    T operator*(const Vector3D<O>& o) const { return X * o.X + Y * o.Y + Z * o.Z; }
    template <typename O>
    Vector3D operator%(const Vector3D<O>& o) const {
        return Vector3D(Y * o.Z - Z * o.Y, Z * o.X - X * o.Z, X * o.Y - Y * o.X);
    }
    Vector3D operator*(T s) const { return Vector3D(X * s, Y * s, Z * s); }
    Vector3D operator/(T s) const { return Vector3D(X / s, Y / s, Z / s); }
    template <typename O>
    Vector3D& operator+=(const Vector3D<O>& o) { X += o.X; Y += o.Y; Z += o.Z; return *this; }
    template <typename O>
    Vector3D& operator-=(const Vector3D<O>& o) { X -= o.X; Y -= o.Y; Z -= o.Z; return *this; }
    Vector3D& operator*=(T s) { X *= s; Y *= s; Z *= s; return *this; }
    Vector3D& operator/=(T s) { X /= s; Y /= s; Z /= s; return *this; }

    double Length() const { return std::sqrt(static_cast<double>(X * X + Y * Y + Z * Z)); }
    Vector3D Normalize() const {
        const double l = Length();
        return l > 0 ? Vector3D(X / l, Y / l, Z / l) : *this;
    }

    friend double compute_cos_angle(const Vector3D& a, const Vector3D& b) {
        const double np = a.Length() * b.Length();
        return (np < 1.0e-13) ? 1.0 : (a * b) / np;
    }
    friend std::ostream& operator<<(std::ostream& os, const Vector3D& v) {
        return os << "(" << v.X << ", " << v.Y << ", " << v.Z << ")";
    }
};

template <typename T>
Vector3D<T> operator*(T s, const Vector3D<T>& v) { return v * s; }

template <typename T>
bool operator==(const Vector3D<T>& a, const Vector3D<T>& b) { return a.X == b.X && a.Y == b.Y && a.Z == b.Z; }

template <typename T>
bool operator<(const Vector3D<T>& a, const Vector3D<T>& b) { return std::tie(a.X, a.Y, a.Z) < std::tie(b.X, b.Y, b.Z); }
