// Small fp32 linear algebra for the host scene surface.
//
// Matrices are column-major (m[col][row]), the convention of the reference's glm types
// (src/lalgebra.h:6-15), so .scn transforms compose in the same order. Only what the scene
// surface needs: affine transforms, a 4x4 inverse, quaternion <-> rotation matrix and the
// perspective matrix.
#pragma once

#include <cmath>
#include <cstdint>

namespace pth {

struct V2 {
    float x = 0, y = 0;
};

struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};

inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// glm::normalize computes v * inversesqrt(dot(v, v)); 1/sqrt keeps the same rounding path.
inline V3 normalize(V3 a) { return a * (1.0f / std::sqrt(dot(a, a))); }

struct V4 {
    float x = 0, y = 0, z = 0, w = 0;
    V4() = default;
    V4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    float& operator[](int i) { return (&x)[i]; }
    float operator[](int i) const { return (&x)[i]; }
};

struct M4 {
    V4 c[4];  // columns
    static M4 identity()
    {
        M4 m;
        for (int i = 0; i < 4; i++)
            m.c[i][i] = 1.0f;
        return m;
    }
    float& at(int col, int row) { return c[col][row]; }
    float at(int col, int row) const { return c[col][row]; }
};

inline M4 operator*(const M4& a, const M4& b)
{
    M4 r;
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++)
                s += a.at(k, row) * b.at(col, k);
            r.at(col, row) = s;
        }
    return r;
}

inline V4 operator*(const M4& a, V4 v)
{
    V4 r;
    for (int row = 0; row < 4; row++)
        r[row] = a.at(0, row) * v.x + a.at(1, row) * v.y + a.at(2, row) * v.z + a.at(3, row) * v.w;
    return r;
}

inline M4 transpose(const M4& a)
{
    M4 r;
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++)
            r.at(col, row) = a.at(row, col);
    return r;
}

// General 4x4 inverse by cofactors (adjugate / determinant), fp32.
M4 inverse(const M4& m);

// Affine helpers with the semantics of src/lalgebra.cpp:5-58.
M4 translation(V3 t);
M4 rotation(V3 axis, float angle);  // identity when angle < 1e-5; throws on a ~zero axis
M4 scaling(V3 s);

struct Quat {
    float w = 1, x = 0, y = 0, z = 0;
};

// Rotation matrix (columns) -> unit quaternion (Shoemake's largest-component method) and
// back; together they give the camera rotation the reference stores as a quaternion.
Quat quat_from_mat3(V3 col0, V3 col1, V3 col2);
M4 mat4_from_quat(Quat q);

// Right-handed perspective with depth range [0, 1] (the reference builds with
// GLM_FORCE_DEPTH_ZERO_TO_ONE, Makefile:9), before the y flip of camera.cpp:24.
M4 perspective_rh_zo(float fovy, float aspect, float znear, float zfar);

inline float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

}  // namespace pth
