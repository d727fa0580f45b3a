#include "mathx.h"

#include <stdexcept>

namespace pth {

M4 inverse(const M4& m)
{
    // Work on a row-major copy a[r][c]; cofactors by 2x2 sub-determinant pairs.
    float a[4][4];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            a[r][c] = m.at(c, r);

    const float s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const float s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const float s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const float s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const float s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const float s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];

    const float c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const float c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const float c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const float c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const float c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const float c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];

    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    const float inv_det = 1.0f / det;

    float b[4][4];
    b[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * inv_det;
    b[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * inv_det;
    b[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * inv_det;
    b[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * inv_det;

    b[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * inv_det;
    b[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * inv_det;
    b[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * inv_det;
    b[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * inv_det;

    b[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * inv_det;
    b[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * inv_det;
    b[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * inv_det;
    b[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * inv_det;

    b[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * inv_det;
    b[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * inv_det;
    b[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * inv_det;
    b[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * inv_det;

    M4 r;
    for (int row = 0; row < 4; row++)
        for (int col = 0; col < 4; col++)
            r.at(col, row) = b[row][col];
    return r;
}

M4 translation(V3 t)
{
    M4 m = M4::identity();
    m.at(3, 0) = t.x;
    m.at(3, 1) = t.y;
    m.at(3, 2) = t.z;
    return m;
}

M4 rotation(V3 u, float angle)
{
    // src/lalgebra.cpp:19-58: Rodrigues' rotation about a normalised axis.
    if (angle < 1e-5)
        return M4::identity();
    if (length(u) < 1e-5)
        throw std::runtime_error("Rotation axis was (close to) zero");
    u = normalize(u);
    const float c = cosf(angle), s = sinf(angle);
    M4 m = M4::identity();
    m.at(0, 0) = u.x * u.x + (1 - u.x * u.x) * c;
    m.at(0, 1) = u.y * u.x * (1 - c) + u.z * s;
    m.at(0, 2) = u.z * u.x * (1 - c) - u.y * s;
    m.at(1, 0) = u.x * u.y * (1 - c) - u.z * s;
    m.at(1, 1) = u.y * u.y + (1 - u.y * u.y) * c;
    m.at(1, 2) = u.z * u.y * (1 - c) + u.x * s;
    m.at(2, 0) = u.x * u.z * (1 - c) + u.y * s;
    m.at(2, 1) = u.y * u.z * (1 - c) - u.x * s;
    m.at(2, 2) = u.z * u.z + (1 - u.z * u.z) * c;
    return m;
}

M4 scaling(V3 s)
{
    M4 m = M4::identity();
    m.at(0, 0) = s.x;
    m.at(1, 1) = s.y;
    m.at(2, 2) = s.z;
    return m;
}

Quat quat_from_mat3(V3 c0, V3 c1, V3 c2)
{
    // Matrix element m[col][row]; the four candidates 4w^2-1, 4x^2-1, ... and the largest wins.
    const float m00 = c0.x, m01 = c0.y, m02 = c0.z;
    const float m10 = c1.x, m11 = c1.y, m12 = c1.z;
    const float m20 = c2.x, m21 = c2.y, m22 = c2.z;
    const float fw = m00 + m11 + m22;
    const float fx = m00 - m11 - m22;
    const float fy = m11 - m00 - m22;
    const float fz = m22 - m00 - m11;
    int big = 0;
    float fbig = fw;
    if (fx > fbig) { fbig = fx; big = 1; }
    if (fy > fbig) { fbig = fy; big = 2; }
    if (fz > fbig) { fbig = fz; big = 3; }
    const float bigv = std::sqrt(fbig + 1.0f) * 0.5f;
    const float mult = 0.25f / bigv;
    Quat q;
    switch (big) {
    case 0: q = {bigv, (m12 - m21) * mult, (m20 - m02) * mult, (m01 - m10) * mult}; break;
    case 1: q = {(m12 - m21) * mult, bigv, (m01 + m10) * mult, (m20 + m02) * mult}; break;
    case 2: q = {(m20 - m02) * mult, (m01 + m10) * mult, bigv, (m12 + m21) * mult}; break;
    default: q = {(m01 - m10) * mult, (m20 + m02) * mult, (m12 + m21) * mult, bigv}; break;
    }
    return q;
}

M4 mat4_from_quat(Quat q)
{
    const float qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    const float qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    const float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    M4 m = M4::identity();
    m.at(0, 0) = 1.0f - 2.0f * (qyy + qzz);
    m.at(0, 1) = 2.0f * (qxy + qwz);
    m.at(0, 2) = 2.0f * (qxz - qwy);
    m.at(1, 0) = 2.0f * (qxy - qwz);
    m.at(1, 1) = 1.0f - 2.0f * (qxx + qzz);
    m.at(1, 2) = 2.0f * (qyz + qwx);
    m.at(2, 0) = 2.0f * (qxz + qwy);
    m.at(2, 1) = 2.0f * (qyz - qwx);
    m.at(2, 2) = 1.0f - 2.0f * (qxx + qyy);
    return m;
}

M4 perspective_rh_zo(float fovy, float aspect, float znear, float zfar)
{
    const float t = tanf(fovy / 2.0f);
    M4 m;
    m.at(0, 0) = 1.0f / (aspect * t);
    m.at(1, 1) = 1.0f / t;
    m.at(2, 2) = zfar / (znear - zfar);
    m.at(2, 3) = -1.0f;
    m.at(3, 2) = -(zfar * znear) / (zfar - znear);
    return m;
}

}  // namespace pth
