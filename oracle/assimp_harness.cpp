// Golden vectors for the glTF import's arithmetic, computed with Assimp 5.4.3's own inline
// math (the reference vendors it: include/assimp/matrix4x4.inl, matrix3x3.inl, vector3.inl,
// quaternion.inl; compiled here from where it lies under the reference checkout).
//
// Test infrastructure only: built by oracle/Makefile into oracle/_ref/aiharness; it prints JSON
// on stdout, which tests/golden/make_golden.py stores as tests/golden/assimp_math.json.
//
// The scene it describes (tests/test_import_steps.py writes it as a glTF file):
//   root R (no mesh) -> A (mesh 0), B (no mesh) -> C (mesh 1), D1 (mesh 2), D2 (mesh 2)
// meshes are single triangles with normals; materials of meshes 0 and 1 differ only in name.
// What the post-processing steps make of it, in the arithmetic of those headers:
//   * node transforms: identity * Translation * Rotation(quaternion) * Scaling (glTF TRS)
//   * OptimizeGraph: A, C, D1, D2 pulled up under the root with R * A, R * (B * C), R * D1,
//     R * D2; C joined into A: its triangle baked by inverse(R * A) * (R * (B * C)), its
//     normals by that matrix's inverse-transposed 3x3
//   * face normals ((v1 - v0) ^ (v2 - v0)).NormalizeSafe() for triangles without normals
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <assimp/vector3.h>
#include <assimp/matrix3x3.h>
#include <assimp/matrix4x4.h>
#include <assimp/quaternion.h>
#include <assimp/vector3.inl>
#include <assimp/matrix3x3.inl>
#include <assimp/matrix4x4.inl>
#include <assimp/quaternion.inl>

namespace {

uint32_t state = 0x9e3779b9u;
float uniform(float lo, float hi)  // xorshift32, enough for test inputs
{
    state ^= state << 13;
    state ^= state >> 17;
    state ^= state << 5;
    return lo + (hi - lo) * float(state >> 8) * (1.0f / 16777216.0f);
}

uint32_t bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

void print_list(const std::vector<float>& v)
{
    printf("[");
    for (size_t i = 0; i < v.size(); i++)
        printf("%s%u", i ? "," : "", bits(v[i]));
    printf("]");
}

std::vector<float> rows(const aiMatrix4x4& m)
{
    return {m.a1, m.a2, m.a3, m.a4, m.b1, m.b2, m.b3, m.b4, m.c1, m.c2, m.c3, m.c4, m.d1, m.d2, m.d3, m.d4};
}

struct Trs {
    float t[3], q[4], s[3];  // q = x, y, z, w (glTF order)
};

Trs random_trs(bool mirror)
{
    Trs n;
    for (float& x : n.t)
        x = uniform(-3.0f, 3.0f);
    float q[4], len2 = 0.0f;
    for (float& x : q) {
        x = uniform(-1.0f, 1.0f);
        len2 += x * x;
    }
    const float inv = 1.0f / std::sqrt(len2);
    for (int i = 0; i < 4; i++)
        n.q[i] = q[i] * inv;
    for (float& x : n.s)
        x = uniform(0.5f, 2.0f);
    if (mirror)
        n.s[1] = -n.s[1];
    return n;
}

aiMatrix4x4 node_matrix(const Trs& n)
{
    aiMatrix4x4 m, t, s;
    aiMatrix4x4::Translation(aiVector3D(n.t[0], n.t[1], n.t[2]), t);
    m = m * t;
    aiQuaternion rot;
    rot.x = n.q[0];
    rot.y = n.q[1];
    rot.z = n.q[2];
    rot.w = n.q[3];
    m = m * aiMatrix4x4(rot.GetMatrix());
    aiMatrix4x4::Scaling(aiVector3D(n.s[0], n.s[1], n.s[2]), s);
    m = m * s;
    return m;
}

struct Tri {
    aiVector3D p[3], n[3];
};

Tri random_tri()
{
    Tri t;
    for (int k = 0; k < 3; k++) {
        t.p[k] = aiVector3D(uniform(-1, 1), uniform(-1, 1), uniform(-1, 1));
        t.n[k] = aiVector3D(uniform(-1, 1), uniform(-1, 1), uniform(-1, 1)).Normalize();
    }
    return t;
}

void print_trs(const char* name, const Trs& n)
{
    printf("\"%s\": {\"t\": ", name);
    print_list({n.t[0], n.t[1], n.t[2]});
    printf(", \"q\": ");
    print_list({n.q[0], n.q[1], n.q[2], n.q[3]});
    printf(", \"s\": ");
    print_list({n.s[0], n.s[1], n.s[2]});
    printf("}");
}

void print_tri(const char* name, const Tri& t)
{
    std::vector<float> p, n;
    for (int k = 0; k < 3; k++) {
        p.insert(p.end(), {t.p[k].x, t.p[k].y, t.p[k].z});
        n.insert(n.end(), {t.n[k].x, t.n[k].y, t.n[k].z});
    }
    printf("\"%s\": {\"pos\": ", name);
    print_list(p);
    printf(", \"nrm\": ");
    print_list(n);
    printf("}");
}

}  // namespace

int main()
{
    const Trs R = random_trs(false), A = random_trs(false), B = random_trs(false), C = random_trs(false),
              D1 = random_trs(false), D2 = random_trs(true);
    const Tri t0 = random_tri(), t1 = random_tri(), t2 = random_tri();

    const aiMatrix4x4 mR = node_matrix(R), mA = node_matrix(A), mB = node_matrix(B), mC = node_matrix(C),
                      mD1 = node_matrix(D1), mD2 = node_matrix(D2);
    const aiMatrix4x4 wA = mR * mA, wC = mR * (mB * mC), wD1 = mR * mD1, wD2 = mR * mD2;
    aiMatrix4x4 inv = wA;
    inv.Inverse();
    const aiMatrix4x4 bake = inv * wC;
    const aiMatrix3x3 it = aiMatrix3x3(bake).Inverse().Transpose();

    std::vector<float> merged_pos, merged_nrm;
    for (int k = 0; k < 3; k++) {  // the joined mesh (C's triangle, baked) first, A's own last
        aiVector3D p = t1.p[k], n = t1.n[k];
        p *= bake;
        n *= it;
        merged_pos.insert(merged_pos.end(), {p.x, p.y, p.z});
        merged_nrm.insert(merged_nrm.end(), {n.x, n.y, n.z});
    }
    for (int k = 0; k < 3; k++) {
        merged_pos.insert(merged_pos.end(), {t0.p[k].x, t0.p[k].y, t0.p[k].z});
        merged_nrm.insert(merged_nrm.end(), {t0.n[k].x, t0.n[k].y, t0.n[k].z});
    }

    printf("{\"nodes\": {");
    print_trs("R", R);
    printf(", ");
    print_trs("A", A);
    printf(", ");
    print_trs("B", B);
    printf(", ");
    print_trs("C", C);
    printf(", ");
    print_trs("D1", D1);
    printf(", ");
    print_trs("D2", D2);
    printf("}, \"tris\": {");
    print_tri("t0", t0);
    printf(", ");
    print_tri("t1", t1);
    printf(", ");
    print_tri("t2", t2);
    printf("}, \"merged_pos\": ");
    print_list(merged_pos);
    printf(", \"merged_nrm\": ");
    print_list(merged_nrm);
    printf(", \"instance_A\": ");
    print_list(rows(wA));
    printf(", \"instance_D1\": ");
    print_list(rows(wD1));
    printf(", \"instance_D2\": ");
    print_list(rows(wD2));
    printf(", \"determinant_D2\": ");
    print_list({mD2.Determinant()});

    // face normals of random triangles (GenNormals for primitives without NORMAL)
    printf(", \"face_normals\": [");
    for (int i = 0; i < 16; i++) {
        Tri t = random_tri();
        if (i == 15)  // a degenerate one: zero length, left unnormalised
            t.p[2] = t.p[1];
        const aiVector3D n = ((t.p[1] - t.p[0]) ^ (t.p[2] - t.p[0])).NormalizeSafe();
        std::vector<float> p;
        for (int k = 0; k < 3; k++)
            p.insert(p.end(), {t.p[k].x, t.p[k].y, t.p[k].z});
        printf("%s{\"pos\": ", i ? ", " : "");
        print_list(p);
        printf(", \"normal\": ");
        print_list({n.x, n.y, n.z});
        printf("}");
    }
    printf("]}\n");
    return 0;
}
