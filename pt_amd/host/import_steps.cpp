#include "import_steps.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <stdexcept>

namespace pth {
namespace steps {

// ---------------------------------------------------------------------------------------
// Arithmetic. Each result must carry the bits Assimp's inline code gives (the reference
// bakes vertices and instance transforms with it), so every expression below keeps the
// operand order and grouping of include/assimp/matrix4x4.inl, matrix3x3.inl, vector3.inl
// and quaternion.inl, and this file is compiled without floating-point contraction.

namespace {
struct Rows4 {  // a1..d4 names of an aiMatrix4x4
    float a1, a2, a3, a4, b1, b2, b3, b4, c1, c2, c3, c4, d1, d2, d3, d4;
    explicit Rows4(const AiMat4& m)
        : a1(m.m[0][0]), a2(m.m[0][1]), a3(m.m[0][2]), a4(m.m[0][3]),
          b1(m.m[1][0]), b2(m.m[1][1]), b3(m.m[1][2]), b4(m.m[1][3]),
          c1(m.m[2][0]), c2(m.m[2][1]), c3(m.m[2][2]), c4(m.m[2][3]),
          d1(m.m[3][0]), d2(m.m[3][1]), d3(m.m[3][2]), d4(m.m[3][3]) {}
};
}  // namespace

AiMat4 ai_identity()
{
    AiMat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            r.m[i][j] = i == j ? 1.0f : 0.0f;
    return r;
}

AiMat4 ai_mul(const AiMat4& a, const AiMat4& b)
{
    // element (r, c) = b(0,c) a(r,0) + b(1,c) a(r,1) + b(2,c) a(r,2) + b(3,c) a(r,3), summed left to right
    AiMat4 r;
    for (int row = 0; row < 4; row++)
        for (int col = 0; col < 4; col++)
            r.m[row][col] = b.m[0][col] * a.m[row][0] + b.m[1][col] * a.m[row][1] + b.m[2][col] * a.m[row][2] +
                            b.m[3][col] * a.m[row][3];
    return r;
}

float ai_determinant(const AiMat4& m)
{
    const Rows4 e(m);
    return e.a1 * e.b2 * e.c3 * e.d4 - e.a1 * e.b2 * e.c4 * e.d3 + e.a1 * e.b3 * e.c4 * e.d2 -
           e.a1 * e.b3 * e.c2 * e.d4 + e.a1 * e.b4 * e.c2 * e.d3 - e.a1 * e.b4 * e.c3 * e.d2 -
           e.a2 * e.b3 * e.c4 * e.d1 + e.a2 * e.b3 * e.c1 * e.d4 - e.a2 * e.b4 * e.c1 * e.d3 +
           e.a2 * e.b4 * e.c3 * e.d1 - e.a2 * e.b1 * e.c3 * e.d4 + e.a2 * e.b1 * e.c4 * e.d3 +
           e.a3 * e.b4 * e.c1 * e.d2 - e.a3 * e.b4 * e.c2 * e.d1 + e.a3 * e.b1 * e.c2 * e.d4 -
           e.a3 * e.b1 * e.c4 * e.d2 + e.a3 * e.b2 * e.c4 * e.d1 - e.a3 * e.b2 * e.c1 * e.d4 -
           e.a4 * e.b1 * e.c2 * e.d3 + e.a4 * e.b1 * e.c3 * e.d2 - e.a4 * e.b2 * e.c3 * e.d1 +
           e.a4 * e.b2 * e.c1 * e.d3 - e.a4 * e.b3 * e.c1 * e.d2 + e.a4 * e.b3 * e.c2 * e.d1;
}

AiMat4 ai_inverse(const AiMat4& m)
{
    const float det = ai_determinant(m);
    AiMat4 r;
    if (det == 0.0f) {
        for (auto& row : r.m)
            for (float& x : row)
                x = std::numeric_limits<float>::quiet_NaN();
        return r;
    }
    const float k = 1.0f / det;
    const Rows4 e(m);
    // cofactors by expansion over the third and fourth rows' 2x2 minors
    r.m[0][0] = k * (e.b2 * (e.c3 * e.d4 - e.c4 * e.d3) + e.b3 * (e.c4 * e.d2 - e.c2 * e.d4) + e.b4 * (e.c2 * e.d3 - e.c3 * e.d2));
    r.m[0][1] = -k * (e.a2 * (e.c3 * e.d4 - e.c4 * e.d3) + e.a3 * (e.c4 * e.d2 - e.c2 * e.d4) + e.a4 * (e.c2 * e.d3 - e.c3 * e.d2));
    r.m[0][2] = k * (e.a2 * (e.b3 * e.d4 - e.b4 * e.d3) + e.a3 * (e.b4 * e.d2 - e.b2 * e.d4) + e.a4 * (e.b2 * e.d3 - e.b3 * e.d2));
    r.m[0][3] = -k * (e.a2 * (e.b3 * e.c4 - e.b4 * e.c3) + e.a3 * (e.b4 * e.c2 - e.b2 * e.c4) + e.a4 * (e.b2 * e.c3 - e.b3 * e.c2));
    r.m[1][0] = -k * (e.b1 * (e.c3 * e.d4 - e.c4 * e.d3) + e.b3 * (e.c4 * e.d1 - e.c1 * e.d4) + e.b4 * (e.c1 * e.d3 - e.c3 * e.d1));
    r.m[1][1] = k * (e.a1 * (e.c3 * e.d4 - e.c4 * e.d3) + e.a3 * (e.c4 * e.d1 - e.c1 * e.d4) + e.a4 * (e.c1 * e.d3 - e.c3 * e.d1));
    r.m[1][2] = -k * (e.a1 * (e.b3 * e.d4 - e.b4 * e.d3) + e.a3 * (e.b4 * e.d1 - e.b1 * e.d4) + e.a4 * (e.b1 * e.d3 - e.b3 * e.d1));
    r.m[1][3] = k * (e.a1 * (e.b3 * e.c4 - e.b4 * e.c3) + e.a3 * (e.b4 * e.c1 - e.b1 * e.c4) + e.a4 * (e.b1 * e.c3 - e.b3 * e.c1));
    r.m[2][0] = k * (e.b1 * (e.c2 * e.d4 - e.c4 * e.d2) + e.b2 * (e.c4 * e.d1 - e.c1 * e.d4) + e.b4 * (e.c1 * e.d2 - e.c2 * e.d1));
    r.m[2][1] = -k * (e.a1 * (e.c2 * e.d4 - e.c4 * e.d2) + e.a2 * (e.c4 * e.d1 - e.c1 * e.d4) + e.a4 * (e.c1 * e.d2 - e.c2 * e.d1));
    r.m[2][2] = k * (e.a1 * (e.b2 * e.d4 - e.b4 * e.d2) + e.a2 * (e.b4 * e.d1 - e.b1 * e.d4) + e.a4 * (e.b1 * e.d2 - e.b2 * e.d1));
    r.m[2][3] = -k * (e.a1 * (e.b2 * e.c4 - e.b4 * e.c2) + e.a2 * (e.b4 * e.c1 - e.b1 * e.c4) + e.a4 * (e.b1 * e.c2 - e.b2 * e.c1));
    r.m[3][0] = -k * (e.b1 * (e.c2 * e.d3 - e.c3 * e.d2) + e.b2 * (e.c3 * e.d1 - e.c1 * e.d3) + e.b3 * (e.c1 * e.d2 - e.c2 * e.d1));
    r.m[3][1] = k * (e.a1 * (e.c2 * e.d3 - e.c3 * e.d2) + e.a2 * (e.c3 * e.d1 - e.c1 * e.d3) + e.a3 * (e.c1 * e.d2 - e.c2 * e.d1));
    r.m[3][2] = -k * (e.a1 * (e.b2 * e.d3 - e.b3 * e.d2) + e.a2 * (e.b3 * e.d1 - e.b1 * e.d3) + e.a3 * (e.b1 * e.d2 - e.b2 * e.d1));
    r.m[3][3] = k * (e.a1 * (e.b2 * e.c3 - e.b3 * e.c2) + e.a2 * (e.b3 * e.c1 - e.b1 * e.c3) + e.a3 * (e.b1 * e.c2 - e.b2 * e.c1));
    return r;
}

AiVec ai_transform_point(const AiMat4& m, AiVec v)
{
    AiVec r;
    r.x = m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z + m.m[0][3];
    r.y = m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z + m.m[1][3];
    r.z = m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z + m.m[2][3];
    return r;
}

AiMat3 ai_normal_matrix(const AiMat4& m4)
{
    const float a1 = m4.m[0][0], a2 = m4.m[0][1], a3 = m4.m[0][2];
    const float b1 = m4.m[1][0], b2 = m4.m[1][1], b3 = m4.m[1][2];
    const float c1 = m4.m[2][0], c2 = m4.m[2][1], c3 = m4.m[2][2];
    AiMat3 inv;
    const float det = a1 * b2 * c3 - a1 * b3 * c2 + a2 * b3 * c1 - a2 * b1 * c3 + a3 * b1 * c2 - a3 * b2 * c1;
    if (det == 0.0f) {
        for (auto& row : inv.m)
            for (float& x : row)
                x = std::numeric_limits<float>::quiet_NaN();
        return inv;
    }
    const float k = 1.0f / det;
    inv.m[0][0] = k * (b2 * c3 - b3 * c2);
    inv.m[0][1] = -k * (a2 * c3 - a3 * c2);
    inv.m[0][2] = k * (a2 * b3 - a3 * b2);
    inv.m[1][0] = -k * (b1 * c3 - b3 * c1);
    inv.m[1][1] = k * (a1 * c3 - a3 * c1);
    inv.m[1][2] = -k * (a1 * b3 - a3 * b1);
    inv.m[2][0] = k * (b1 * c2 - b2 * c1);
    inv.m[2][1] = -k * (a1 * c2 - a2 * c1);
    inv.m[2][2] = k * (a1 * b2 - a2 * b1);
    AiMat3 t;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t.m[i][j] = inv.m[j][i];
    return t;
}

AiVec ai_transform_vector(const AiMat3& m, AiVec v)
{
    AiVec r;
    r.x = m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z;
    r.y = m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z;
    r.z = m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z;
    return r;
}

AiMat4 ai_rotation(float x, float y, float z, float w)
{
    AiMat4 r = ai_identity();
    r.m[0][0] = 1.0f - 2.0f * (y * y + z * z);
    r.m[0][1] = 2.0f * (x * y - z * w);
    r.m[0][2] = 2.0f * (x * z + y * w);
    r.m[1][0] = 2.0f * (x * y + z * w);
    r.m[1][1] = 1.0f - 2.0f * (x * x + z * z);
    r.m[1][2] = 2.0f * (y * z - x * w);
    r.m[2][0] = 2.0f * (x * z - y * w);
    r.m[2][1] = 2.0f * (y * z + x * w);
    r.m[2][2] = 1.0f - 2.0f * (x * x + y * y);
    return r;
}

AiMat4 ai_translation(float x, float y, float z)
{
    AiMat4 r = ai_identity();
    r.m[0][3] = x;
    r.m[1][3] = y;
    r.m[2][3] = z;
    return r;
}

AiMat4 ai_scaling(float x, float y, float z)
{
    AiMat4 r = ai_identity();
    r.m[0][0] = x;
    r.m[1][1] = y;
    r.m[2][2] = z;
    return r;
}

AiVec ai_face_normal(AiVec a, AiVec b, AiVec c)
{
    const AiVec e1{b.x - a.x, b.y - a.y, b.z - a.z};
    const AiVec e2{c.x - a.x, c.y - a.y, c.z - a.z};
    AiVec n{e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
    const float len = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    if (len > 0.0f) {  // scaled by the reciprocal of the length, not divided
        const float inv = 1.0f / len;
        n.x *= inv;
        n.y *= inv;
        n.z *= inv;
    }
    return n;
}

// ---------------------------------------------------------------------------------------
// Helpers over the model

namespace {

bool vec_ne(const AiVec& a, const AiVec& b) { return a.x != b.x || a.y != b.y || a.z != b.z; }
bool vec_eq(const AiVec& a, const AiVec& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

float sq_dist(const AiVec& a, const AiVec& b)
{
    const AiVec d{a.x - b.x, a.y - b.y, a.z - b.z};
    return d.x * d.x + d.y * d.y + d.z * d.z;
}

float colour_dist(const std::array<float, 4>& a, const std::array<float, 4>& b)
{
    const float r = a[0] - b[0], g = a[1] - b[1], bl = a[2] - b[2], al = a[3] - b[3];
    return r * r + g * g + bl * bl + al * al;
}

// Node references per mesh over the graph below the root.
std::vector<uint32_t> reference_counts(const Model& m)
{
    std::vector<uint32_t> refs(m.meshes.size(), 0);
    std::function<void(uint32_t, int)> walk = [&](uint32_t n, int depth) {
        if (depth > 4096)
            throw std::runtime_error("glTF node hierarchy too deep (cycle?)");
        for (uint32_t mi : m.nodes[n].meshes)
            refs[mi]++;
        for (uint32_t c : m.nodes[n].children)
            walk(c, depth + 1);
    };
    walk(m.root, 0);
    return refs;
}

void for_each_node(Model& m, const std::function<void(Node&)>& fn)
{
    std::function<void(uint32_t)> walk = [&](uint32_t n) {
        fn(m.nodes[n]);
        const std::vector<uint32_t> children = m.nodes[n].children;
        for (uint32_t c : children)
            walk(c);
    };
    walk(m.root);
}

void flip_winding(Mesh& mesh)
{
    for (Face& f : mesh.faces)
        for (uint32_t b = 0; b < f.n / 2; b++)
            std::swap(f.i[b], f.i[f.n - 1 - b]);
}

bool special(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return (u & 0x7f800000u) == 0x7f800000u;
}

// ValidateArrayContents<aiVector3D>: the reason the array is invalid, or nullptr.
const char* invalid_array(const std::vector<AiVec>& arr, const std::vector<bool>& dirty, bool may_be_identical,
                          bool may_be_zero)
{
    bool differs = false;
    size_t counted = 0;
    for (size_t i = 0; i < arr.size(); i++) {
        if (dirty[i])
            continue;
        counted++;
        const AiVec& v = arr[i];
        if (special(v.x) || special(v.y) || special(v.z))
            return "INF/NAN was found in a vector component";
        if (!may_be_zero && !v.x && !v.y && !v.z)
            return "Found zero-length vector";
        if (i && vec_ne(v, arr[i - 1]))
            differs = true;
    }
    if (counted > 1 && !differs && !may_be_identical)
        return "All vectors are identical";
    return nullptr;
}

// Removes the meshes whose flag is set and renumbers every node's references.
void drop_meshes(Model& m, const std::vector<bool>& drop)
{
    std::vector<uint32_t> remap(m.meshes.size(), UINT32_MAX);
    std::vector<Mesh> kept;
    for (size_t i = 0; i < m.meshes.size(); i++)
        if (!drop[i]) {
            remap[i] = uint32_t(kept.size());
            kept.push_back(std::move(m.meshes[i]));
        }
    m.meshes.swap(kept);
    for (Node& n : m.nodes) {
        std::vector<uint32_t> refs;
        for (uint32_t r : n.meshes)
            if (remap[r] != UINT32_MAX)
                refs.push_back(remap[r]);
        n.meshes.swap(refs);
    }
}

}  // namespace

uint32_t vertex_format(const Mesh& mesh)
{
    uint32_t f = 1;
    if (!mesh.nrm.empty())
        f |= 0x2;
    if (!mesh.tan.empty() && !mesh.bitan.empty())
        f |= 0x4;
    for (size_t p = 0; p < mesh.uv.size() && p < 8; p++) {
        f |= 0x100u << p;
        if (mesh.uv_comps[p] == 3)
            f |= 0x10000u << p;
    }
    for (size_t p = 0; p < mesh.col.size() && p < 8; p++)
        f |= 0x1000000u << p;
    return f;
}

// ---------------------------------------------------------------------------------------
// RemoveRedundantMaterials: materials no mesh uses are dropped; a material whose properties
// equal an earlier used one's is replaced by it. Survivors keep their order.

void remove_redundant_materials(Model& m)
{
    const size_t n = m.material_keys.size();
    if (n == 0)
        return;
    std::vector<bool> used(n, false);
    for (const Mesh& mesh : m.meshes)
        used.at(mesh.material) = true;
    std::vector<uint32_t> slot(n, 0);
    uint32_t kept = 0;
    for (size_t i = 0; i < n; i++) {
        if (!used[i])
            continue;
        bool duplicate = false;
        for (size_t a = 0; a < i; a++)
            if (used[a] && m.material_keys[a] == m.material_keys[i]) {
                slot[i] = slot[a];
                duplicate = true;
                break;
            }
        if (!duplicate)
            slot[i] = kept++;
    }
    if (kept == n)
        return;
    if (kept == 0)
        throw std::runtime_error("No materials remaining");
    std::vector<std::string> keys(kept);
    std::vector<int> sources(kept, -2);
    for (size_t p = 0; p < n; p++)
        if (used[p] && sources[slot[p]] == -2) {
            keys[slot[p]] = m.material_keys[p];
            sources[slot[p]] = m.material_source[p];
        }
    for (Mesh& mesh : m.meshes)
        mesh.material = slot[mesh.material];
    m.material_keys.swap(keys);
    m.material_source.swap(sources);
}

// ---------------------------------------------------------------------------------------
// FindInstances: a mesh equal to an earlier one (same vertex format, counts, material and
// primitive types; positions within 1e-4 of the mesh's extent, normals and tangents alike,
// texture coordinates and colours within 1e-3 squared; the same face structure) is replaced by
// the nearest such earlier mesh in every node.

void find_instances(Model& m)
{
    const size_t n = m.meshes.size();
    if (n == 0)
        return;
    std::vector<uint32_t> remap(n, 0);
    std::vector<bool> gone(n, false);
    uint32_t out = 0;
    for (size_t i = 0; i < n; i++) {
        const Mesh& inst = m.meshes[i];
        AiVec lo{1e10f, 1e10f, 1e10f}, hi{-1e10f, -1e10f, -1e10f};
        for (const AiVec& p : inst.pos) {
            lo = AiVec{std::min(p.x, lo.x), std::min(p.y, lo.y), std::min(p.z, lo.z)};
            hi = AiVec{std::max(p.x, hi.x), std::max(p.y, hi.y), std::max(p.z, hi.z)};
        }
        const AiVec ext{hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
        float eps = std::sqrt(ext.x * ext.x + ext.y * ext.y + ext.z * ext.z) * 1e-4f;
        eps *= eps;
        const float uv_eps = 10e-4f;
        auto same = [](const std::vector<AiVec>& a, const std::vector<AiVec>& b, float e) {
            for (size_t k = 0; k < a.size(); k++)
                if (sq_dist(a[k], b[k]) >= e)
                    return false;
            return true;
        };
        for (long a = long(i) - 1; a >= 0; a--) {
            if (gone[size_t(a)])
                continue;
            const Mesh& orig = m.meshes[size_t(a)];
            if (vertex_format(orig) != vertex_format(inst) || orig.bones || inst.bones ||
                orig.faces.size() != inst.faces.size() || orig.pos.size() != inst.pos.size() ||
                orig.material != inst.material || orig.ptypes != inst.ptypes)
                continue;
            if (!same(orig.pos, inst.pos, eps))
                continue;
            if (!orig.nrm.empty() && !same(orig.nrm, inst.nrm, eps))
                continue;
            if (!orig.tan.empty() && (!same(orig.tan, inst.tan, eps) || !same(orig.bitan, inst.bitan, eps)))
                continue;
            bool uv_ok = true;
            for (size_t c = 0; c < orig.uv.size() && uv_ok; c++)
                uv_ok = same(orig.uv[c], inst.uv[c], uv_eps);
            if (!uv_ok)
                continue;
            bool col_ok = true;
            for (size_t c = 0; c < orig.col.size() && col_ok; c++)
                for (size_t k = 0; k < orig.col[c].size() && col_ok; k++)
                    col_ok = colour_dist(orig.col[c][k], inst.col[c][k]) < uv_eps;
            if (!col_ok)
                continue;
            // face structure: every corner of the instance's face t must be a vertex whose last
            // face in the original is t
            std::vector<uint32_t> last_face(orig.pos.size(), 0);
            for (uint32_t t = 0; t < orig.faces.size(); t++)
                for (uint32_t k = 0; k < orig.faces[t].n; k++)
                    last_face[orig.faces[t].i[k]] = t;
            bool faces_ok = true;
            for (uint32_t t = 0; t < inst.faces.size() && faces_ok; t++)
                for (uint32_t k = 0; k < inst.faces[t].n && faces_ok; k++)
                    faces_ok = last_face[inst.faces[t].i[k]] == t;
            if (!faces_ok)
                continue;
            remap[i] = remap[size_t(a)];
            gone[i] = true;
            break;
        }
        if (!gone[i])
            remap[i] = out++;
    }
    if (out == n)
        return;
    std::vector<Mesh> kept;
    for (size_t i = 0; i < n; i++)
        if (!gone[i])
            kept.push_back(std::move(m.meshes[i]));
    m.meshes.swap(kept);
    for (Node& node : m.nodes)
        for (uint32_t& r : node.meshes)
            r = remap[r];
}

// ---------------------------------------------------------------------------------------
// OptimizeGraph. Below a locked parent (the graph gets a locked dummy root) every unlocked
// node is replaced by its descendants, each carrying the product of the transforms on the way
// (the parent's on the left), descendants first, the node itself after them if it has meshes
// or locked children left. Among a locked node's children, the leaves without instanced
// meshes are joined into the first of them: their meshes are baked into its frame (positions
// by inverse(first) x own transform, normals and tangents by the inverse transpose of its
// 3x3 part, face winding reversed under a mirroring transform) and appended to its mesh list.

namespace {
const char* const kDummyRoot = "$Reserved_And_Evil";

struct GraphOptimizer {
    Model& m;
    std::vector<uint32_t> refs;
    unsigned merged = 0;

    bool locked(uint32_t n) const { return m.locked.count(m.nodes[n].name) != 0; }

    void bake(Mesh& mesh, const AiMat4& t)
    {
        if (ai_determinant(t) < 0.0f)
            flip_winding(mesh);
        const AiMat3 it = ai_normal_matrix(t);
        for (size_t a = 0; a < mesh.pos.size(); a++) {
            mesh.pos[a] = ai_transform_point(t, mesh.pos[a]);
            if (!mesh.nrm.empty())
                mesh.nrm[a] = ai_transform_vector(it, mesh.nrm[a]);
            if (!mesh.tan.empty() && !mesh.bitan.empty()) {
                mesh.tan[a] = ai_transform_vector(it, mesh.tan[a]);
                mesh.bitan[a] = ai_transform_vector(it, mesh.bitan[a]);
            }
        }
    }

    void collect(uint32_t nd, std::vector<uint32_t>& out, int depth)
    {
        if (depth > 4096)
            throw std::runtime_error("glTF node hierarchy too deep (cycle?)");
        std::vector<uint32_t> kids;
        const std::vector<uint32_t> children = m.nodes[nd].children;
        for (uint32_t c : children)
            collect(c, kids, depth + 1);
        if (!locked(nd)) {
            std::vector<uint32_t> stay;
            for (uint32_t k : kids) {
                if (!locked(k)) {
                    m.nodes[k].transform = ai_mul(m.nodes[nd].transform, m.nodes[k].transform);
                    out.push_back(k);
                } else {
                    stay.push_back(k);
                }
            }
            kids.swap(stay);
            if (m.nodes[nd].meshes.empty() && kids.empty())
                return;  // the node goes
            out.push_back(nd);
        } else {
            out.push_back(nd);
            long master = -1;
            AiMat4 inv{};
            std::vector<uint32_t> join, stay;
            for (uint32_t k : kids) {
                Node& c = m.nodes[k];
                if (c.children.empty() && !locked(k) &&
                    std::none_of(c.meshes.begin(), c.meshes.end(), [&](uint32_t mi) { return refs[mi] > 1; })) {
                    if (master < 0) {
                        master = long(k);
                        inv = ai_inverse(c.transform);
                    } else {
                        c.transform = ai_mul(inv, c.transform);
                        join.push_back(k);
                        continue;
                    }
                }
                stay.push_back(k);
            }
            kids.swap(stay);
            if (master >= 0 && !join.empty()) {
                Node& first = m.nodes[size_t(master)];
                first.name = "$MergedNode_" + std::to_string(merged++);
                for (uint32_t j : join)
                    for (uint32_t mi : m.nodes[j].meshes) {
                        bake(m.meshes[mi], m.nodes[j].transform);
                        m.nodes[size_t(master)].meshes.push_back(mi);
                    }
            }
        }
        m.nodes[nd].children = kids;
    }
};
}  // namespace

void optimize_graph(Model& m)
{
    GraphOptimizer g{m, reference_counts(m)};
    m.locked.insert(kDummyRoot);
    Node dummy;
    dummy.name = kDummyRoot;
    dummy.transform = ai_identity();
    dummy.children = {m.root};
    const std::string root_name = m.nodes[m.root].name;
    m.nodes.push_back(dummy);
    const uint32_t d = uint32_t(m.nodes.size() - 1);
    std::vector<uint32_t> top;
    g.collect(d, top, 0);
    const std::vector<uint32_t>& kids = m.nodes[d].children;
    if (kids.empty())
        throw std::runtime_error("After optimizing the scene graph, no data remains");
    if (kids.size() > 1) {
        m.nodes[d].name = root_name;
        m.root = d;
    } else {
        m.root = kids[0];
    }
    m.locked.erase(kDummyRoot);
}

// ---------------------------------------------------------------------------------------
// OptimizeMeshes: meshes used by several nodes come first, in their order. Then, node by node
// (depth first, a node's meshes before its children), each singly used mesh absorbs the later
// meshes of the same node that have its vertex format and material: the later list is scanned
// in order, an absorbed entry replaced by the node's last entry and scanned again; the merged
// mesh holds the absorbed meshes in that order and then the absorbing mesh.

namespace {
Mesh merge_meshes(const std::vector<const Mesh*>& list)
{
    Mesh out;
    const Mesh& first = *list.front();
    out.material = first.material;
    size_t nv = 0;
    for (size_t k = 0; k < list.size(); k++) {
        out.name += list[k]->name;
        if (k + 1 < list.size())
            out.name += ".";
        nv += list[k]->pos.size();
        out.ptypes |= list[k]->ptypes;
        out.bones = out.bones || list[k]->bones;
    }
    auto cat = [&](std::vector<AiVec> Mesh::*field, bool wanted) {
        if (!wanted)
            return;
        std::vector<AiVec>& dst = out.*field;
        dst.reserve(nv);
        for (const Mesh* mesh : list) {
            const std::vector<AiVec>& src = mesh->*field;
            if (src.empty())
                dst.resize(dst.size() + mesh->pos.size());
            else
                dst.insert(dst.end(), src.begin(), src.end());
        }
    };
    cat(&Mesh::pos, true);
    cat(&Mesh::nrm, !first.nrm.empty());
    cat(&Mesh::tan, !first.tan.empty() && !first.bitan.empty());
    cat(&Mesh::bitan, !first.tan.empty() && !first.bitan.empty());
    out.uv.resize(first.uv.size());
    out.uv_comps = first.uv_comps;
    for (size_t c = 0; c < first.uv.size(); c++)
        for (const Mesh* mesh : list) {
            if (c < mesh->uv.size())
                out.uv[c].insert(out.uv[c].end(), mesh->uv[c].begin(), mesh->uv[c].end());
            else
                out.uv[c].resize(out.uv[c].size() + mesh->pos.size());
        }
    out.col.resize(first.col.size());
    for (size_t c = 0; c < first.col.size(); c++)
        for (const Mesh* mesh : list) {
            if (c < mesh->col.size())
                out.col[c].insert(out.col[c].end(), mesh->col[c].begin(), mesh->col[c].end());
            else
                out.col[c].resize(out.col[c].size() + mesh->pos.size(), std::array<float, 4>{0, 0, 0, 0});
        }
    uint32_t offset = 0;
    for (const Mesh* mesh : list) {
        for (Face f : mesh->faces) {
            for (uint32_t k = 0; k < f.n; k++)
                f.i[k] += offset;
            out.faces.push_back(f);
        }
        offset += uint32_t(mesh->pos.size());
    }
    return out;
}
}  // namespace

void optimize_meshes(Model& m)
{
    if (m.meshes.size() <= 1)
        return;
    const std::vector<uint32_t> uses = reference_counts(m);
    std::vector<uint32_t> fmt(m.meshes.size()), out_id(m.meshes.size(), UINT32_MAX);
    std::vector<Mesh> output;
    for (size_t i = 0; i < m.meshes.size(); i++) {
        fmt[i] = vertex_format(m.meshes[i]);
        if (uses[i] > 1) {
            out_id[i] = uint32_t(output.size());
            output.push_back(m.meshes[i]);
        }
    }
    auto joinable = [&](uint32_t a, uint32_t b) {
        const Mesh& x = m.meshes[a];
        const Mesh& y = m.meshes[b];
        return fmt[a] == fmt[b] && x.material == y.material && !x.bones && !y.bones;
    };
    std::function<void(uint32_t)> visit = [&](uint32_t nd) {
        std::vector<uint32_t>& list = m.nodes[nd].meshes;
        for (size_t i = 0; i < list.size(); i++) {
            const uint32_t im = list[i];
            if (uses[im] > 1) {
                list[i] = out_id[im];
                continue;
            }
            std::vector<const Mesh*> merge;
            for (size_t a = i + 1; a < list.size(); a++) {
                const uint32_t am = list[a];
                if (uses[am] == 1 && joinable(im, am)) {
                    merge.push_back(&m.meshes[am]);
                    list[a] = list.back();
                    list.pop_back();
                    a--;
                }
            }
            if (!merge.empty()) {
                merge.push_back(&m.meshes[im]);
                output.push_back(merge_meshes(merge));
            } else {
                output.push_back(m.meshes[im]);
            }
            list[i] = uint32_t(output.size() - 1);
        }
        const std::vector<uint32_t> children = m.nodes[nd].children;
        for (uint32_t c : children)
            visit(c);
    };
    visit(m.root);
    if (output.empty())
        throw std::runtime_error("OptimizeMeshes: No meshes remaining");
    m.meshes.swap(output);
}

// ---------------------------------------------------------------------------------------
// FindDegenerates (degenerate faces kept, PP_FD_REMOVE off by default, include/assimp/
// config.h:251-254): a corner whose position equals (==) an earlier corner's of the same face is
// taken out, so a triangle with coincident corners becomes a line or a point; the mesh's
// primitive types are recounted.

void find_degenerates(Model& m)
{
    for (Mesh& mesh : m.meshes) {
        mesh.ptypes = 0;
        for (Face& f : mesh.faces) {
            for (uint32_t i = 0; i < f.n; i++) {
                uint32_t limit = f.n;
                for (uint32_t t = i + 1; t < limit; t++) {
                    if (vec_eq(mesh.pos[f.i[i]], mesh.pos[f.i[t]])) {
                        f.n--;
                        limit--;
                        for (uint32_t k = t; k < f.n; k++)
                            f.i[k] = f.i[k + 1];
                        f.i[f.n] = 0xdeadbeefu;
                        t--;
                    }
                }
            }
            mesh.ptypes |= f.n == 1 ? PRIM_POINT : f.n == 2 ? PRIM_LINE : f.n == 3 ? PRIM_TRIANGLE : PRIM_POLYGON;
        }
    }
}

// ---------------------------------------------------------------------------------------
// FindInvalidData, over the vertices some face uses: positions with an Inf/NaN or all equal
// delete the mesh; a texture-coordinate set with an Inf/NaN or all equal is deleted with the
// sets after it; normals with an Inf/NaN or a zero vector are deleted (GenNormals then makes
// face normals), as are tangents/bitangents with an Inf/NaN or all equal. Meshes of points and
// lines only keep their normals unchecked; in mixed meshes the corners of points and lines are
// left out of the normal checks.

void find_invalid_data(Model& m)
{
    std::vector<bool> drop(m.meshes.size(), false);
    for (size_t mi = 0; mi < m.meshes.size(); mi++) {
        Mesh& mesh = m.meshes[mi];
        const size_t nv = mesh.pos.size();
        std::vector<bool> dirty(nv, !mesh.faces.empty());
        for (const Face& f : mesh.faces)
            for (uint32_t k = 0; k < f.n; k++)
                dirty[f.i[k]] = false;
        if (invalid_array(mesh.pos, dirty, false, true)) {
            drop[mi] = true;
            continue;
        }
        for (size_t c = 0; c < mesh.uv.size(); c++)
            if (invalid_array(mesh.uv[c], dirty, false, true)) {
                mesh.uv.resize(c);
                mesh.uv_comps.resize(c);
                break;
            }
        if (mesh.nrm.empty() && mesh.tan.empty())
            continue;
        if (mesh.ptypes & (PRIM_POINT | PRIM_LINE)) {
            if (!(mesh.ptypes & (PRIM_TRIANGLE | PRIM_POLYGON)))
                continue;
            for (const Face& f : mesh.faces)
                if (f.n < 3) {
                    dirty[f.i[0]] = true;
                    if (f.n == 2)
                        dirty[f.i[1]] = true;
                }
        }
        if (!mesh.nrm.empty() && invalid_array(mesh.nrm, dirty, true, false))
            mesh.nrm.clear();
        if (!mesh.tan.empty() && invalid_array(mesh.tan, dirty, false, true)) {
            mesh.tan.clear();
            mesh.bitan.clear();
        }
        if (!mesh.bitan.empty() && invalid_array(mesh.bitan, dirty, false, true)) {
            mesh.bitan.clear();
            mesh.tan.clear();
        }
    }
    if (std::find(drop.begin(), drop.end(), true) == drop.end())
        return;
    drop_meshes(m, drop);
    if (m.meshes.empty())
        throw std::runtime_error("No meshes remaining");
}

// ---------------------------------------------------------------------------------------
// GenNormals (face normals) for meshes without normals that hold triangles: every corner of a
// face gets ((v1 - v0) x (v_last - v0)) normalised; corners of points and lines get NaN.

void gen_face_normals(Model& m)
{
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    for (Mesh& mesh : m.meshes) {
        if (!mesh.nrm.empty() || !(mesh.ptypes & (PRIM_TRIANGLE | PRIM_POLYGON)))
            continue;
        mesh.nrm.assign(mesh.pos.size(), AiVec{});
        for (const Face& f : mesh.faces) {
            if (f.n < 3) {
                for (uint32_t k = 0; k < f.n; k++)
                    mesh.nrm[f.i[k]] = AiVec{qnan, qnan, qnan};
                continue;
            }
            const AiVec n = ai_face_normal(mesh.pos[f.i[0]], mesh.pos[f.i[1]], mesh.pos[f.i[f.n - 1]]);
            for (uint32_t k = 0; k < f.n; k++)
                mesh.nrm[f.i[k]] = n;
        }
    }
}

// ---------------------------------------------------------------------------------------
// JoinIdenticalVertices: the vertices some face uses, in order; one whose position is exactly
// that of an earlier kept vertex (candidates looked up by position, newest first) and whose
// other attributes are each within 1e-5 (squared distance 1e-10; absent attributes count as
// zero) is replaced by it. Faces are renumbered; unused vertices vanish.

void join_identical_vertices(Model& m)
{
    const float eps = 1e-5f;
    const float eps2 = eps * eps;
    for (Mesh& mesh : m.meshes) {
        const size_t nv = mesh.pos.size();
        if (nv == 0 || mesh.faces.empty())
            continue;
        std::vector<bool> used(nv, false);
        for (const Face& f : mesh.faces)
            for (uint32_t k = 0; k < f.n; k++)
                used[f.i[k]] = true;
        auto key = [&](const AiVec& p) {
            float c[3] = {p.x == 0.0f ? 0.0f : p.x, p.y == 0.0f ? 0.0f : p.y, p.z == 0.0f ? 0.0f : p.z};
            uint32_t u[3];
            memcpy(u, c, 12);
            return std::array<uint32_t, 3>{u[0], u[1], u[2]};
        };
        auto same = [&](size_t a, size_t b) {
            if (sq_dist(mesh.pos[a], mesh.pos[b]) > eps2)
                return false;
            if (!mesh.nrm.empty() && sq_dist(mesh.nrm[a], mesh.nrm[b]) > eps2)
                return false;
            if (!mesh.tan.empty() && sq_dist(mesh.tan[a], mesh.tan[b]) > eps2)
                return false;
            if (!mesh.bitan.empty() && sq_dist(mesh.bitan[a], mesh.bitan[b]) > eps2)
                return false;
            for (const auto& set : mesh.uv)
                if (sq_dist(set[a], set[b]) > eps2)
                    return false;
            for (const auto& set : mesh.col)
                if (colour_dist(set[a], set[b]) > eps2)
                    return false;
            return true;
        };
        std::map<std::array<uint32_t, 3>, std::vector<uint32_t>> by_pos;
        std::vector<uint32_t> kept, replace(nv, 0);
        for (size_t a = 0; a < nv; a++) {
            if (!used[a])
                continue;
            std::vector<uint32_t>& cands = by_pos[key(mesh.pos[a])];
            long found = -1;
            for (size_t c = cands.size(); c-- > 0;)
                if (same(a, kept[cands[c]])) {
                    found = long(cands[c]);
                    break;
                }
            if (found >= 0) {
                replace[a] = uint32_t(found);
            } else {
                replace[a] = uint32_t(kept.size());
                cands.push_back(uint32_t(kept.size()));
                kept.push_back(uint32_t(a));
            }
        }
        auto pick = [&](std::vector<AiVec>& v) {
            if (v.empty())
                return;
            std::vector<AiVec> r(kept.size());
            for (size_t k = 0; k < kept.size(); k++)
                r[k] = v[kept[k]];
            v.swap(r);
        };
        pick(mesh.pos);
        pick(mesh.nrm);
        pick(mesh.tan);
        pick(mesh.bitan);
        for (auto& set : mesh.uv)
            pick(set);
        for (auto& set : mesh.col) {
            std::vector<std::array<float, 4>> r(kept.size());
            for (size_t k = 0; k < kept.size(); k++)
                r[k] = set[kept[k]];
            set.swap(r);
        }
        for (Face& f : mesh.faces)
            for (uint32_t k = 0; k < f.n; k++)
                f.i[k] = replace[f.i[k]];
    }
}

void run_steps(Model& m)
{
    remove_redundant_materials(m);
    find_instances(m);
    optimize_graph(m);
    optimize_meshes(m);
    find_degenerates(m);
    find_invalid_data(m);
    gen_face_normals(m);
    join_identical_vertices(m);
}

}  // namespace steps
}  // namespace pth
