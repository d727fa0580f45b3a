// The post-processing the reference asks Assimp 5.4.3 for on every glTF file it imports
// (src/scene/scene.cpp:537-550: Triangulate | FlipUVs | GenNormals | JoinIdenticalVertices |
// RemoveRedundantMaterials | FindDegenerates | FindInvalidData | FindInstances |
// OptimizeMeshes | OptimizeGraph), restated over a model of what Assimp's glTF importer
// hands those steps: every face with its own vertices ("verbose" meshes), one material per
// mesh, a node graph referencing meshes by index.
//
// The steps decide mesh, instance and primitive numbering (which meshes merge, in which
// order their faces are concatenated, which geometry is baked into another node's frame).
// Their arithmetic is that of Assimp's vendored headers (include/assimp/matrix4x4.inl,
// matrix3x3.inl, vector3.inl, quaternion.inl); the steps themselves live in Assimp's sources,
// which the reference does not ship, and are restated from Assimp's published algorithms
// (PostProcessing/RemoveRedundantMaterials.cpp, FindInstancesProcess.cpp, OptimizeGraph.cpp,
// OptimizeMeshes.cpp, FindDegenerates.cpp, FindInvalidDataProcess.cpp,
// GenFaceNormalsProcess.cpp, JoinVerticesProcess.cpp, Common/SceneCombiner.cpp MergeMeshes).
#pragma once

#include <array>
#include <cstdint>
#include <set>
#include <string>
#include <vector>

namespace pth {
namespace steps {

// aiMatrix4x4: m[row][col] (a1..a4 is row 0), column vectors, translation in column 3.
struct AiMat4 {
    float m[4][4];
};
struct AiMat3 {
    float m[3][3];
};
struct AiVec {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

AiMat4 ai_identity();
AiMat4 ai_mul(const AiMat4& a, const AiMat4& b);  // a * b (aiMatrix4x4::operator*=)
float ai_determinant(const AiMat4& a);
AiMat4 ai_inverse(const AiMat4& a);              // NaN-filled when singular, as Assimp's
AiVec ai_transform_point(const AiMat4& m, AiVec v);
AiMat3 ai_normal_matrix(const AiMat4& m);        // aiMatrix3x3(m).Inverse().Transpose()
AiVec ai_transform_vector(const AiMat3& m, AiVec v);
AiMat4 ai_rotation(float x, float y, float z, float w);  // aiMatrix4x4(aiQuaternion(w,x,y,z).GetMatrix())
AiMat4 ai_translation(float x, float y, float z);
AiMat4 ai_scaling(float x, float y, float z);
AiVec ai_face_normal(AiVec a, AiVec b, AiVec c);  // ((b - a) ^ (c - a)).NormalizeSafe()

enum : uint32_t { PRIM_POINT = 1, PRIM_LINE = 2, PRIM_TRIANGLE = 4, PRIM_POLYGON = 8 };

struct Face {
    uint32_t n = 0;
    uint32_t i[3] = {0, 0, 0};
};

struct Mesh {
    std::string name;
    std::vector<AiVec> pos, nrm, tan, bitan;         // nrm / tan / bitan empty when absent
    std::vector<std::vector<AiVec>> uv;               // texture coordinate sets (z = 0)
    std::vector<uint32_t> uv_comps;                   // components per set
    std::vector<std::vector<std::array<float, 4>>> col;  // colour sets
    std::vector<Face> faces;
    uint32_t material = 0;
    uint32_t ptypes = 0;
    bool bones = false;
};

struct Node {
    std::string name;
    AiMat4 transform;
    std::vector<uint32_t> meshes;
    std::vector<uint32_t> children;  // indices into Model::nodes
};

struct Model {
    std::vector<Mesh> meshes;
    // One entry per material: equal keys mean equal material properties (the name left out,
    // as Assimp's material hash leaves out '?'-keys); the source is the caller's own index.
    std::vector<std::string> material_keys;
    std::vector<int> material_source;
    std::vector<Node> nodes;
    uint32_t root = 0;
    std::set<std::string> locked;  // names OptimizeGraph keeps: cameras, lights, animated and bone nodes
};

uint32_t vertex_format(const Mesh& mesh);  // GetMeshVFormatUnique

void remove_redundant_materials(Model& model);
void find_instances(Model& model);
void optimize_graph(Model& model);
void optimize_meshes(Model& model);
void find_degenerates(Model& model);
void find_invalid_data(Model& model);
void gen_face_normals(Model& model);
void join_identical_vertices(Model& model);

// All of the above in the order Assimp's step registry runs them.
void run_steps(Model& model);

}  // namespace steps
}  // namespace pth
