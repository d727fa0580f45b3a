// Golden-vector harness over the reference's own host code.
//
// Test infrastructure only: this file is compiled by oracle/Makefile together with the
// reference's rng.cpp, colour.cpp, lalgebra.cpp, scene/camera.cpp and io/imagewrite.cpp,
// read where they lie under the reference checkout. Nothing in the product links it.
// It prints JSON on stdout; tests/golden/make_golden.py turns that into fixtures.
//
// Reference functions exercised (paths relative to the reference root):
//   Splitmix32::next, Xshiro128::next/next_float      src/rng.cpp:24-60
//   AliasTable<T>::build / copy                         src/aliastable.h:21-77
//   from_srgb / to_srgb                                 src/colour.cpp:3-15
//   translation / rotation / scaling                    src/lalgebra.cpp:5-58
//   Camera::view_matrix / projection_matrix             src/scene/camera.h:16, camera.cpp:21-25
//   write_hdr (stbi_write_hdr)                          src/io/imagewrite.cpp:19-37
// The camera basis and the InstanceData transform packing are the few glm lines of
// src/scene/scenebuild.cpp:184-194 and src/scene/scenebuffer.cpp:131-132, which cannot be
// linked on their own (they sit in files that need Assimp/Vulkan).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "aliastable.h"
#include "colour.h"
#include "io/imagewrite.h"
#include "lalgebra.h"
#include "rng.h"
#include "scene/camera.h"

static void print_u32s(const std::vector<uint32_t>& v)
{
    printf("[");
    for (size_t i = 0; i < v.size(); i++)
        printf("%s%u", i ? "," : "", v[i]);
    printf("]");
}

static void print_floats(const float* v, size_t n)
{
    printf("[");
    for (size_t i = 0; i < n; i++) {
        uint32_t bits;
        memcpy(&bits, &v[i], 4);
        printf("%s%u", i ? "," : "", bits);  // exact bit patterns
    }
    printf("]");
}

static void print_mat4(const Mat4& m)
{
    float f[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            f[c * 4 + r] = m[c][r];  // column-major, as glm stores it
    print_floats(f, 16);
}

struct CameraCase
{
    Vec3 position, look_dir, up;
    float fov_deg, near, far, aspect;
};

int main(int argc, char** argv)
{
    const std::string tmp_dir = argc > 1 ? argv[1] : "/tmp";
    printf("{\n");

    // --- rng.cpp ---------------------------------------------------------------------
    printf("\"rng\": [\n");
    const uint32_t seeds[] = { 0u, 1u, 0xbd7a0d48u, 12345u, 0xdeadbeefu };
    for (size_t si = 0; si < sizeof(seeds) / sizeof(seeds[0]); si++) {
        Splitmix32 sm(seeds[si]);
        std::vector<uint32_t> sm_out;
        for (int i = 0; i < 8; i++)
            sm_out.push_back(sm.next());
        Splitmix32 seeder(seeds[si]);
        uint32_t s0 = seeder.next(), s1 = seeder.next(), s2 = seeder.next(), s3 = seeder.next();
        Xshiro128 x(s0, s1, s2, s3);
        std::vector<uint32_t> xs_out;
        for (int i = 0; i < 16; i++)
            xs_out.push_back(x.next());
        Xshiro128 y(s0, s1, s2, s3);
        std::vector<float> fl;
        for (int i = 0; i < 16; i++)
            fl.push_back(y.next_float());
        printf("{\"seed\": %u, \"splitmix\": ", seeds[si]);
        print_u32s(sm_out);
        printf(", \"xshiro\": ");
        print_u32s(xs_out);
        printf(", \"xshiro_float_bits\": ");
        print_floats(fl.data(), fl.size());
        printf("}%s\n", si + 1 < sizeof(seeds) / sizeof(seeds[0]) ? "," : "");
    }
    printf("],\n");

    // --- aliastable.h ----------------------------------------------------------------
    {
        // The seeding expression of calculate_object_intensity (src/scene/lightsample.cpp:43-46),
        // compiled as the reference compiles it (g++): the order in which the four
        // seeder.next() arguments are evaluated decides the generator's state.
        Splitmix32 seeder(0xbd7a0d48u);
        Xshiro128 rng(seeder.next(), seeder.next(), seeder.next(), seeder.next());
        std::vector<float> fl;
        for (int i = 0; i < 8; i++)
            fl.push_back(rng.next_float());
        printf("\"intensity_rng_float_bits\": ");
        print_floats(fl.data(), fl.size());
        printf(",\n");
    }
    printf("\"alias\": [\n");
    std::vector<std::vector<float>> weight_sets = {
        { 1.0f },
        { 1.0f, 1.0f, 1.0f, 1.0f },
        { 1.0f, 2.0f, 3.0f, 4.0f },
        { 0.1f, 5.0f, 0.0f, 2.5f, 7.25f, 0.01f, 3.0f },
        { 10.0f, 0.0f, 0.0f, 0.0f, 1.0f },
    };
    {
        // A longer, seeded set: emitter-area-like weights.
        Splitmix32 seeder(777u);
        Xshiro128 r(seeder.next(), seeder.next(), seeder.next(), seeder.next());
        std::vector<float> w;
        for (int i = 0; i < 100; i++)
            w.push_back(r.next_float() * 3.0f + 0.001f);
        weight_sets.push_back(w);
    }
    for (size_t wi = 0; wi < weight_sets.size(); wi++) {
        const std::vector<float>& w = weight_sets[wi];
        std::vector<uint32_t> values(w.size());
        for (size_t i = 0; i < w.size(); i++)
            values[i] = 1000u + (uint32_t)i;
        AliasTable<uint32_t> table(values.data(), w.data(), w.size());
        std::vector<uint32_t> words(table.size_in_bytes() / 4);
        table.copy(words.data());
        printf("{\"weights_bits\": ");
        print_floats(w.data(), w.size());
        printf(", \"words\": ");
        print_u32s(words);
        printf("}%s\n", wi + 1 < weight_sets.size() ? "," : "");
    }
    printf("],\n");

    // --- colour.cpp ------------------------------------------------------------------
    {
        std::vector<float> xs, from, to;
        for (int i = 0; i <= 64; i++)
            xs.push_back(i / 64.0f);
        const float extra[] = { 0.7f, 0.04045f, 0.0031308f, 0.00031308f, 0.375f, 0.45f, 0.9f, 1.5f, 0.2f, 0.8f };
        for (float e : extra)
            xs.push_back(e);
        for (float x : xs) {
            from.push_back(from_srgb(x));
            to.push_back(to_srgb(x));
        }
        printf("\"srgb\": {\"x_bits\": ");
        print_floats(xs.data(), xs.size());
        printf(", \"from_bits\": ");
        print_floats(from.data(), from.size());
        printf(", \"to_bits\": ");
        print_floats(to.data(), to.size());
        printf("},\n");
    }

    // --- lalgebra.cpp: transforms + InstanceData packing (scenebuffer.cpp:131-132) ----
    printf("\"transforms\": [\n");
    struct XformCase
    {
        Vec3 t, axis, s;
        float angle;
    };
    const XformCase xcases[] = {
        { Vec3(1, 2, 3), Vec3(0, 1, 0), Vec3(1, 1, 1), 0.5f },
        { Vec3(-4.5f, 0.25f, 10), Vec3(1, 1, 0), Vec3(2, 2, 2), 2.0f },
        { Vec3(0, 0, 0), Vec3(0.3f, -0.2f, 0.9f), Vec3(0.5f, 0.5f, 0.5f), 4.1f },
        { Vec3(7, -8, 9), Vec3(0, 0, 1), Vec3(3, 3, 3), 0.000001f },
        { Vec3(-30, 41, 12), Vec3(-0.6f, 0.64f, 0.48f), Vec3(1.7f, 1.7f, 1.7f), 5.9f },
    };
    const size_t nx = sizeof(xcases) / sizeof(xcases[0]);
    for (size_t i = 0; i < nx; i++) {
        const XformCase& c = xcases[i];
        // .scn semantics: each line pre-multiplies (scenebuild.cpp:213-221); a block
        // "s ..., r ..., t ..." yields T * R * S.
        Mat4 m = scaling(c.s);
        m = rotation(c.axis, c.angle) * m;
        m = translation(c.t) * m;
        Mat4 transform = glm::transpose(m);
        Mat4 inv = glm::inverse(transform);
        Mat3x4 normal = glm::transpose(inv);
        float nf[12];
        for (int col = 0; col < 3; col++)
            for (int r = 0; r < 4; r++)
                nf[col * 4 + r] = normal[col][r];
        printf("{\"t\": [%.9g,%.9g,%.9g], \"axis\": [%.9g,%.9g,%.9g], \"angle\": %.9g, \"s\": [%.9g,%.9g,%.9g], ",
               c.t.x, c.t.y, c.t.z, c.axis.x, c.axis.y, c.axis.z, c.angle, c.s.x, c.s.y, c.s.z);
        printf("\"rotation\": ");
        print_mat4(rotation(c.axis, c.angle));
        printf(", \"matrix\": ");
        print_mat4(m);
        printf(", \"instance_transform\": ");
        print_mat4(transform);
        printf(", \"instance_normal\": ");
        print_floats(nf, 12);
        printf("}%s\n", i + 1 < nx ? "," : "");
    }
    printf("],\n");

    // --- camera (scenebuild.cpp:184-194, camera.h:16, camera.cpp:21-25,
    //             pathtrace.cpp:381-396) ---------------------------------------------
    printf("\"cameras\": [\n");
    const CameraCase ccases[] = {
        { Vec3(0, 1, -3), Vec3(0, 0, 1), Vec3(0, 1, 0), 60.0f, 0.1f, 1000.0f, 1.0f },
        { Vec3(0, 0, 120), Vec3(0, 0, -1), Vec3(0, 1, 0), 60.0f, 0.1f, 1000.0f, 1920.0f / 1080.0f },
        { Vec3(3, -2, 5), Vec3(0.3f, 0.2f, -0.9f), Vec3(0, 1, 0), 45.0f, 0.1f, 1000.0f, 1.5f },
        { Vec3(-10, 4, 2), Vec3(1, -0.5f, 0.25f), Vec3(0, 1, 0), 75.0f, 0.5f, 200.0f, 0.6f },
    };
    const size_t nc = sizeof(ccases) / sizeof(ccases[0]);
    for (size_t i = 0; i < nc; i++) {
        const CameraCase& c = ccases[i];
        Vec3 x = glm::normalize(c.look_dir);
        Vec3 y = glm::normalize(glm::cross(c.up, x));
        Vec3 z = glm::cross(x, y);
        Camera cam(c.position, Quaternion(Mat3(y, z, x)), glm::radians(c.fov_deg), 1.0, c.near, c.far);
        cam.aspect_ratio = c.aspect;
        Mat4 view = cam.view_matrix();
        Mat4 proj = cam.projection_matrix();
        Mat4 inv_proj = glm::transpose(glm::inverse(proj));
        Mat4 inv_view = glm::transpose(glm::inverse(view));
        printf("{\"position\": [%.9g,%.9g,%.9g], \"lookdir\": [%.9g,%.9g,%.9g], \"up\": [%.9g,%.9g,%.9g], "
               "\"fov\": %.9g, \"near\": %.9g, \"far\": %.9g, \"aspect_bits\": ",
               c.position.x, c.position.y, c.position.z, c.look_dir.x, c.look_dir.y, c.look_dir.z,
               c.up.x, c.up.y, c.up.z, c.fov_deg, c.near, c.far);
        print_floats(&c.aspect, 1);
        printf(", \"inv_view_rowmajor\": ");
        print_mat4(inv_view);  // transposed glm matrix printed column-major == row-major original
        printf(", \"inv_proj_rowmajor\": ");
        print_mat4(inv_proj);
        printf("}%s\n", i + 1 < nc ? "," : "");
    }
    printf("],\n");

    // --- imagewrite.cpp -> stbi_write_hdr ----------------------------------------------
    printf("\"hdr\": [\n");
    struct HdrCase
    {
        uint32_t w, h, seed;
    };
    const HdrCase hcases[] = { { 3, 2, 1 }, { 9, 3, 2 }, { 40, 5, 3 }, { 16, 16, 4 } };
    const size_t nh = sizeof(hcases) / sizeof(hcases[0]);
    for (size_t i = 0; i < nh; i++) {
        const HdrCase& c = hcases[i];
        Splitmix32 seeder(c.seed);
        Xshiro128 r(seeder.next(), seeder.next(), seeder.next(), seeder.next());
        std::vector<float> img(c.w * c.h * 4);
        for (size_t k = 0; k < img.size(); k++) {
            float v = r.next_float();
            // a mix of magnitudes, zeros and runs so RLE and exponent paths are exercised
            if (k % 11 == 0)
                v = 0.0f;
            else if (k % 7 == 0)
                v = v * 1000.0f;
            else if (k % 5 == 0)
                v = v * 1e-3f;
            if ((k / 4) % 13 < 3)
                v = 0.5f;  // runs of identical pixels
            img[k] = v;
        }
        std::string path = tmp_dir + "/golden_hdr_" + std::to_string(i) + ".hdr";
        write_hdr((const uint8_t*)img.data(), path, c.w, c.h, c.w * 16);
        // A perturbed twin of the same size, for the meansq fixture pairs.
        std::vector<float> twin(img);
        for (size_t k = 0; k < twin.size(); k++)
            twin[k] = twin[k] * (1.0f + 0.25f * r.next_float()) + 0.01f * (float)(k % 3);
        write_hdr((const uint8_t*)twin.data(), tmp_dir + "/golden_hdr_" + std::to_string(i) + "_b.hdr", c.w, c.h, c.w * 16);
        std::ifstream f(path, std::ios::binary);
        std::vector<unsigned char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        printf("{\"w\": %u, \"h\": %u, \"rgba_bits\": ", c.w, c.h);
        print_floats(img.data(), img.size());
        printf(", \"file_hex\": \"");
        for (unsigned char b : bytes)
            printf("%02x", b);
        printf("\", \"path\": \"%s\"}%s\n", path.c_str(), i + 1 < nh ? "," : "");
    }
    printf("]\n");

    printf("}\n");
    return 0;
}
