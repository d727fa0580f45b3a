// Test infrastructure: decodes an image with the reference's vendored stb_image 2.30
// (include/stb/stb_image.h), the way TextureData does (src/scene/texturedata.cpp:30-46):
// stbi_info for the channel count, three widened to four, stbi_load with that count.
// Usage: stbharness in.png out.bin -> out.bin = u32 width, height, channels, then the pixels.
#define STB_IMAGE_IMPLEMENTATION
#include <stb/stb_image.h>

#include <cstdint>
#include <cstdio>

int main(int argc, char** argv)
{
    if (argc != 3)
        return 2;
    int x, y, comp;
    if (!stbi_info(argv[1], &x, &y, &comp)) {
        std::fprintf(stderr, "stbi_info failed: %s\n", stbi_failure_reason());
        return 1;
    }
    const int want = comp == 3 ? 4 : comp;
    unsigned char* px = stbi_load(argv[1], &x, &y, &comp, want);
    if (!px) {
        std::fprintf(stderr, "stbi_load failed: %s\n", stbi_failure_reason());
        return 1;
    }
    FILE* f = std::fopen(argv[2], "wb");
    if (!f)
        return 1;
    const uint32_t hdr[3] = {uint32_t(x), uint32_t(y), uint32_t(want)};
    std::fwrite(hdr, 4, 3, f);
    std::fwrite(px, 1, size_t(x) * y * want, f);
    std::fclose(f);
    stbi_image_free(px);
    return 0;
}
