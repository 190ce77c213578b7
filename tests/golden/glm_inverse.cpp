// Fixture program: glm::inverse of float 4x4 matrices, as PushEntity stores
// the entity inverse (Source/Core/BVH/Intersector.h:209).  Reads N matrices
// (16 floats each, column-major, glm's own memory order) from stdin and writes
// their inverses in the same layout to stdout.  Built by make_glm_inverse.py
// against the reference's vendored glm 0.9.8.5 headers; no GL and no other
// reference code is involved.
#include <cstdio>
#include <cstring>
#include <vector>

#include <glm/glm.hpp>

int main() {
    std::vector<float> in;
    float buf[16];
    while (std::fread(buf, sizeof(float), 16, stdin) == 16) in.insert(in.end(), buf, buf + 16);
    for (size_t k = 0; k + 16 <= in.size(); k += 16) {
        glm::mat4 m;
        std::memcpy(&m[0][0], &in[k], sizeof(buf));
        const glm::mat4 inv = glm::inverse(m);
        std::fwrite(&inv[0][0], sizeof(float), 16, stdout);
    }
    return 0;
}
