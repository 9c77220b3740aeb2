// CPU driver of the library's host code for the sanitizer build (tests/test_sanitizers.py):
// the WGSL translator (wgsl.cpp) on every body given on the command line or in files, plus
// every prefix of each (truncated bodies exercise the error paths), and the host grid build
// (grid.cpp) on the scene files given, at several resolutions.  Built with g++ and
// -fsanitize=address,undefined; any report aborts the process.
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "grid.hpp"
#include "wgsl.hpp"

static std::string read_file(const char* path) {
    std::ifstream f(path, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

int main(int argc, char** argv) {
    long translated = 0, rejected = 0, grids = 0;
    for (int a = 1; a < argc; ++a) {
        const std::string arg = argv[a];
        if (arg.rfind("wgsl:", 0) == 0) {  // a file holding one WGSL body
            const std::string body = read_file(arg.c_str() + 5);
            for (size_t n = 0; n <= body.size(); ++n) {
                std::string out, err;
                if (linon::translate_wgsl_body(body.substr(0, n), out, err)) {
                    ++translated;
                    (void)linon::user_field_source(out);
                } else {
                    ++rejected;
                }
            }
        } else if (arg.rfind("mesh:", 0) == 0) {  // binary: uint32 nv, nf; float xyz[3 nv]; uint32 abc[3 nf]
            const std::string raw = read_file(arg.c_str() + 5);
            if (raw.size() < 8) return 2;
            uint32_t nv = 0, nf = 0;
            std::memcpy(&nv, raw.data(), 4);
            std::memcpy(&nf, raw.data() + 4, 4);
            if (raw.size() != 8 + 12ull * nv + 12ull * nf) return 2;
            const float* xyz = reinterpret_cast<const float*>(raw.data() + 8);
            const uint32_t* abc = reinterpret_cast<const uint32_t*>(raw.data() + 8 + 12ull * nv);
            for (int res : {4, 17, 64, 128}) {
                linon::GridBuild g;
                std::string err;
                if (!linon::build_grid(xyz, nv, abc, nf, res, g, err)) {
                    std::fprintf(stderr, "grid build failed: %s\n", err.c_str());
                    return 3;
                }
                ++grids;
                std::vector<float> packed;
                linon::pack_faces(xyz, abc, nf, packed);
            }
        } else {
            std::fprintf(stderr, "unknown argument %s\n", arg.c_str());
            return 2;
        }
    }
    std::printf("translated %ld, rejected %ld, grids %ld\n", translated, rejected, grids);
    return 0;
}
