// pcd2ply IN OUT: converts a cloud (PCD, PLY or text) to PLY ASCII. The
// counterpart of the reference's src/pcd2ply.cpp:58-70.
#include <cstdio>
#include <string>
#include <vector>

#include "cloud_io.h"
#include "tool_common.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: pcd2ply IN OUT\n");
        return 1;
    }
    std::vector<emsa_io::Point> c;
    std::string err;
    if (!emsa_io::load_cloud(argv[1], c, err)) return emsa_tool::fail(err);
    if (!emsa_io::save_ply_ascii(argv[2], c, err)) return emsa_tool::fail(err);
    return 0;
}
