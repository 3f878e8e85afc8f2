// add_rigid_transform IN OUT RT: writes IN transformed by p' = R p + t, with
// (R, t) read from the text file RT (3 rows of R, then t). The counterpart of
// the reference's src/add_rigid_transform.cpp:70-113, which writes PCD ASCII;
// here the output format follows OUT's extension (.ply: PLY ASCII, .xyz/.txt:
// text, otherwise PCD ASCII).
#include <cstdio>
#include <string>
#include <vector>

#include "cloud_io.h"
#include "tool_common.h"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: add_rigid_transform IN OUT RT_FILE\n");
        return 1;
    }
    std::vector<emsa_io::Point> c;
    std::string err;
    if (!emsa_io::load_cloud(argv[1], c, err)) return emsa_tool::fail(err);
    float R[9], t[3];
    if (!emsa_io::load_rt(argv[3], R, t)) {
        std::fprintf(stderr, "Fail to read RT from file [%s]\n", argv[3]);
        return 1;
    }
    emsa_io::transform(c, R, t);
    return emsa_tool::save_by_extension(argv[2], c);
}
