// merge A B OUT: writes the points of A followed by those of B. The
// counterpart of the reference's src/merge.cpp:58-77 (PCD ASCII output); here
// the output format follows OUT's extension as in add_rigid_transform.
#include <cstdio>
#include <string>
#include <vector>

#include "cloud_io.h"
#include "tool_common.h"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: merge A B OUT\n");
        return 1;
    }
    std::vector<emsa_io::Point> a, b;
    std::string err;
    if (!emsa_io::load_cloud(argv[1], a, err) || !emsa_io::load_cloud(argv[2], b, err))
        return emsa_tool::fail(err);
    a.insert(a.end(), b.begin(), b.end());
    return emsa_tool::save_by_extension(argv[3], a);
}
