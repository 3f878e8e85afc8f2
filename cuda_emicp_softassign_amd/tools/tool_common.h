// Shared helpers of the file tools (add_rigid_transform, merge, pcd2ply).
#ifndef EMSA_TOOL_COMMON_H_
#define EMSA_TOOL_COMMON_H_

#include <cstdio>
#include <string>
#include <vector>

#include "cloud_io.h"

namespace emsa_tool {

inline int fail(const std::string& err) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
}

inline bool has_suffix(const std::string& s, const char* suf) {
    const std::string x(suf);
    return s.size() >= x.size() && s.compare(s.size() - x.size(), x.size(), x) == 0;
}

// .ply: PLY ASCII; .xyz / .txt: text; anything else: PCD ASCII.
inline int save_by_extension(const std::string& path, const std::vector<emsa_io::Point>& c) {
    std::string err;
    bool ok;
    if (has_suffix(path, ".ply"))
        ok = emsa_io::save_ply_ascii(path, c, err);
    else if (has_suffix(path, ".xyz") || has_suffix(path, ".txt"))
        ok = emsa_io::save_xyz(path, c, err);
    else
        ok = emsa_io::save_pcd_ascii(path, c, err);
    return ok ? 0 : fail(err);
}

}  // namespace emsa_tool

#endif
