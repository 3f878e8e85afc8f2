// Point-cloud files and the input preparation of the command-line harness.
//
// The reference reads clouds with PCL's loadPolygonFile (VTK readers; NaN
// points removed, src/main.cpp:57-74), reduces them by random sampling
// (src/main.cpp:77-99), optionally normalises their scale
// (alignScaleOnce, src/main.cpp:124-161) and reads / writes (R, t) as text
// (src/main.cpp:197-228). This header is the same functionality without PCL or
// VTK: readers for the formats the reference's data and tools use, a seeded
// reduction, and the scale normalisation.
#ifndef EMSA_CLOUD_IO_H_
#define EMSA_CLOUD_IO_H_

#include <cstdint>
#include <string>
#include <vector>

namespace emsa_io {

struct Point {
    float x, y, z;
};

// Reads a cloud. Format by content and extension:
//  - PLY ("ply" magic): ascii, binary_little_endian or binary_big_endian; the
//    x, y, z properties of element "vertex" (float, double or integer types;
//    other properties and elements are skipped);
//  - PCD (".pcd"): DATA ascii or binary, fields x y z (size 4 or 8);
//  - otherwise text: the first three numbers of every line that has them
//    (".txt", ".xyz"; blank lines and lines starting with '#' skipped).
// Line ends may be LF or CRLF. Points with a non-finite coordinate are
// dropped, as the reference's removeNaNFromPointCloud does. Returns false and
// sets err on failure (missing file, bad header, short data).
bool load_cloud(const std::string& path, std::vector<Point>& out, std::string& err);

// ASCII writers (the reference's tools write PCD ASCII and PLY).
bool save_pcd_ascii(const std::string& path, const std::vector<Point>& cloud, std::string& err);
bool save_ply_ascii(const std::string& path, const std::vector<Point>& cloud, std::string& err);
bool save_xyz(const std::string& path, const std::vector<Point>& cloud, std::string& err);

// Keeps each point with probability percent / 100, in order (the reference
// deletes a point when rand()/RAND_MAX >= percent/100, seeded from the clock,
// src/main.cpp:77-99). Here the draws come from splitmix64 seeded with
// `seed`, so a run is reproducible.
void points_reduction(std::vector<Point>& cloud, float percent, uint64_t seed);

// Trace of the covariance of a cloud (mean squared distance to its centroid;
// the covariance normalised by the point count, as PCL's
// computeMeanAndCovarianceMatrix).
double covariance_trace(const std::vector<Point>& cloud);

// Scales each cloud about the origin by sqrt(0.003 / trace(cov)) (the
// reference's alignScaleOnce, src/main.cpp:124-161). Returns the traces
// before and after in tr[0..3] = {target before, source before, target after,
// source after}.
void align_scale_once(std::vector<Point>& target, std::vector<Point>& source, double tr[4]);

// (R, t) text files: 3 rows of R then t, "%f " separated (src/main.cpp:197-228).
bool save_rt(const std::string& path, const float R[9], const float t[3]);
bool load_rt(const std::string& path, float R[9], float t[3]);

// p' = R p + t for every point.
void transform(std::vector<Point>& cloud, const float R[9], const float t[3]);

}  // namespace emsa_io

#endif
