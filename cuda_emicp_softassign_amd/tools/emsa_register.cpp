// Command-line registration harness: the counterpart of the reference's
// src/main.cpp (flags :245-327, loading :348-352, alignScaleOnce :356-359,
// pointsReduction :363-380, (R, t) files :434-440 and :475-479, the timed
// call of the selected algorithms :448-470), written against the drop-in API
// of include/3dregistration.h.
//
//   emsa_register -pointFileX=target.ply -pointFileY=source.ply [-emicp]
//                 [-softassign] [-emicpcpu] [-icp] [parameters] [options]
//
// Flags are "-name" or "--name"; values are "-name=value". A numeric
// parameter that is absent or 0 takes its default (the reference's
// zero-means-default rule). Algorithms run in the order icp, emicp,
// emicpcpu, softassign, each from the (R, t) the previous one left, as in the
// reference; with none selected, softassign runs.
//
// Differences from the reference, all additive: -seed=N seeds the point
// reduction (the reference seeds rand() from the clock, so X and Y get the same
// keep pattern when reduced in the same second; here both use the same seed
// too, so a run is reproducible); -logdomain selects the log-domain Softassign;
// -saveRTtrace=FILE writes every iteration's (R, t) (through the per-iteration
// hook that drives the reference's viewer); -saveCloudsTo=PREFIX writes the
// prepared clouds as PREFIX_X.xyz and PREFIX_Y.xyz. There is no PCL viewer
// here: without -noviewer a note is printed and the run continues as with it.
// The elapsed time is wall clock (the reference's clock() is process CPU time).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "3dregistration.h"
#include "cloud_io.h"

namespace {

// The value part of argv[i] for flag `name`: "" for a bare "-name", the text
// after '=' for "-name=value", nullptr when argv[i] is another flag.
const char* match(const char* arg, const char* name) {
    while (*arg == '-') ++arg;
    const size_t n = std::strlen(name);
    if (std::strncmp(arg, name, n) != 0) return nullptr;
    if (arg[n] == 0) return arg + n;
    if (arg[n] == '=') return arg + n + 1;
    return nullptr;
}

struct Args {
    int argc;
    char** argv;
    const char* value(const char* name) const {
        for (int i = 1; i < argc; ++i)
            if (argv[i][0] == '-') {
                const char* v = match(argv[i], name);
                if (v) return v;
            }
        return nullptr;
    }
    bool flag(const char* name) const { return value(name) != nullptr; }
    const char* str(const char* name) const {
        const char* v = value(name);
        return v && *v ? v : nullptr;
    }
    int i(const char* name) const {
        const char* v = value(name);
        return v ? std::atoi(v) : 0;
    }
    float f(const char* name) const {
        const char* v = value(name);
        return v ? (float)std::atof(v) : 0.f;
    }
};

pcl::PointCloud<pcl::PointXYZ>::Ptr to_cloud(const std::vector<emsa_io::Point>& v) {
    pcl::PointCloud<pcl::PointXYZ>::Ptr c(new pcl::PointCloud<pcl::PointXYZ>());
    c->points.reserve(v.size());
    for (const emsa_io::Point& p : v) c->points.push_back(pcl::PointXYZ(p.x, p.y, p.z));
    return c;
}

void trace_line(int algorithm, int iteration, const float* R, const float* t, void* user) {
    FILE* f = static_cast<FILE*>(user);
    std::fprintf(f, "%d %d", algorithm, iteration);
    for (int k = 0; k < 9; ++k) std::fprintf(f, " %.9g", R[k]);
    for (int k = 0; k < 3; ++k) std::fprintf(f, " %.9g", t[k]);
    std::fprintf(f, "\n");
}

void usage() {
    std::fprintf(stderr,
                 "Wrong arguments.\n"
                 "usage: emsa_register -pointFileX=FILE -pointFileY=FILE [-icp] [-emicp] "
                 "[-emicpcpu] [-softassign]\n"
                 "  softassign: -JMAX= -I0= -I1= -alpha= -T_0= -TFACTOR= -moutlier= "
                 "[-logdomain]\n"
                 "  EM-ICP:     -sigma_p2= -sigma_inf= -sigma_factor= -d_02=\n"
                 "  ICP:        -maxIteration=\n"
                 "  input:      -alignScaleOnce -pointsReductionRate= -pointsReductionRateX= "
                 "-pointsReductionRateY= -seed=\n"
                 "  (R, t):     -loadRTfromFile= -saveRTtoFile= -saveRTtrace=\n"
                 "  misc:       -noviewer -nostop -notimer -saveCloudsTo=PREFIX\n"
                 "min ||X - (R*Y+t)||\n");
}

}  // namespace

int main(int argc, char** argv) {
    const Args a{argc, argv};
    const char* fileX = a.str("pointFileX");
    const char* fileY = a.str("pointFileY");
    if (!fileX || !fileY) {
        usage();
        return 1;
    }
    std::printf("option: pointFileX= %s\noption: pointFileY= %s\n", fileX, fileY);

    const bool isICP = a.flag("icp"), isEMICP = a.flag("emicp"), isCPU = a.flag("emicpcpu");
    bool isSA = a.flag("softassign");
    if (!isICP && !isEMICP && !isCPU && !isSA) isSA = true;  // default algorithm

    emsa_params d;
    emsa_default_params(&d);
    registrationParameters param{};
    param.JMAX = d.JMAX;
    param.I0 = d.I0;
    param.I1 = d.I1;
    param.alpha = d.alpha;
    param.T_0 = d.T_0;
    param.TFACTOR = d.TFACTOR;
    param.moutlier = d.moutlier;
    param.sigma_p2 = d.sigma_p2;
    param.sigma_inf = d.sigma_inf;
    param.sigma_factor = d.sigma_factor;
    param.d_02 = d.d_02;
    param.maxIteration = d.maxIteration;
    if (isSA) {
        if (int v = a.i("JMAX")) param.JMAX = v;
        if (int v = a.i("I0")) param.I0 = v;
        if (int v = a.i("I1")) param.I1 = v;
        if (float v = a.f("alpha")) param.alpha = v;
        if (float v = a.f("T_0")) param.T_0 = v;
        if (float v = a.f("TFACTOR")) param.TFACTOR = v;
        // default slack value from the T_0 in effect (src/main.cpp:286)
        param.moutlier = (float)(1 / std::sqrt(param.T_0) * std::exp(-1.0f));
        if (float v = a.f("moutlier")) param.moutlier = v;
        std::printf("softassign parameters\nJMAX %d\nI0 %d\nI1 %d\nalpha %g\nT_0 %g\n"
                    "TFACTOR %g\nmoutlier %g\n",
                    param.JMAX, param.I0, param.I1, param.alpha, param.T_0, param.TFACTOR,
                    param.moutlier);
        if (a.flag("logdomain")) emsa_set_softassign_variant(EMSA_SOFTASSIGN_LOGDOMAIN);
    }
    if (isICP) {
        if (int v = a.i("maxIteration")) param.maxIteration = v;
        std::printf("ICP parameters\nmaxIteration %d\n", param.maxIteration);
    }
    if (isEMICP || isCPU) {
        if (float v = a.f("sigma_p2")) param.sigma_p2 = v;
        if (float v = a.f("sigma_inf")) param.sigma_inf = v;
        if (float v = a.f("sigma_factor")) param.sigma_factor = v;
        if (float v = a.f("d_02")) param.d_02 = v;
        std::printf("EM-ICP parameters\nsigma_p2 %g\nsigma_inf %g\nsigma_factor %g\nd_02 %g\n",
                    param.sigma_p2, param.sigma_inf, param.sigma_factor, param.d_02);
    }
    param.noviewer = a.flag("noviewer");
    param.notimer = a.flag("notimer");
    param.nostop = a.flag("nostop");
    param.argc = argc;
    param.argv = argv;

    std::vector<emsa_io::Point> X, Y;
    std::string err;
    if (!emsa_io::load_cloud(fileX, X, err) || !emsa_io::load_cloud(fileY, Y, err)) {
        std::fprintf(stderr, "loadFile failed: %s\n", err.c_str());
        return 1;
    }
    if (a.flag("alignScaleOnce")) {
        double tr[4];
        emsa_io::align_scale_once(X, Y, tr);
        std::printf("scale1 = %g\nscale2 = %g\nscale1 = %g\nscale2 = %g\n", tr[0], tr[1], tr[2],
                    tr[3]);
    }
    const int seedArg = a.i("seed");
    const uint64_t seed = seedArg ? (uint64_t)(int64_t)seedArg : 1;
    if (float r = a.f("pointsReductionRate")) {
        emsa_io::points_reduction(X, r, seed);
        emsa_io::points_reduction(Y, r, seed);
        std::printf("number of points X,Y are reduced to %g%% of original.\n", r);
    } else {
        if (float r = a.f("pointsReductionRateX")) {
            emsa_io::points_reduction(X, r, seed);
            std::printf("number of points X are reduced to %g%% of original.\n", r);
        }
        if (float r = a.f("pointsReductionRateY")) {
            emsa_io::points_reduction(Y, r, seed);
            std::printf("number of points Y are reduced to %g%% of original.\n", r);
        }
    }
    std::printf("Xsize: %zu\nYsize: %zu\n", X.size(), Y.size());
    if (X.empty() || Y.empty()) {
        std::fprintf(stderr, "empty point cloud after loading and reduction\n");
        return 1;
    }
    if (const char* pre = a.str("saveCloudsTo")) {
        if (!emsa_io::save_xyz(std::string(pre) + "_X.xyz", X, err) ||
            !emsa_io::save_xyz(std::string(pre) + "_Y.xyz", Y, err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            return 1;
        }
    }
    if (!param.noviewer)
        std::printf("no 3D viewer in this build: running as with -noviewer "
                    "(-saveRTtrace=FILE records every iteration)\n");

    const pcl::PointCloud<pcl::PointXYZ>::Ptr target = to_cloud(X), source = to_cloud(Y);
    float h_R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, h_t[3] = {0, 0, 0};
    if (const char* f = a.str("loadRTfromFile")) {
        FILE* probe = std::fopen(f, "r");
        if (probe) {
            std::fclose(probe);
            if (!emsa_io::load_rt(f, h_R, h_t)) {
                std::fprintf(stderr, "Fail to read RT from file [%s]\n", f);
                return 1;
            }
        } else {  // the reference keeps the identity when the file cannot be opened
            std::fprintf(stderr, "cannot open [%s]: starting from the identity\n", f);
        }
    }
    printRT(h_R, h_t);

    FILE* traceFile = nullptr;
    if (const char* f = a.str("saveRTtrace")) {
        traceFile = std::fopen(f, "w");
        if (!traceFile) {
            std::fprintf(stderr, "cannot write [%s]\n", f);
            return 1;
        }
        emsa_set_iteration_callback(trace_line, traceFile);
    }

    int failed = 0;
    auto ran = [&]() {
        if (emsa_last_error()[0]) failed = 1;
    };
    const auto t0 = std::chrono::steady_clock::now();
    if (isICP) icp(target, source, h_R, h_t, param), ran();
    if (isEMICP) emicp(target, source, h_R, h_t, param), ran();
    if (isCPU) emicp_cpu(target, source, h_R, h_t, param), ran();
    if (isSA) softassign(target, source, h_R, h_t, param), ran();
    const double el =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (traceFile) {
        emsa_set_iteration_callback(nullptr, nullptr);
        std::fclose(traceFile);
    }
    std::printf("elapsed %f\n", el);
    printRT(h_R, h_t);
    if (const char* f = a.str("saveRTtoFile")) {
        if (!emsa_io::save_rt(f, h_R, h_t)) {
            std::fprintf(stderr, "cannot write [%s]\n", f);
            return 1;
        }
    }
    return failed ? 2 : 0;
}
