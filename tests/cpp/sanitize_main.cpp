// Host-code exercise for the sanitizer build (AddressSanitizer +
// UndefinedBehaviorSanitizer on the host side of libemsa): the drop-in
// header's host functions, the FP32 host EM-ICP with its trace and
// per-iteration hook, and the argument checks of every host entry point that
// fail before any device work. Needs no GPU: GPU entry points given valid
// arguments return EMSA_ENODEV here, or run when a device is present. Exit
// status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "3dregistration.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s (%s)\n", what, emsa_last_error());
        ++failures;
    }
}

int hook_calls = 0;
void hook(int, int, const float*, const float*, void*) { ++hook_calls; }

}  // namespace

int main() {
    const int n = 257;
    pcl::PointCloud<pcl::PointXYZ>::Ptr target(new pcl::PointCloud<pcl::PointXYZ>);
    pcl::PointCloud<pcl::PointXYZ>::Ptr source(new pcl::PointCloud<pcl::PointXYZ>);
    unsigned seed = 7u;
    auto rnd = [&seed]() {
        seed = seed * 1664525u + 1013904223u;
        return (seed >> 8) / 16777216.0f - 0.5f;
    };
    for (int i = 0; i < n; ++i) {
        pcl::PointXYZ y(rnd(), rnd(), rnd());
        source->points.push_back(y);
        target->points.push_back(pcl::PointXYZ(y.x + 0.01f, y.y - 0.02f, y.z));
    }

    // drop-in host functions
    float* X = nullptr;
    int Xsize = 0;
    cloud2data(target, &X, Xsize);
    expect(Xsize == n && X[n] == target->points[0].y, "cloud2data float**");
    boost::shared_array<float> Ys;
    int Ysize = 0;
    cloud2data(source, Ys, Ysize);
    expect(Ysize == n && Ys[2 * n + 3] == source->points[3].z, "cloud2data shared_array");
    std::vector<float> Y(Ys.get(), Ys.get() + 3 * n);

    registrationParameters param{};
    emsa_params d;
    emsa_default_params(&d);
    param.sigma_p2 = d.sigma_p2;
    param.sigma_inf = 1e-3f;  // a short schedule
    param.sigma_factor = d.sigma_factor;
    param.d_02 = d.d_02;
    param.noviewer = 1;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t[3] = {0, 0, 0};
    emicp_cpu(target, source, R, t, param);
    expect(std::fabs(t[0] - 0.01f) < 1e-3f && std::fabs(t[1] + 0.02f) < 1e-3f, "emicp_cpu");
    printRT(R, t);

    // C ABI: FP32 host EM-ICP with trace and hook
    emsa_params p = d;
    p.sigma_inf = 1e-3f;
    const int steps = emsa_emicp_iterations(&p);
    expect(steps > 0, "emicp_iterations");
    std::vector<double> trace((size_t)steps * EMSA_TRACE_W);
    emsa_set_iteration_callback(hook, nullptr);
    float R2[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, t2[3] = {0, 0, 0};
    const int rc = emsa_emicp_cpu(X, n, Y.data(), n, R2, t2, &p, trace.data(), steps - 1);
    emsa_set_iteration_callback(nullptr, nullptr);
    expect(rc == steps - 1 && hook_calls == steps - 1, "emicp_cpu trace capacity and hook");

    // argument checks: rejected before any device work
    const float nan = std::numeric_limits<float>::quiet_NaN();
    float Rb[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, tb[3] = {0, 0, 0};
    expect(emsa_emicp(nullptr, n, Y.data(), n, Rb, tb, &p) == EMSA_EINVAL, "emicp null X");
    expect(emsa_emicp(X, 0, Y.data(), n, Rb, tb, &p) == EMSA_EINVAL, "emicp N = 0");
    expect(emsa_emicp(X, n, Y.data(), n, Rb, tb, nullptr) == EMSA_EINVAL, "emicp null params");
    emsa_params bad = p;
    bad.sigma_factor = 1.0f;
    expect(emsa_emicp(X, n, Y.data(), n, Rb, tb, &bad) == EMSA_EINVAL, "emicp factor 1");
    bad = p;
    bad.sigma_inf = nan;
    expect(emsa_emicp(X, n, Y.data(), n, Rb, tb, &bad) < 0, "emicp NaN sigma_inf");
    expect(emsa_emicp_cpu(X, n, Y.data(), n, Rb, tb, &bad, nullptr, 0) < 0, "emicp_cpu NaN");
    std::vector<float> Xn(X, X + 3 * n);
    Xn[5] = nan;
    expect(emsa_emicp(Xn.data(), n, Y.data(), n, Rb, tb, &p) == EMSA_ENONFINITE, "emicp NaN X");
    expect(emsa_emicp_cpu(Xn.data(), n, Y.data(), n, Rb, tb, &p, nullptr, 0) == EMSA_ENONFINITE,
           "emicp_cpu NaN X");
    expect(emsa_emicp_traced(X, n, Y.data(), n, Rb, tb, &p, nullptr, 1) == EMSA_EINVAL,
           "emicp_traced null trace");
    expect(emsa_softassign(X, n, Y.data(), n, Rb, tb, nullptr, 0) == EMSA_EINVAL,
           "softassign null params");
    emsa_params sa = d;
    sa.I1 = 0;
    expect(emsa_softassign(X, n, Y.data(), n, Rb, tb, &sa, 0) == EMSA_EINVAL, "softassign I1 0");
    sa = d;
    expect(emsa_softassign(X, n, Y.data(), n, Rb, tb, &sa, 7) == EMSA_EINVAL,
           "softassign variant");
    expect(emsa_icp(X, n, Y.data(), -1, Rb, tb, &p) == EMSA_EINVAL, "icp M < 0");
    expect(emsa_find_rt_from_s(nullptr, nullptr, nullptr, Rb, tb) == EMSA_EINVAL,
           "findRTfromS null");
    for (int i = 0; i < 9; ++i) expect(Rb[i] == (i % 4 == 0 ? 1.f : 0.f), "R untouched");

    // valid arguments: runs with a GPU, EMSA_ENODEV without one
    const int g = emsa_emicp(X, n, Y.data(), n, Rb, tb, &p);
    expect(g == EMSA_OK || g == EMSA_ENODEV, "emicp valid arguments");
    delete[] X;
    if (failures) return 1;
    std::printf("sanitize OK\n");
    return 0;
}
