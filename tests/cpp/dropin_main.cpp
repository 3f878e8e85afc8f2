// A caller written against the reference API (src/3dregistration.h): builds
// two clouds, runs emicp(), emicp_cpu(), softassign() and icp() through the
// drop-in header and checks that the known rigid transform comes back, and
// that the per-iteration hook sees every ICP iteration. Exit status 0 on
// success.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "3dregistration.h"

int main() {
    pcl::PointCloud<pcl::PointXYZ>::Ptr target(new pcl::PointCloud<pcl::PointXYZ>);
    pcl::PointCloud<pcl::PointXYZ>::Ptr source(new pcl::PointCloud<pcl::PointXYZ>);
    // rotation of 0.2 rad about z, translation (0.03, -0.02, 0.01)
    const float c = std::cos(0.2f), s = std::sin(0.2f);
    const float Rt[9] = {c, -s, 0, s, c, 0, 0, 0, 1}, tt[3] = {0.03f, -0.02f, 0.01f};
    unsigned seed = 12345u;
    auto rnd = [&seed]() { seed = seed * 1664525u + 1013904223u; return (seed >> 8) / 16777216.0f - 0.5f; };
    for (int i = 0; i < 800; ++i) {
        pcl::PointXYZ y(rnd(), rnd(), rnd());
        source->points.push_back(y);
        pcl::PointXYZ x(Rt[0] * y.x + Rt[1] * y.y + Rt[2] * y.z + tt[0],
                        Rt[3] * y.x + Rt[4] * y.y + Rt[5] * y.z + tt[1],
                        Rt[6] * y.x + Rt[7] * y.y + Rt[8] * y.z + tt[2]);
        target->points.push_back(x);
    }
    registrationParameters param;
    param.JMAX = 100; param.I0 = 5; param.I1 = 3; param.alpha = 3.0f; param.T_0 = 100.0f;
    param.TFACTOR = 0.95f; param.moutlier = 1 / sqrtf(param.T_0) * expf(-1.0f);
    param.sigma_p2 = 0.01f; param.sigma_inf = 0.00001f; param.sigma_factor = 0.9f;
    param.d_02 = 0.01f; param.maxIteration = 30;
    param.noviewer = 1; param.nostop = 1; param.notimer = 1; param.argc = 0; param.argv = 0;

    float h_R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, h_t[3] = {0, 0, 0};
    emicp(target, source, h_R, h_t, param);
    printRT(h_R, h_t);
    double err = 0;
    for (int i = 0; i < 9; ++i) err = std::fmax(err, std::fabs(h_R[i] - Rt[i]));
    for (int i = 0; i < 3; ++i) err = std::fmax(err, std::fabs(h_t[i] - tt[i]));
    std::printf("emicp max |error| %g\n", err);
    if (!(err < 1e-4)) return 1;

    // emicp_cpu(): the same registration on the host CPU in FP32
    float c_R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, c_t[3] = {0, 0, 0};
    emicp_cpu(target, source, c_R, c_t, param);
    double errc = 0;
    for (int i = 0; i < 9; ++i) errc = std::fmax(errc, std::fabs(c_R[i] - Rt[i]));
    for (int i = 0; i < 3; ++i) errc = std::fmax(errc, std::fabs(c_t[i] - tt[i]));
    std::printf("emicp_cpu max |error| %g\n", errc);
    if (!(errc < 1e-4)) return 6;

    float* X = nullptr;
    int Xsize = 0;
    cloud2data(target, &X, Xsize);
    const bool layout_ok = Xsize == 800 && X[0] == target->points[0].x && X[800] == target->points[0].y;
    delete[] X;
    if (!layout_ok) return 2;
    boost::shared_array<float> Xs;
    int Xs_size = 0;
    cloud2data(source, Xs, Xs_size);
    if (!(Xs_size == 800 && Xs[1] == source->points[1].x && Xs[1600 + 1] == source->points[1].z))
        return 2;

    // findRTfromS on the exact correspondence
    float Xc[3] = {0, 0, 0}, Yc[3] = {0, 0, 0}, S[9] = {0};
    for (int i = 0; i < 800; ++i) {
        Xc[0] += target->points[i].x / 800; Xc[1] += target->points[i].y / 800; Xc[2] += target->points[i].z / 800;
        Yc[0] += source->points[i].x / 800; Yc[1] += source->points[i].y / 800; Yc[2] += source->points[i].z / 800;
    }
    for (int i = 0; i < 800; ++i) {
        const float xt[3] = {target->points[i].x - Xc[0], target->points[i].y - Xc[1], target->points[i].z - Xc[2]};
        const float ys[3] = {source->points[i].x - Yc[0], source->points[i].y - Yc[1], source->points[i].z - Yc[2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) S[3 * a + b] += ys[a] * xt[b];  // h_S[3a+b] = sum source_a target_b
    }
    float R2[9], t2[3];
    findRTfromS(Xc, Yc, S, R2, t2);
    double err2 = 0;
    for (int i = 0; i < 9; ++i) err2 = std::fmax(err2, std::fabs(R2[i] - Rt[i]));
    std::printf("findRTfromS max |error| %g\n", err2);
    if (!(err2 < 1e-5)) return 3;

    // softassign(), both variants: no slack (moutlier 0) and Sinkhorn run to
    // convergence (I1 = 30), annealed until second-nearest partners no longer
    // count, so the soft correspondences become the exact permutation and the
    // known transform comes back
    const int variants[2] = {EMSA_SOFTASSIGN_MATERIALISED, EMSA_SOFTASSIGN_LOGDOMAIN};
    for (int v = 0; v < 2; ++v) {
        float s_R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, s_t[3] = {0, 0, 0};
        param.T_0 = 0.5f; param.TFACTOR = 0.8f; param.alpha = 0.0f; param.JMAX = 48;
        param.I0 = 2; param.I1 = 30; param.moutlier = 0.0f;
        emsa_set_softassign_variant(variants[v]);
        softassign(target, source, s_R, s_t, param);
        printRT(s_R, s_t);
        double err4 = 0;
        for (int i = 0; i < 9; ++i) err4 = std::fmax(err4, std::fabs(s_R[i] - Rt[i]));
        for (int i = 0; i < 3; ++i) err4 = std::fmax(err4, std::fabs(s_t[i] - tt[i]));
        std::printf("softassign (variant %d) max |error| %g\n", variants[v], err4);
        if (!(err4 < 1e-4)) return 4;
    }
    emsa_set_softassign_variant(EMSA_SOFTASSIGN_MATERIALISED);

    // icp(): noiseless pair, every source point has its exact partner; a
    // viewer-style hook sees every iteration's transform, the last one equal
    // to the result
    struct Seen {
        int n = 0, alg = 0;
        float R[9], t[3];
    } seen;
    emsa_set_iteration_callback(
        [](int alg, int, const float* R, const float* t, void* u) {
            Seen* s = static_cast<Seen*>(u);
            ++s->n;
            s->alg = alg;
            for (int i = 0; i < 9; ++i) s->R[i] = R[i];
            for (int i = 0; i < 3; ++i) s->t[i] = t[i];
        },
        &seen);
    float i_R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, i_t[3] = {0, 0, 0};
    param.maxIteration = 60;
    icp(target, source, i_R, i_t, param);
    emsa_set_iteration_callback(nullptr, nullptr);
    std::printf("iteration hook: %d calls\n", seen.n);
    if (seen.n != 60 || seen.alg != EMSA_ALG_ICP) return 7;
    for (int i = 0; i < 9; ++i)
        if (seen.R[i] != i_R[i]) return 7;
    for (int i = 0; i < 3; ++i)
        if (seen.t[i] != i_t[i]) return 7;
    printRT(i_R, i_t);
    double err3 = 0;
    for (int i = 0; i < 9; ++i) err3 = std::fmax(err3, std::fabs(i_R[i] - Rt[i]));
    for (int i = 0; i < 3; ++i) err3 = std::fmax(err3, std::fabs(i_t[i] - tt[i]));
    std::printf("icp max |error| %g\n", err3);
    if (!(err3 < 1e-4)) return 5;
    std::printf("dropin OK\n");
    return 0;
}
