// C++ drop-in API of src/3dregistration.h, forwarding to the C ABI (emsa.h).
// Error behaviour: the reference functions return void and ignore CUDA
// errors; here a failure is reported on stderr and h_R, h_t are left as they
// were on entry.
#include "3dregistration.h"

#include <atomic>
#include <cstdio>
#include <vector>

namespace {

std::atomic<int> g_variant{EMSA_SOFTASSIGN_MATERIALISED};

emsa_params to_params(const registrationParameters& p) {
    emsa_params q;
    q.JMAX = p.JMAX;
    q.I0 = p.I0;
    q.I1 = p.I1;
    q.alpha = p.alpha;
    q.T_0 = p.T_0;
    q.TFACTOR = p.TFACTOR;
    q.moutlier = p.moutlier;
    q.sigma_p2 = p.sigma_p2;
    q.sigma_inf = p.sigma_inf;
    q.sigma_factor = p.sigma_factor;
    q.d_02 = p.d_02;
    q.maxIteration = p.maxIteration;
    q.noviewer = p.noviewer;
    q.nostop = p.nostop;
    q.notimer = p.notimer;
    return q;
}

// cloud2data layout (src/cloud2data.cpp:29-47) into a caller-owned vector.
std::vector<float> to_soa(const pcl::PointCloud<pcl::PointXYZ>::Ptr& cloud) {
    const std::size_t n = cloud ? cloud->size() : 0;
    std::vector<float> X(3 * n);
    for (std::size_t i = 0; i < n; ++i) {
        X[i] = cloud->points[i].x;
        X[n + i] = cloud->points[i].y;
        X[2 * n + i] = cloud->points[i].z;
    }
    return X;
}

void report(const char* fn, int rc) {
    if (rc != EMSA_OK) std::fprintf(stderr, "%s failed (%d): %s\n", fn, rc, emsa_last_error());
}

#ifdef EMSA_HAVE_PCL_VIEWER
// The reference redraws the transformed source after every iteration when
// the viewer is on (src/emicp.cu:652-664, src/softassign.cu:805-815,
// src/icp.cpp:272-282); here the per-iteration hook of the C ABI does it.
void redraw(int, int, const float* R, const float* t, void* user) {
    const registrationParameters& param = *static_cast<const registrationParameters*>(user);
    Eigen::Matrix4f T = Eigen::Matrix4f::Identity();
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) T(r, c) = R[3 * r + c];
        T(r, 3) = t[r];
    }
    pcl::transformPointCloud(*param.cloud_source, *param.cloud_source_trans, T);
    param.viewer->updatePointCloud(param.cloud_source_trans, *param.source_trans_color,
                                   "source trans");
    param.viewer->spinOnce();
}

// Installs the redraw for one call when the caller's parameters carry a
// viewer and -noviewer is off (it replaces a hook the caller set on this
// thread for that call, and removes it afterwards).
struct ViewerScope {
    bool on;
    explicit ViewerScope(const registrationParameters& p)
        : on(!p.noviewer && p.viewer && p.cloud_source && p.cloud_source_trans &&
             p.source_trans_color) {
        if (on) emsa_set_iteration_callback(redraw, const_cast<registrationParameters*>(&p));
    }
    ~ViewerScope() {
        if (on) emsa_set_iteration_callback(nullptr, nullptr);
    }
};
#else
// Without PCL's viewer there is nothing to redraw; a hook the caller set with
// emsa_set_iteration_callback stays in effect.
struct ViewerScope {
    explicit ViewerScope(const registrationParameters&) {}
};
#endif

}  // namespace

void emsa_set_softassign_variant(int variant) { g_variant = variant; }

void emicp(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
           const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
           const registrationParameters& param) {
    const std::vector<float> X = to_soa(cloud_target), Y = to_soa(cloud_source);
    const emsa_params p = to_params(param);
    const ViewerScope viewer(param);
    report("emicp", emsa_emicp(X.data(), (int)(X.size() / 3), Y.data(), (int)(Y.size() / 3), h_R,
                               h_t, &p));
}

void emicp_cpu(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
               const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
               const registrationParameters& param) {
    const std::vector<float> X = to_soa(cloud_target), Y = to_soa(cloud_source);
    const emsa_params p = to_params(param);
    const ViewerScope viewer(param);
    const int rc = emsa_emicp_cpu(X.data(), (int)(X.size() / 3), Y.data(), (int)(Y.size() / 3),
                                  h_R, h_t, &p, nullptr, 0);
    report("emicp_cpu", rc < 0 ? rc : EMSA_OK);
}

void icp(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
         const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
         const registrationParameters& param) {
    const std::vector<float> X = to_soa(cloud_target), Y = to_soa(cloud_source);
    const emsa_params p = to_params(param);
    const ViewerScope viewer(param);
    report("icp", emsa_icp(X.data(), (int)(X.size() / 3), Y.data(), (int)(Y.size() / 3), h_R, h_t,
                           &p));
}

void softassign(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
                const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
                const registrationParameters& param) {
    const std::vector<float> X = to_soa(cloud_target), Y = to_soa(cloud_source);
    const emsa_params p = to_params(param);
    const ViewerScope viewer(param);
    report("softassign", emsa_softassign(X.data(), (int)(X.size() / 3), Y.data(),
                                         (int)(Y.size() / 3), h_R, h_t, &p, g_variant.load()));
}

void findRTfromS(const float* h_Xc, const float* h_Yc, const float* h_S, float* h_R,
                 float* h_t) {
    report("findRTfromS", emsa_find_rt_from_s(h_Xc, h_Yc, h_S, h_R, h_t));
}

// src/main.cpp:184-195
void printRT(const float* R, const float* t) {
    std::printf("R\n");
    for (int r = 0; r < 9; r++) {
        std::printf("%f ", R[r]);
        if ((r + 1) % 3 == 0) std::printf("\n");
    }
    std::printf("t\n");
    for (int r = 0; r < 3; r++) std::printf("%f ", t[r]);
    std::printf("\n");
}

// src/cloud2data.cpp:29-47: the caller owns (delete[]) the returned array.
void cloud2data(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud, float** X, int& Xsize) {
    const std::vector<float> v = to_soa(cloud);
    Xsize = (int)(v.size() / 3);
    float* h = new float[v.size() > 0 ? v.size() : 1];
    for (std::size_t i = 0; i < v.size(); ++i) h[i] = v[i];
    *X = h;
}
