/*
 * Drop-in replacement for the reference header src/3dregistration.h
 * (tttamaki/cuda_emicp_softassign): same functions, same signatures, same
 * registrationParameters fields in the same order. The GPU functions forward
 * to the C ABI of libemsa (emsa.h), which runs hand-written HIP kernels on
 * an AMD Instinct MI355X (gfx950).
 *
 * With PCL available (and EMSA_NO_PCL not defined) the real
 * pcl::PointCloud<pcl::PointXYZ> and the viewer members are used. Without
 * PCL a minimal pcl::PointXYZ / pcl::PointCloud with the members the API
 * touches (points, size(), Ptr) is provided so that callers compile
 * unchanged.
 */
#ifndef _3DREGISTRATION_H_
#define _3DREGISTRATION_H_

#include <cstddef>
#include <memory>
#include <vector>

#include "emsa.h"

#if !defined(EMSA_NO_PCL) && defined(__has_include)
#if __has_include(<pcl/point_cloud.h>) && __has_include(<pcl/point_types.h>)
#include <pcl/point_cloud.h>
#include <pcl/point_types.h>
#define EMSA_HAVE_PCL 1
#if __has_include(<pcl/visualization/cloud_viewer.h>)
#include <pcl/common/transforms.h>
#include <pcl/visualization/cloud_viewer.h>
#define EMSA_HAVE_PCL_VIEWER 1
#endif
#endif
#endif

#ifndef EMSA_HAVE_PCL
namespace boost {
// The members of boost::shared_array that cloud2data's second overload uses.
template <class T>
class shared_array {
    std::shared_ptr<T> p_;

public:
    shared_array() = default;
    explicit shared_array(T* p) : p_(p, std::default_delete<T[]>()) {}
    void reset(T* p = nullptr) { p_.reset(p, std::default_delete<T[]>()); }
    T* get() const { return p_.get(); }
    T& operator[](std::ptrdiff_t i) const { return p_.get()[i]; }
    explicit operator bool() const { return (bool)p_; }
};
}  // namespace boost
namespace pcl {
struct PointXYZ {
    float x, y, z;
    PointXYZ() : x(0), y(0), z(0) {}
    PointXYZ(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};
template <class PointT>
struct PointCloud {
    typedef std::shared_ptr<PointCloud<PointT> > Ptr;
    typedef std::shared_ptr<const PointCloud<PointT> > ConstPtr;
    std::vector<PointT> points;
    std::size_t size() const { return points.size(); }
    void push_back(const PointT& p) { points.push_back(p); }
};
}  // namespace pcl
#endif

// threads in a 2D block of the reference kernels (src/3dregistration.h:36);
// kept for source compatibility, unused by the HIP kernels.
#define BLOCK_SIZE 16

typedef struct {
    // for softassign
    int JMAX;       // annealing iterations, default 100
    int I0;         // iterations per temperature, default 5
    int I1;         // Sinkhorn row/column normalisations, default 3
    float alpha;    // outlier parameter, default 3.0
    float T_0;      // initial temperature, default 100.0
    float TFACTOR;  // temperature factor, default 0.95
    float moutlier; // slack row/column value, default 1/sqrtf(T_0)*expf(-1.0f)

    // for EM-ICP
    float sigma_p2;     // initial sigma^2, default 0.01
    float sigma_inf;    // minimum sigma^2, default 0.00001
    float sigma_factor; // sigma^2 factor, default 0.9
    float d_02;         // outlier term, default 0.01

    // for ICP
    int maxIteration;  // default 30

    // misc
    int noviewer;
    int nostop;
    int notimer;

    int argc;
    char** argv;

#ifdef EMSA_HAVE_PCL_VIEWER
    boost::shared_ptr<pcl::visualization::PCLVisualizer> viewer;
    pcl::visualization::PointCloudColorHandlerCustom<pcl::PointXYZ>::Ptr source_trans_color;
    pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source;
    pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source_trans;
    pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target;
#endif
} registrationParameters;

// Selects the Softassign variant used by softassign() (not in the reference):
// EMSA_SOFTASSIGN_MATERIALISED (default) or EMSA_SOFTASSIGN_LOGDOMAIN.
void emsa_set_softassign_variant(int variant);

void softassign(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
                const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
                const registrationParameters& param);
void emicp(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
           const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
           const registrationParameters& param);
// EM-ICP on the host CPU in FP32 (src/emicp_cpu.cpp:186-547).
void emicp_cpu(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
               const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
               const registrationParameters& param);
// ICP with exact nearest-neighbour correspondences (src/icp.cpp:122-289,
// declared at src/3dregistration.h:96-99): param.maxIteration iterations.
void icp(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_target,
         const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud_source, float* h_R, float* h_t,
         const registrationParameters& param);

void findRTfromS(const float* h_Xc, const float* h_Yc, const float* h_S, float* h_R, float* h_t);

void printRT(const float* R, const float* t);

void cloud2data(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud, float** X, int& Xsize);

// src/cloud2data.cpp:50-67: the SoA copy owned by a shared_array. Inline, so
// that it works with Boost's shared_array (PCL builds) as with the one above.
inline void cloud2data(const pcl::PointCloud<pcl::PointXYZ>::Ptr cloud,
                       boost::shared_array<float>& X, int& Xsize) {
    Xsize = cloud ? (int)cloud->size() : 0;
    float* h = new float[Xsize > 0 ? 3 * (std::size_t)Xsize : 1];
    for (int i = 0; i < Xsize; ++i) {
        h[i] = cloud->points[i].x;
        h[Xsize + i] = cloud->points[i].y;
        h[2 * Xsize + i] = cloud->points[i].z;
    }
    X.reset(h);
}

#define CUDA_SAFE_CALL(a) a
#define CUT_SAFE_CALL(a) a

#endif
