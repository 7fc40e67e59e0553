// kernels.h -- smoothing-kernel classes with the reference's API (src/kernels.h:8-129).
//
// The GPU step evaluates the cubic spline inside its own kernels; these host classes are the
// reference's public kernel API (create_kernel, W / gradW / laplacianW), for callers that
// evaluate kernels themselves.  Expressions follow src/kernels.cpp in FP32 as written.
#pragma once

#include <memory>

#include "vec3.h"

namespace sph {

// Two-float vector for the 2-D cubic spline (glm::vec2's layout).
struct vec2 {
    float x = 0.0f, y = 0.0f;
    constexpr vec2() = default;
    constexpr vec2(float a, float b) : x(a), y(b) {}
};

class Kernel {
   protected:
    float h_;
    float h_inv_;
    float h_sq_;
    float norm_ = 0.0f;

   public:
    explicit Kernel(float h);
    virtual ~Kernel() = default;

    virtual float W(const vec3& r) const = 0;
    virtual vec3 gradW(const vec3& r) const = 0;
    virtual float laplacianW(const vec3& r) const = 0;

    float get_smoothing_length() const { return h_; }
    void set_smoothing_length(float h);
    float get_support_radius() const { return 2.0f * h_; }
    float get_cutoff_radius() const { return get_support_radius(); }
};

// Cubic spline (src/kernels.cpp:23-164): sigma_3d = 1/(pi h^3), sigma_2d = 10/(7 pi h^2),
// sigma_1d = 2/(3 h).
class CubicSplineKernel : public Kernel {
    float sigma_1d_, sigma_2d_, sigma_3d_;

   public:
    explicit CubicSplineKernel(float h);
    float W(const vec3& r) const override;
    vec3 gradW(const vec3& r) const override;
    float laplacianW(const vec3& r) const override;
    float W_3D(const vec3& r) const { return W(r); }
    float W_2D(const vec2& r) const;
    float W_1D(float r) const;
};

// Wendland C2 (src/kernels.cpp:166-199): norm 21/(2 pi h^3), (1 - q/2)^4 (2q + 1) on q < 2.
class WendlandC2Kernel : public Kernel {
    float norm_factor_;

   public:
    explicit WendlandC2Kernel(float h);
    float W(const vec3& r) const override;
    vec3 gradW(const vec3& r) const override;
    float laplacianW(const vec3& r) const override;
};

// Gaussian (src/kernels.cpp:201-222): (pi h^2)^(-3/2) exp(-|r|^2/h^2), unbounded support.
class GaussianKernel : public Kernel {
    float sigma_sq_inv_;

   public:
    explicit GaussianKernel(float h);
    float W(const vec3& r) const override;
    vec3 gradW(const vec3& r) const override;
    float laplacianW(const vec3& r) const override;
};

enum class KernelType { CUBIC_SPLINE, WENDLAND_C2, GAUSSIAN };

// src/kernels.cpp:224-236; throws std::invalid_argument for an unknown type.
std::unique_ptr<Kernel> create_kernel(KernelType type, float smoothing_length);

namespace kernel_utils {

// q = |r| / h (src/kernels.h:117-119)
inline float compute_q(const vec3& r, float h) { return length(r) / h; }

// |r|^2 <= (f h)^2 (src/kernels.h:122-124)
inline bool within_support(const vec3& r, float h, float support_factor = 2.0f) {
    return dot(r, r) <= (support_factor * h) * (support_factor * h);
}

// spacing * (neighbours / 50)^(1/3) (src/kernels.cpp:296-300)
float estimate_smoothing_length(float particle_spacing, float neighbor_count = 50.0f);

}  // namespace kernel_utils

}  // namespace sph
