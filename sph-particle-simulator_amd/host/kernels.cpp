// kernels.cpp -- smoothing kernels of the reference's public API (src/kernels.cpp), FP32.
#include "kernels.h"

#include <cmath>
#include <stdexcept>

namespace sph {

namespace {

constexpr float kPi = 3.14159265358979323846f;

// Cubic-spline shape and its derivative terms for q = |r|/h, times sigma
// (src/kernels.cpp:40-66, 68-148).
float cubic_w(float sigma, float q) {
    if (q >= 0.0f && q <= 1.0f) return sigma * (2.0f / 3.0f - q * q + 0.5f * q * q * q);
    if (q > 1.0f && q <= 2.0f) {
        const float t = 2.0f - q;
        return sigma * (1.0f / 6.0f * t * t * t);
    }
    return 0.0f;
}

vec3 cubic_grad(float sigma, float h, const vec3& r, float q) {
    const float r_len = length(r);
    if (r_len < 1e-6f) return vec3(0.0f);
    vec3 g(0.0f);
    if (q >= 0.0f && q <= 1.0f) {
        g = sigma * (-2.0f * q + 1.5f * q * q) * (r / r_len);
    } else if (q > 1.0f && q <= 2.0f) {
        const float t = 2.0f - q;
        g = -sigma * (0.5f * t * t) * (r / r_len);
    }
    return g / h;  // dW/dr = dW/dq * r / (|r| h)
}

float cubic_lap(float sigma, float h_sq, float q) {
    if (q >= 0.0f && q <= 1.0f) return sigma * (-2.0f + 3.0f * q) / h_sq;
    if (q > 1.0f && q <= 2.0f) return sigma * (2.0f - q) / h_sq;
    return 0.0f;
}

}  // namespace

Kernel::Kernel(float h) : h_(h), h_inv_(1.0f / h), h_sq_(h * h) {}

void Kernel::set_smoothing_length(float h) {
    h_ = h;
    h_inv_ = 1.0f / h;
    h_sq_ = h * h;
}

CubicSplineKernel::CubicSplineKernel(float h) : Kernel(h) {
    sigma_3d_ = 1.0f / (kPi * h * h * h);
    sigma_2d_ = 10.0f / (7.0f * kPi * h * h);
    sigma_1d_ = 2.0f / (3.0f * h);
    norm_ = sigma_3d_;
}

float CubicSplineKernel::W(const vec3& r) const {
    return cubic_w(sigma_3d_, kernel_utils::compute_q(r, h_));
}

vec3 CubicSplineKernel::gradW(const vec3& r) const {
    return cubic_grad(sigma_3d_, h_, r, kernel_utils::compute_q(r, h_));
}

float CubicSplineKernel::laplacianW(const vec3& r) const {
    return cubic_lap(sigma_3d_, h_sq_, kernel_utils::compute_q(r, h_));
}

float CubicSplineKernel::W_2D(const vec2& r) const {
    // glm::length of a vec2: sqrt(x*x + y*y)
    return cubic_w(sigma_2d_, std::sqrt(r.x * r.x + r.y * r.y) / h_);
}

float CubicSplineKernel::W_1D(float r) const { return cubic_w(sigma_1d_, std::fabs(r) / h_); }

WendlandC2Kernel::WendlandC2Kernel(float h) : Kernel(h) {
    norm_factor_ = 21.0f / (2.0f * kPi * h * h * h);
}

float WendlandC2Kernel::W(const vec3& r) const {
    const float q = kernel_utils::compute_q(r, h_);
    if (q >= 2.0f) return 0.0f;
    const float t = 1.0f - 0.5f * q;
    return norm_factor_ * t * t * t * t * (2.0f * q + 1.0f);
}

vec3 WendlandC2Kernel::gradW(const vec3& r) const {
    const float r_len = length(r);
    if (r_len < 1e-6f) return vec3(0.0f);
    const float q = r_len / h_;
    if (q >= 2.0f) return vec3(0.0f);
    const float t = 1.0f - 0.5f * q;
    const float dw_dq = -5.0f * t * t * t * q;
    return norm_factor_ * dw_dq * (r / (r_len * h_));
}

float WendlandC2Kernel::laplacianW(const vec3& r) const {
    const float q = kernel_utils::compute_q(r, h_);
    if (q >= 2.0f) return 0.0f;
    const float t = 1.0f - 0.5f * q;
    return norm_factor_ * (5.0f / h_sq_) * t * t * (5.0f * q - 3.0f);
}

GaussianKernel::GaussianKernel(float h) : Kernel(h) {
    sigma_sq_inv_ = 1.0f / (h * h);
    norm_ = 1.0f / std::pow(kPi * h * h, 1.5f);
}

float GaussianKernel::W(const vec3& r) const {
    return norm_ * std::exp(-dot(r, r) * sigma_sq_inv_);
}

vec3 GaussianKernel::gradW(const vec3& r) const {
    const float e = std::exp(-dot(r, r) * sigma_sq_inv_);
    return -2.0f * sigma_sq_inv_ * norm_ * e * r;
}

float GaussianKernel::laplacianW(const vec3& r) const {
    const float r_sq = dot(r, r);
    const float e = std::exp(-r_sq * sigma_sq_inv_);
    return 2.0f * sigma_sq_inv_ * norm_ * e * (2.0f * sigma_sq_inv_ * r_sq - 3.0f);
}

std::unique_ptr<Kernel> create_kernel(KernelType type, float smoothing_length) {
    switch (type) {
        case KernelType::CUBIC_SPLINE:
            return std::make_unique<CubicSplineKernel>(smoothing_length);
        case KernelType::WENDLAND_C2:
            return std::make_unique<WendlandC2Kernel>(smoothing_length);
        case KernelType::GAUSSIAN:
            return std::make_unique<GaussianKernel>(smoothing_length);
    }
    throw std::invalid_argument("Unknown kernel type");
}

namespace kernel_utils {

float estimate_smoothing_length(float particle_spacing, float neighbor_count) {
    return particle_spacing * std::pow(neighbor_count / 50.0f, 1.0f / 3.0f);
}

}  // namespace kernel_utils

}  // namespace sph
