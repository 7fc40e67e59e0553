// Sanitizer run of the host-side kernel classes (host/kernels.cpp), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/native/Makefile.
#include <cmath>
#include <cstdio>
#include <memory>

#include "kernels.h"

int main() {
    using namespace sph;
    const KernelType types[] = {KernelType::CUBIC_SPLINE, KernelType::WENDLAND_C2,
                                KernelType::GAUSSIAN};
    double acc = 0.0;
    for (KernelType t : types) {
        std::unique_ptr<Kernel> k = create_kernel(t, 0.02f);
        for (int i = 0; i < 1000; ++i) {
            const float x = 0.05f * (float)(i - 500) / 500.0f;
            const vec3 r(x, 0.5f * x, -0.25f * x);
            const vec3 g = k->gradW(r);
            acc += k->W(r) + g.x + g.y + g.z + k->laplacianW(r);
        }
        k->set_smoothing_length(0.03f);
        acc += k->W(vec3(0.0f)) + k->get_support_radius();
    }
    CubicSplineKernel cubic(0.5f);
    acc += cubic.W_2D(vec2(0.1f, 0.2f)) + cubic.W_1D(-0.3f);
    acc += kernel_utils::estimate_smoothing_length(0.01f, 60.0f);
    acc += kernel_utils::within_support(vec3(0.01f), 0.02f) ? 1.0 : 0.0;
    if (!std::isfinite(acc)) {
        std::fprintf(stderr, "non-finite kernel values\n");
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
