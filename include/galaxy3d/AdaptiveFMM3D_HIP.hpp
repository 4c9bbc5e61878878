// Header-only ForceSolver3D over the C ABI (include/afmm.h): the MI355X
// replacement for galaxy3d::AdaptiveFMM3D (src3d/AdaptiveFMM3D.hpp:20-133) and
// for the ad-hoc wrappers of gpu::adaptiveFmm3dGpu (gpu/galaxy_sim.cu:41-52,
// gpu/adaptive_gpu_harness.cu:95-102).
//
//   galaxy3d::AdaptiveFMM3D_HIP solver(KernelParams3D{G, eps}, p, threshold);
//   Accelerations3D a = solver.computeAccelerations(bodies);
//
// Same constructor arguments and result (input order, G*(far + near)). The
// reference has no error channel; errors here throw std::runtime_error. The
// device context (buffers, stream, tables) lives as long as the solver, so
// repeated calls reuse device memory. Like the reference (whose stats are
// `mutable`), one solver must not be used from two threads at once.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../afmm.h"
#include "ForceSolver3D.hpp"

namespace galaxy3d {

class AdaptiveFMM3D_HIP final : public ForceSolver3D {
public:
    AdaptiveFMM3D_HIP(KernelParams3D params, int p, int threshold = 64, int device = 0,
                      int precision = AFMM_PRECISION_FP64)
        : params_(params), precision_(precision) {
        afmm_params prm{params.G, params.eps, p, threshold, precision, device};
        afmm_ctx* c = nullptr;
        const int rc = afmm_create(&prm, &c);
        if (rc) throw std::runtime_error(std::string("afmm_create: ") + afmm_strerror(rc));
        ctx_.reset(c);
    }

    Accelerations3D computeAccelerations(const Bodies3D& bodies) const override {
        static_assert(sizeof(Vec3) == 3 * sizeof(double) && std::is_standard_layout<Vec3>::value,
                      "Vec3 must be three packed doubles");
        const std::size_t n = bodies.size();
        Accelerations3D acc(n, Vec3{});
        if (n == 0) return acc;  // AdaptiveFMM3D.hpp:34
        const int rc = afmm_eval(ctx_.get(), (int64_t)n, reinterpret_cast<const double*>(bodies.pos.data()),
                                 bodies.mass.data(), reinterpret_cast<double*>(acc.data()));
        if (rc)
            throw std::runtime_error(std::string("afmm_eval: ") + afmm_strerror(rc) + ": " +
                                     afmm_last_error(ctx_.get()));
        afmm_stats st{};
        afmm_get_stats(ctx_.get(), &st);
        lastMaxLeafOcc = st.max_leaf_occ;  // AdaptiveFMM3D.hpp:26-27
        lastNumBoxes = (int)st.n_boxes;
        return acc;
    }

    const char* name() const override {
        switch (precision_) {
            case AFMM_PRECISION_MIXED: return "AdaptiveFMM3D(CGR, MI355X HIP mixed split-FP32)";
            case AFMM_PRECISION_MIXED_BF16: return "AdaptiveFMM3D(CGR, MI355X HIP mixed bf16-split)";
            case AFMM_PRECISION_M2L_BF16: return "AdaptiveFMM3D(CGR, MI355X HIP bf16-split M2L, FP64 P2P)";
            case AFMM_PRECISION_P2P_FP32: return "AdaptiveFMM3D(CGR, MI355X HIP FP64 M2L, FP32 P2P)";
            default: return "AdaptiveFMM3D(CGR, MI355X HIP FP64)";
        }
    }

    afmm_ctx* context() const { return ctx_.get(); }

    mutable int lastMaxLeafOcc = 0;
    mutable int lastNumBoxes = 0;

private:
    struct Deleter {
        void operator()(afmm_ctx* c) const { afmm_destroy(c); }
    };
    KernelParams3D params_;
    int precision_;
    std::unique_ptr<afmm_ctx, Deleter> ctx_;
};

}  // namespace galaxy3d
