// ============================================================================
// direct.hip -- FP64 direct summation onto a target subset: the accuracy
// yardstick of DirectSolver3D (DirectSolver3D.hpp:15-34) and the softened
// pair potential of Kernel3D.hpp:39-42, on the device.
//
//   acc[k] = G sum_{j != t_k} m_j d / (|d|^2 + eps^2)^{3/2},  d = x_j - x_{t_k}
//   phi[k] = -G sum_{j != t_k} m_j / sqrt(|d|^2 + eps^2)
//
// Grid: (target blocks of 64) x (source splits). A wave holds 64 targets in
// lanes and streams its split's sources through LDS 64 at a time; the
// per-split partial sums are then added in split order, so results do not
// depend on scheduling.
//
// The partial sums are the only scratch memory. direct_targets takes them from
// the context's pool; direct_targets_noalloc takes them from `scratch`, a
// static device array every GPU holds for the life of the process, so that a
// caller that cannot allocate (graph capture, serving loops) can run it. Both
// use the same split count for the same (n, ntgt) and give the same bits.
// ============================================================================
#include <algorithm>
#include <mutex>

#include "afmm_internal.h"

namespace afmm {
namespace {

// Static scratch of direct_targets_noalloc: the per-split partial sums of up
// to kNoAllocMaxTargets targets (split count as in split_count below, at most
// 64 * (4096 + 1024) target-splits of 4 doubles: 10 MiB).
constexpr size_t kScratchDoubles = (size_t)64 * (4096 + kNoAllocMaxTargets / 64) * 4;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Source splits of the direct sum: enough blocks to fill the GPU (about 4096
// target-block x split pairs), at least 256 sources per split.
int split_count(int n, int ntgt) {
    const int ntb = (ntgt + 63) / 64;
    return std::max(1, std::min((n + 255) / 256, (4096 + ntb - 1) / ntb));
}

// Target indices outside [0, n) (only possible on the allocation-free entry,
// whose indices are device data the host never sees) read nothing and give NaN.
template <class I>
__global__ void __launch_bounds__(64) k_direct(int n, const double* __restrict__ pos,
                                              const double* __restrict__ mass, int ntgt,
                                              const I* __restrict__ tgt, int chunk, double eps2,
                                              double* __restrict__ part) {
    __shared__ double sx[64], sy[64], sz[64], sm[64];
    const int lane = threadIdx.x;
    const int k = blockIdx.x * 64 + lane;
    const int split = blockIdx.y;
    const int j0 = split * chunk, j1 = min(n, j0 + chunk);
    const I ti = k < ntgt ? tgt[k] : I(-1);
    const bool act = ti >= 0 && ti < n;
    const int i = act ? (int)ti : -1;
    const double xi = act ? pos[3 * (size_t)i] : 0.0, yi = act ? pos[3 * (size_t)i + 1] : 0.0,
                 zi = act ? pos[3 * (size_t)i + 2] : 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0, ph = 0.0;
    for (int jb = j0; jb < j1; jb += 64) {
        const int j = jb + lane;
        if (j < j1) {
            sx[lane] = pos[3 * (size_t)j];
            sy[lane] = pos[3 * (size_t)j + 1];
            sz[lane] = pos[3 * (size_t)j + 2];
            sm[lane] = mass[j];
        }
        __syncthreads();
        const int cnt = min(64, j1 - jb);
        for (int q = 0; q < cnt; ++q) {
            const double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double r = sqrt(r2);
            const double inv = 1.0 / (r2 * r);
            if (jb + q != i) {
                const double s = sm[q] * inv;
                ax += s * dx;
                ay += s * dy;
                az += s * dz;
                ph -= sm[q] / r;
            }
        }
        __syncthreads();
    }
    if (k < ntgt) {
        const double nan = __builtin_nan("");
        double* o = part + ((size_t)split * ntgt + k) * 4;
        o[0] = act ? ax : nan;
        o[1] = act ? ay : nan;
        o[2] = act ? az : nan;
        o[3] = act ? ph : nan;
    }
}

__global__ void k_direct_reduce(int ntgt, int nsplit, double G, const double* __restrict__ part,
                                double* __restrict__ acc, double* __restrict__ phi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntgt) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < nsplit; ++q)
        for (int c = 0; c < 4; ++c) s[c] += part[((size_t)q * ntgt + k) * 4 + c];
    acc[3 * (size_t)k] = G * s[0];
    acc[3 * (size_t)k + 1] = G * s[1];
    acc[3 * (size_t)k + 2] = G * s[2];
    if (phi) phi[k] = G * s[3];
}

template <class I>
void launch_direct(hipStream_t s, const double* d_pos, const double* d_mass, int n, const I* d_tgt, int ntgt, double G,
                   double eps, double* part, double* d_acc, double* d_phi) {
    const int ntb = (ntgt + 63) / 64;
    const int nsplit = split_count(n, ntgt);
    const int chunk = (n + nsplit - 1) / nsplit;
    hipLaunchKernelGGL(k_direct<I>, dim3(ntb, nsplit), dim3(64), 0, s, n, d_pos, d_mass, ntgt, d_tgt, chunk,
                       eps * eps, part);
    hipLaunchKernelGGL(k_direct_reduce, dim3((ntgt + 255) / 256), dim3(256), 0, s, ntgt, nsplit, G, part, d_acc,
                       d_phi);
    AFMM_CHECK(hipGetLastError());
}

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

void direct_targets(Pool& pool, hipStream_t s, const double* d_pos, const double* d_mass, int n, const int* d_tgt,
                    int ntgt, double G, double eps, double* d_acc, double* d_phi) {
    if (ntgt <= 0 || n <= 0) return;
    double* part = pool.get<double>("direct_part", (size_t)split_count(n, ntgt) * ntgt * 4);
    launch_direct(s, d_pos, d_mass, n, d_tgt, ntgt, G, eps, part, d_acc, d_phi);
}

void direct_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    // the caller has made `device` current; the symbol's address is per device
    void* p = nullptr;
    AFMM_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(p);
}

void direct_targets_noalloc(int device, hipStream_t s, const double* d_pos, const double* d_mass, int n,
                            const int64_t* d_tgt, int ntgt, double G, double eps, double* d_acc, double* d_phi) {
    if (ntgt <= 0 || n <= 0) return;
    if (ntgt > kNoAllocMaxTargets) throw std::runtime_error("more targets than the static scratch holds");
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        part = device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
    }
    if (!part) throw std::runtime_error("static scratch not set up on this device");
    if ((size_t)split_count(n, ntgt) * ntgt * 4 > kScratchDoubles)
        throw std::runtime_error("static scratch too small for this split");
    launch_direct(s, d_pos, d_mass, n, d_tgt, ntgt, G, eps, part, d_acc, d_phi);
}

}  // namespace afmm
