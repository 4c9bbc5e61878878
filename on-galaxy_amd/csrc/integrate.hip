// ============================================================================
// integrate.hip -- device-resident velocity Verlet (VelocityVerlet3D,
// Integrator3D.hpp:19-40) for a set of independent systems, and their
// diagnostics (Diagnostics3D.hpp:20-38).
//
// A step of system q (time step dt_q) is the reference's, per component and
// in its operation order (no contraction, so results are bit-reproducible):
//   x <- x + (v*dt + a*(0.5*dt*dt))          drift
//   a' <- forces at the new positions         (one batched evaluation)
//   v <- v + (a + a')*(0.5*dt);  a <- a'      kick
// A system not yet initialised (VelocityVerlet3D::initialize) skips its drift
// and kick and only takes the accelerations of that evaluation, so a new
// system joins the same batched evaluation as running ones.
//
// Bodies of a step are addressed through a job table: batch body k belongs to
// system q with boff[q] <= k < boff[q+1] and is body k - boff[q] of that
// system's own arrays, which stay resident between steps.
// ============================================================================
#include <algorithm>

#include "afmm_internal.h"

#pragma clang fp contract(off)

namespace afmm {
namespace {

__device__ __forceinline__ int job_of(int k, const int* boff, int njobs) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Drift of initialised systems, then gather of positions and masses into the
// batch arrays of the evaluation.
__global__ void k_drift_gather(int n, int njobs, const int* __restrict__ boff, const JobRef* __restrict__ jobs,
                               double* __restrict__ bpos, double* __restrict__ bmass) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int q = job_of(k, boff, njobs);
    const JobRef j = jobs[q];
    const size_t i = (size_t)(k - boff[q]);
    double x[3];
    for (int c = 0; c < 3; ++c) x[c] = j.pos[3 * i + c];
    if (j.init && j.drift) {
        const double half_dt2 = 0.5 * j.dt * j.dt;
        for (int c = 0; c < 3; ++c) {
            const double d = j.vel[3 * i + c] * j.dt + j.acc[3 * i + c] * half_dt2;
            x[c] = x[c] + d;
            j.pos[3 * i + c] = x[c];
        }
    }
    for (int c = 0; c < 3; ++c) bpos[3 * (size_t)k + c] = x[c];
    bmass[k] = j.mass[i];
}

// Kick of initialised systems with the new accelerations; every system keeps
// the new accelerations (and potentials when requested).
__global__ void k_kick(int n, int njobs, const int* __restrict__ boff, const JobRef* __restrict__ jobs,
                       const double* __restrict__ bacc, const double* __restrict__ bphi) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int q = job_of(k, boff, njobs);
    const JobRef j = jobs[q];
    const size_t i = (size_t)(k - boff[q]);
    const double half_dt = 0.5 * j.dt;
    for (int c = 0; c < 3; ++c) {
        const double an = bacc[3 * (size_t)k + c];
        if (j.init && j.drift) j.vel[3 * i + c] = j.vel[3 * i + c] + (j.acc[3 * i + c] + an) * half_dt;
        j.acc[3 * i + c] = an;
    }
    if (bphi && j.phi) j.phi[i] = bphi[k];
}

// Diagnostics3D per system: {kinetic, potential, total, momentum xyz, angular
// momentum xyz}, potential = 0.5 sum m phi (the FMM potential of the current
// positions). One workgroup per system, fixed reduction order.
constexpr int kDiagT = 256;
__global__ void __launch_bounds__(kDiagT) k_diagnostics(const JobRef* __restrict__ jobs, const int* __restrict__ nb,
                                                       double* __restrict__ out) {
    const int q = blockIdx.x;
    const JobRef j = jobs[q];
    const int n = nb[q];
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // ke, pe, P xyz, L xyz
    for (int i = threadIdx.x; i < n; i += kDiagT) {
        const double m = j.mass[i];
        const double* r = j.pos + 3 * (size_t)i;
        const double* v = j.vel + 3 * (size_t)i;
        t[0] += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        t[1] += 0.5 * m * j.phi[i];
        t[2] += m * v[0];
        t[3] += m * v[1];
        t[4] += m * v[2];
        t[5] += m * (r[1] * v[2] - r[2] * v[1]);
        t[6] += m * (r[2] * v[0] - r[0] * v[2]);
        t[7] += m * (r[0] * v[1] - r[1] * v[0]);
    }
    __shared__ double red[8][kDiagT];
    for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = t[c];
    __syncthreads();
    for (int off = kDiagT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int c = 0; c < 8; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = out + 9 * (size_t)q;
        o[0] = red[0][0];
        o[1] = red[1][0];
        o[2] = red[0][0] + red[1][0];
        for (int c = 0; c < 6; ++c) o[3 + c] = red[2 + c][0];
    }
}

}  // namespace

void leapfrog_gather(hipStream_t s, int n, int njobs, const int* d_boff, const JobRef* d_jobs, double* bpos,
                     double* bmass) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_drift_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, njobs, d_boff, d_jobs, bpos,
                       bmass);
    AFMM_CHECK(hipGetLastError());
}

void leapfrog_kick(hipStream_t s, int n, int njobs, const int* d_boff, const JobRef* d_jobs, const double* bacc,
                   const double* bphi) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_kick, dim3((n + 255) / 256), dim3(256), 0, s, n, njobs, d_boff, d_jobs, bacc, bphi);
    AFMM_CHECK(hipGetLastError());
}

void diagnostics(hipStream_t s, int njobs, const JobRef* d_jobs, const int* d_n, double* d_out) {
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_diagnostics, dim3(njobs), dim3(kDiagT), 0, s, d_jobs, d_n, d_out);
    AFMM_CHECK(hipGetLastError());
}

}  // namespace afmm
