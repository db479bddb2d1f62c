// admm.hip -- the vector steps of the ADMM reconstruction (TiltRecZ-mpi's
// ADMM, main-mpiz.cpp:603-1013; CuADMMZ, recon_z.cu:362-491), on volumes
// in the row-blocked layout.  Rows past Y in the last row block stay zero
// through every step (zero data, zero start), so sums over whole buffers
// equal sums over the volume.
//
// The operator op(x) = sum_theta A^T [(A x) / w] + mu x is the forward
// projection with the ray-weight quotient (k_forward_g on a zero stack)
// followed by k_back<float, BACK_LIN>; the CG scalars stay on the device,
// so a whole reconstruction is one stream of launches.  Dot products sum
// float products in double, in a fixed order; alpha and beta come from
// them as the reference computes them, and the updates use them in double.
#include <hip/hip_runtime.h>

#include "trec_internal.h"

namespace trec {

// Scalars on the device: S[0] = r.r, S[1] = h.p, S[2] = alpha, S[3] = r.h,
// S[4] = beta.
__global__ __launch_bounds__(WG) void k_admm_vec(AdmmVecs a, int step)
{
#pragma clang fp contract(off)
    const double alpha = a.S[2], beta = a.S[4];
    const float mu = a.mu, soft = a.soft;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < a.n; i += (size_t)gridDim.x * WG) {
        switch (step) {
        case ADMM_HTB: a.htb[i] = a.htb[i] + (a.u[i] - a.d[i]) * mu; break;
        case ADMM_R0: {
            const float r = a.htb[i] - a.ax[i];
            a.r[i] = r;
            a.p[i] = r;
            break;
        }
        case ADMM_STEP: a.vol[i] = (float)((double)a.vol[i] + alpha * (double)a.p[i]); break;
        case ADMM_RES: a.r[i] = a.htb[i] - a.ax[i]; break;
        case ADMM_DIR: a.p[i] = (float)((double)a.r[i] + beta * (double)a.p[i]); break;
        default: {  // ADMM_SOFT: u = soft(vol + d); d += vol - u
            const float v = a.vol[i], t = v + a.d[i];
            const float u = t < -soft ? t + soft : (t > soft ? t - soft : 0.f);
            a.u[i] = u;
            a.d[i] = a.d[i] + (v - u);
            break;
        }
        }
    }
}

// Partial dot products of one pass: work-group b writes sum x0.y0 (and
// sum x1.y1 when x1 is given) over its grid-stride share, float products
// summed in double, then a fixed tree.
__global__ __launch_bounds__(WG) void k_admm_dot(const float *__restrict__ x0, const float *__restrict__ y0,
                                                 const float *__restrict__ x1, const float *__restrict__ y1,
                                                 size_t n, double *__restrict__ part)
{
#pragma clang fp contract(off)
    double s0 = 0.0, s1 = 0.0;
    for (size_t i = (size_t)blockIdx.x * WG + threadIdx.x; i < n; i += (size_t)gridDim.x * WG) {
        const float p0 = x0[i] * y0[i];
        s0 += (double)p0;
        if (x1) {
            const float p1 = x1[i] * y1[i];
            s1 += (double)p1;
        }
    }
    __shared__ double red[2][WG];
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    __syncthreads();
#pragma unroll
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = red[0][0];
        part[gridDim.x + blockIdx.x] = red[1][0];
    }
}

// which = 0: S[0] = r.r, S[1] = h.p, alpha; which = 1: S[3] = r.h, beta.
// alpha and beta are 0 unless |h.p| > 10e-6 (main-mpiz.cpp:880-888, 902-909).
__global__ __launch_bounds__(64) void k_admm_coef(const double *__restrict__ part, int nb, int which,
                                                  double *__restrict__ S)
{
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += 64) {
        a += part[k];
        b += part[nb + k];
    }
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) {
        a += __shfl_down(a, w, 64);
        b += __shfl_down(b, w, 64);
    }
    if (threadIdx.x == 0) {
        if (which == 0) {
            S[0] = a;
            S[1] = b;
            S[2] = (b > 10e-6 || -b > 10e-6) ? a / b : 0.0;
        } else {
            const double d2 = S[1];
            S[3] = a;
            S[4] = (d2 > 10e-6 || -d2 > 10e-6) ? a / d2 : 0.0;
        }
    }
}

static unsigned vec_blocks(size_t n) { return (unsigned)std::min<size_t>((n + WG - 1) / WG, 8192); }

hipError_t launch_admm_vec(const AdmmVecs &a, int step, hipStream_t st)
{
    hipLaunchKernelGGL(k_admm_vec, dim3(vec_blocks(a.n)), dim3(WG), 0, st, a, step);
    return hipGetLastError();
}

hipError_t launch_admm_dots(const AdmmVecs &a, int which, hipStream_t st)
{
    const unsigned nb = ADMM_DOT_BLOCKS;
    if (which == 0)
        hipLaunchKernelGGL(k_admm_dot, dim3(nb), dim3(WG), 0, st, a.r, a.r, a.h, a.p, a.n, a.part);
    else
        hipLaunchKernelGGL(k_admm_dot, dim3(nb), dim3(WG), 0, st, a.r, a.h, (const float *)nullptr,
                           (const float *)nullptr, a.n, a.part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_admm_coef, dim3(1), dim3(64), 0, st, a.part, (int)nb, which, a.S);
    return hipGetLastError();
}

}  // namespace trec
