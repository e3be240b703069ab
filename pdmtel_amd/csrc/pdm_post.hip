// pdm_post.hip -- post-processing on the GPU: the areal force density (PostAFD).
//
// CalcArealForceDensity (pdDataManager3.cpp:442-520) gives the stress tensor at a source
// node s from the bond forces that cross the three coordinate planes through it: for normal
// e_n, the collinear lattice nodes i = x_s - (1 + k) dx e_n (k < ns = int(delta / dx), found
// by exact coordinate equality), and for each the sum of the CCI bond forces of the nodes j
// of its family that lie in the half space (x_j - x_s) . e_n >= 0 (CalcForceOfPlusSpace,
// :522-620), times ds = delta / ns:  sigma[3 n + c] = sum_k ds sum_j f_ij,c.
// The bond force is CalcBondForceCCI's bond overload (pdDataManager2.cpp:1087-1154) on the
// node displacements: zero past ecrit.  One workgroup per source node; the collinear ids come
// from the host (exact-equality lookup).  Cut bonds (slits, an extension) carry no force.
#include "pdm_geom.h"
#include "pdm_internal.h"

namespace pdm {

namespace {

constexpr int kAfdThreads = 256;

__device__ __forceinline__ double afd_block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = kAfdThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// colinear[(t * 3 + n) * ns + k]: the collinear node k of direction n of source t (-1: none).
__global__ __launch_bounds__(kAfdThreads) void k_afd(AfdArgs a, const int64_t* __restrict__ src,
                                                     const int32_t* __restrict__ colinear, double* __restrict__ sigma)
{
    __shared__ double sh[kAfdThreads];
    const int64_t t = blockIdx.x;
    const int64_t s = src[t];
    const d4 xs = a.X[s];
    double sig[9];
    for (int q = 0; q < 9; ++q) sig[q] = 0.0;
    for (int n = 0; n < 3; ++n) {
        const double cs = (n == 0) ? xs.x : (n == 1) ? xs.y : xs.z;
        for (int k = 0; k < a.ns; ++k) {
            const int32_t i = colinear[(t * 3 + n) * a.ns + k];
            if (i < 0) continue;
            const d4 xi = a.X[i];
            const d4 ui = a.U[i];
            const int mi = a.mat[i];
            double f0 = 0.0, f1 = 0.0, f2 = 0.0;
            for (int64_t e = a.rp[i] + threadIdx.x; e < a.rp[i + 1]; e += kAfdThreads) {
                if (a.cut[e]) continue;
                const int32_t j = a.col[e];
                const d4 xj = a.X[j];
                const double cj = (n == 0) ? xj.x : (n == 1) ? xj.y : xj.z;
                if (!(cj - cs >= 0.0)) continue;  // j outside the R+ half space of s
                const d4 uj = a.U[j];
                const int mj = a.mat[j];
                const double xi1 = xj.x - xi.x, xi2 = xj.y - xi.y, xi3 = xj.z - xi.z;
                const double e1 = uj.x - ui.x, e2 = uj.y - ui.y, e3 = uj.z - ui.z;
                const double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
                const double a1 = xi1 + e1, a2 = xi2 + e2, a3 = xi3 + e3;
                const double p = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
                const double ecrit = fmin(a.ecrit[mi], a.ecrit[mj]);
                const double spring = fmin(a.spring[mi], a.spring[mj]);
                const double bij = 0.5 * (a.be[i] + a.be[j]);
                const double u = p - r;
                const double st = u / r;
                if (u <= 0.0 || (u > 0.0 && st <= ecrit)) {
                    const double beta = g_beta(r, a.horizon, a.dx, a.original_cci != 0);
                    const double w = bij * a.vol * beta * spring * st;
                    f0 += w * a1 / p;
                    f1 += w * a2 / p;
                    f2 += w * a3 / p;
                }
            }
            f0 = afd_block_sum(f0, sh);
            f1 = afd_block_sum(f1, sh);
            f2 = afd_block_sum(f2, sh);
            sig[3 * n] += f0 * a.ds;
            sig[3 * n + 1] += f1 * a.ds;
            sig[3 * n + 2] += f2 * a.ds;
        }
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 9; ++q) sigma[9 * t + q] = sig[q];
}

}  // namespace

hipError_t launch_afd(const AfdArgs& a, const int64_t* src, int64_t nsrc, const int32_t* colinear, double* sigma,
                      hipStream_t s)
{
    if (nsrc == 0) return hipSuccess;
    k_afd<<<dim3((unsigned)nsrc), dim3(kAfdThreads), 0, s>>>(a, src, colinear, sigma);
    return hipGetLastError();
}

}  // namespace pdm
