// pdm_kernels.hip -- bond-force gather fused with the RK4 stage update.
//
// One 64-lane wavefront per node family (4 families per 256-thread workgroup).
// Lanes stride over the family's CSR entries (col int32 + coef), gathering the
// neighbour's 64-byte NodeRec (position and this stage's displacement) through
// L2; the 3-component sum is reduced across the wavefront with xor shuffles and
// lane 0 applies the stage update: velocity buffer, boundary conditions, k, and
// the next stage's displacement into the other NodeRec buffer.  One launch is
// one stage of RK4 (pdDataManager.cpp:1994-2134), so a step is four launches
// and the state never leaves HBM.
//
// Per-bond arithmetic:
//   PDM_FP64:  CalcBondForceCCI's FP64 expression (pdDataManager2.cpp:1293-1351):
//              r = |xi|, p = |xi+eta|, s = (p - r)/r, f = ((coef*s)*(xi+eta))/p,
//              with coef = bdryEff_ij * dx^3 * beta * c_ij precomputed in the same
//              association order, so each bond force equals the reference's when
//              pow(x, 2.0) is the correctly rounded square.
//   PDM_FP32C: xi and eta formed in FP64 and rounded to FP32; the stretch from the
//              cancellation-free form s = (2 xi.eta + eta.eta) / ((p + r) r); FP32
//              coefficients; family sums kept as two-float (hi, lo) pairs, also across
//              lanes; the stage update in FP64.
#include <cmath>

#include "pdm_internal.h"

namespace pdm {
namespace {

constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;

__device__ __forceinline__ int64_t xcd_block(int remap)
{
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of families so its L2 holds one region's neighbours.
    int64_t b = blockIdx.x;
    if (!remap) return b;
    int64_t nb = gridDim.x;  // a multiple of 8 when remapping
    return (b & 7) * (nb >> 3) + (b >> 3);
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

struct Acc64 {
    double f[3];
    double el;
};

// FP64 family gather: reference arithmetic.
template <bool ENERGY>
__device__ __forceinline__ void gather_fp64(const NodeRec* __restrict__ rin, const int32_t* __restrict__ col,
                                            const double* __restrict__ coef, int64_t e0, int64_t e1, int lane,
                                            const NodeRec& ri, double inv_vcube, Acc64& acc)
{
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, el = 0.0;
    const double ecrit_i = ri.x.w;
    for (int64_t e = e0 + lane; e < e1; e += 64) {
        int32_t j = __builtin_nontemporal_load(&col[e]);
        double c = __builtin_nontemporal_load(&coef[e]);
        const NodeRec rj = rin[j];
        double xi1 = rj.x.x - ri.x.x, xi2 = rj.x.y - ri.x.y, xi3 = rj.x.z - ri.x.z;
        double et1 = rj.u.x - ri.u.x, et2 = rj.u.y - ri.u.y, et3 = rj.u.z - ri.u.z;
        double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
        double a1 = xi1 + et1, a2 = xi2 + et2, a3 = xi3 + et3;
        double p = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
        double u = p - r;
        double s = u / r;
        double ecrit = (rj.x.w < ecrit_i) ? rj.x.w : ecrit_i;
        if (u <= 0.0 || (u > 0.0 && s <= ecrit)) {
            double cs = c * s;
            f0 += (cs * a1) / p;
            f1 += (cs * a2) / p;
            f2 += (cs * a3) / p;
            if (ENERGY) el += 0.25 * (c * inv_vcube) * s * u;
        }
    }
    acc.f[0] = wave_sum(f0);
    acc.f[1] = wave_sum(f1);
    acc.f[2] = wave_sum(f2);
    acc.el = ENERGY ? wave_sum(el) : 0.0;
}

// FP32-compensated family gather.
template <bool ENERGY>
__device__ __forceinline__ void gather_fp32c(const NodeRec* __restrict__ rin, const int32_t* __restrict__ col,
                                             const float* __restrict__ coef, int64_t e0, int64_t e1, int lane,
                                             const NodeRec& ri, double inv_vcube, Acc64& acc)
{
    float h0 = 0.f, h1 = 0.f, h2 = 0.f, l0 = 0.f, l1 = 0.f, l2 = 0.f, el = 0.f;
    const float ecrit_i = (float)ri.x.w;
    const float ivc = (float)inv_vcube;
    for (int64_t e = e0 + lane; e < e1; e += 64) {
        int32_t j = __builtin_nontemporal_load(&col[e]);
        float c = __builtin_nontemporal_load(&coef[e]);
        const NodeRec rj = rin[j];
        float x1 = (float)(rj.x.x - ri.x.x), x2 = (float)(rj.x.y - ri.x.y), x3 = (float)(rj.x.z - ri.x.z);
        float e1_ = (float)(rj.u.x - ri.u.x), e2_ = (float)(rj.u.y - ri.u.y), e3_ = (float)(rj.u.z - ri.u.z);
        float r2 = x1 * x1 + x2 * x2 + x3 * x3;
        float q = 2.0f * (x1 * e1_ + x2 * e2_ + x3 * e3_) + (e1_ * e1_ + e2_ * e2_ + e3_ * e3_);
        float rinv = __builtin_amdgcn_rsqf(r2);
        float r = r2 * rinv;
        float p = __builtin_amdgcn_sqrtf(r2 + q);
        float u = q * __builtin_amdgcn_rcpf(p + r);
        float s = u * rinv;
        float ecrit = ((float)rj.x.w < ecrit_i) ? (float)rj.x.w : ecrit_i;
        if (u <= 0.0f || s <= ecrit) {
            float g = c * s * __builtin_amdgcn_rcpf(p);
            float t, err;
            two_sum(h0, g * (x1 + e1_), t, err); h0 = t; l0 += err;
            two_sum(h1, g * (x2 + e2_), t, err); h1 = t; l1 += err;
            two_sum(h2, g * (x3 + e3_), t, err); h2 = t; l2 += err;
            if (ENERGY) el += 0.25f * (c * ivc) * s * u;
        }
    }
    // cross-lane two-float reduction
    for (int off = 32; off > 0; off >>= 1) {
        float o0 = __shfl_xor(h0, off, 64), o1 = __shfl_xor(h1, off, 64), o2 = __shfl_xor(h2, off, 64);
        float m0 = __shfl_xor(l0, off, 64), m1 = __shfl_xor(l1, off, 64), m2 = __shfl_xor(l2, off, 64);
        float t, err;
        two_sum(h0, o0, t, err); h0 = t; l0 = l0 + m0 + err;
        two_sum(h1, o1, t, err); h1 = t; l1 = l1 + m1 + err;
        two_sum(h2, o2, t, err); h2 = t; l2 = l2 + m2 + err;
        if (ENERGY) el += __shfl_xor(el, off, 64);
    }
    acc.f[0] = (double)h0 + (double)l0;
    acc.f[1] = (double)h1 + (double)l1;
    acc.f[2] = (double)h2 + (double)l2;
    acc.el = (double)el;
}

__device__ __forceinline__ double time_factor_dev(const BCDev& b, double t)
{
    if (b.type == 2) return (t <= b.t_end) ? 1.0 : 0.0;
    if (b.load == 1.0) return (t <= b.t_end) ? t / b.t_end : 1.0;
    return 1.0;
}

// One RK4 stage (STAGE 0..3).  Reads a.rin, writes a.rout.
template <int STAGE, int PREC, bool ENERGY>
__global__ __launch_bounds__(kBlock) void k_stage(StageArgs a, int step_idx)
{
    __shared__ double sh_e[2][kWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = xcd_block(a.xcd_remap) * kWaves + wave;
    const bool active = i < a.n;
    const StepParam sp = a.steps[step_idx];
    const double dt = sp.dt;
    const double dts = (STAGE < 2) ? 0.5 * dt : dt;
    double ke_i = 0.0, ee_i = 0.0;
    if (active) {
        const NodeRec ri = a.rin[i];
        const int64_t e0 = a.row_ptr[i], e1 = a.row_ptr[i + 1];
        Acc64 acc;
        if (PREC == PDM_FP64) gather_fp64<ENERGY>(a.rin, a.col, a.coef64, e0, e1, lane, ri, a.inv_vol_cube, acc);
        else gather_fp32c<ENERGY>(a.rin, a.col, a.coef32, e0, e1, lane, ri, a.inv_vol_cube, acc);
        if (lane == 0) {
            const int m = a.mat[i];
            const int b = a.bc[i];
            const double rho = a.mats[m].rho;
            const double nbi = (double)a.nb[i];
            double Ft[3];
            for (int c = 0; c < 3; ++c) Ft[c] = acc.f[c] + nbi * (a.mats[m].g[c] * rho);
            bool bcnode = false;
            double fac = 1.0;
            if (b >= 0) {
                const BCDev& bd = a.bcs[b];
                bcnode = bd.bcnode != 0;
                fac = time_factor_dev(bd, sp.t);
                if (bd.type == 4)
                    for (int c = 0; c < 3; ++c) Ft[c] += bd.trac[c] * fac;
            }
            d4 vb = (STAGE == 0) ? a.V[i] : a.Vb[i];
            d4 U = a.U[i];
            double* vbp = &vb.x;
            double* Up = &U.x;
            for (int c = 0; c < 3; ++c) vbp[c] += Ft[c] * dts / rho;
            // UpdateNodeStatusWithBC (pdDataManager.cpp:2497-2602)
            if (b >= 0) {
                const BCDev& bd = a.bcs[b];
                const double X[3] = {ri.x.x, ri.x.y, ri.x.z};
                for (int c = 0; c < 3; ++c) {
                    if (bd.f[c] != 1) continue;
                    if (bd.type == 1) { Up[c] = bd.v[c] * fac; vbp[c] = 0.0; }
                    else if (bd.type == 2) { vbp[c] = bd.v[c] * fac; }
                    else if (bd.type == 3) {
                        Up[c] = (bd.v[3 * c] * X[0] + bd.v[3 * c + 1] * X[1] + bd.v[3 * c + 2] * X[2]) * fac;
                        vbp[c] = 0.0;
                    }
                }
            }
            double k[3];
            for (int c = 0; c < 3; ++c) k[c] = vbp[c] * dts;
            d4 nu;  // displacement the next stage gathers
            double* nup = &nu.x;
            nu.w = 0.0;
            if (STAGE == 0) {
                a.k1[i] = d4{k[0], k[1], k[2], 0.0};
                for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + 0.5 * k[c];
                a.Vb[i] = vb;
                if (ENERGY) a.elast[i] = acc.el;
            } else if (STAGE == 1) {
                a.k23[i] = d4{k[0], k[1], k[2], 0.0};
                for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + 0.5 * k[c];
                a.Vb[i] = vb;
                if (ENERGY) a.elast[i] += acc.el;
            } else if (STAGE == 2) {
                d4 k2 = a.k23[i];
                a.k23[i] = d4{k[0] + k2.x, k[1] + k2.y, k[2] + k2.z, 0.0};  // k3 += k24 (pdDataManager.cpp:2052)
                for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + k[c];
                a.Vb[i] = vb;
                if (ENERGY) a.elast[i] += acc.el;
            } else {
                const double dt6 = 1.0 / 6.0;
                d4 k1 = a.k1[i], k23 = a.k23[i];
                d4 V = a.V[i];
                const double* k1p = &k1.x;
                const double* k23p = &k23.x;
                const double* Vp = &V.x;
                for (int c = 0; c < 3; ++c) {
                    if (bcnode) Up[c] = Up[c] + dt6 * (k1p[c] + k[c] + 2.0 * k23p[c]);
                    else Up[c] = Up[c] + dt6 * (k1p[c] + k[c] + 2.0 * k23p[c]) - dt * a.damping * Vp[c] * dt / rho;
                }
                for (int c = 0; c < 3; ++c) vbp[c] += -a.damping * Vp[c] * dt / rho;
                vb.w = 0.0;
                a.V[i] = vb;
                for (int c = 0; c < 3; ++c) nup[c] = Up[c];
                if (ENERGY) {
                    double el = a.elast[i] + acc.el;
                    double velsq = vbp[0] * vbp[0] + vbp[1] * vbp[1] + vbp[2] * vbp[2];
                    ee_i = a.vol * el;
                    ke_i = 0.5 * rho * velsq * a.vol;
                }
            }
            if (b >= 0 || STAGE == 3) {
                U.w = 0.0;
                a.U[i] = U;
            }
            a.rout[i].u = nu;
        }
    }
    if (STAGE == 3 && ENERGY) {
        if (lane == 0) {
            sh_e[0][wave] = ke_i;
            sh_e[1][wave] = ee_i;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double ks = 0.0, es = 0.0;
            for (int w = 0; w < kWaves; ++w) { ks += sh_e[0][w]; es += sh_e[1][w]; }
            a.part[2 * blockIdx.x] = ks;
            a.part[2 * blockIdx.x + 1] = es;
        }
    }
}

__global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ part, int64_t nblocks,
                                                      double* __restrict__ out2)
{
    __shared__ double sk[256], se[256];
    double ks = 0.0, es = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += 256) {
        ks += part[2 * b];
        es += part[2 * b + 1];
    }
    sk[threadIdx.x] = ks;
    se[threadIdx.x] = es;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
            sk[threadIdx.x] += sk[threadIdx.x + off];
            se[threadIdx.x] += se[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[0] = sk[0];
        out2[1] = se[0];
    }
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_forces(ForceArgs a)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * kWaves + wave;
    if (i >= a.n) return;
    const NodeRec ri = a.rin[i];
    const int64_t e0 = a.row_ptr[i], e1 = a.row_ptr[i + 1];
    Acc64 acc;
    if (PREC == PDM_FP64) gather_fp64<true>(a.rin, a.col, a.coef64, e0, e1, lane, ri, a.inv_vol_cube, acc);
    else gather_fp32c<true>(a.rin, a.col, a.coef32, e0, e1, lane, ri, a.inv_vol_cube, acc);
    if (lane == 0) {
        a.F[3 * i] = acc.f[0];
        a.F[3 * i + 1] = acc.f[1];
        a.F[3 * i + 2] = acc.f[2];
        if (a.elast) a.elast[i] = acc.el;
    }
}

// Tangent of the bond force at eta = 0: y_i = sum_j coef_ij xi (xi . (u_i - u_j)) / r^3.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_tangent(ForceArgs a, const double* __restrict__ u3, double* __restrict__ y3)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * kWaves + wave;
    if (i >= a.n) return;
    const d4 xi = a.rin[i].x;
    const double ui0 = u3[3 * i], ui1 = u3[3 * i + 1], ui2 = u3[3 * i + 2];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t e = a.row_ptr[i] + lane; e < a.row_ptr[i + 1]; e += 64) {
        int32_t j = a.col[e];
        const d4 xj = a.rin[j].x;
        if (PREC == PDM_FP64) {
            double c = a.coef64[e];
            double x1 = xj.x - xi.x, x2 = xj.y - xi.y, x3 = xj.z - xi.z;
            double d0 = ui0 - u3[3 * (int64_t)j], d1 = ui1 - u3[3 * (int64_t)j + 1], d2 = ui2 - u3[3 * (int64_t)j + 2];
            double r = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
            double proj = (x1 * d0 + x2 * d1 + x3 * d2) * c / (r * r * r);
            s0 += proj * x1;
            s1 += proj * x2;
            s2 += proj * x3;
        } else {
            float c = a.coef32[e];
            float x1 = (float)(xj.x - xi.x), x2 = (float)(xj.y - xi.y), x3 = (float)(xj.z - xi.z);
            float d0 = (float)(ui0 - u3[3 * (int64_t)j]), d1 = (float)(ui1 - u3[3 * (int64_t)j + 1]),
                  d2 = (float)(ui2 - u3[3 * (int64_t)j + 2]);
            float r2 = x1 * x1 + x2 * x2 + x3 * x3;
            float ri = __builtin_amdgcn_rsqf(r2);
            float proj = (x1 * d0 + x2 * d1 + x3 * d2) * c * (ri * ri * ri);
            s0 += (double)(proj * x1);
            s1 += (double)(proj * x2);
            s2 += (double)(proj * x3);
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        y3[3 * i] = s0;
        y3[3 * i + 1] = s1;
        y3[3 * i + 2] = s2;
    }
}

template <int STAGE, int PREC>
hipError_t launch_stage_t(const StageArgs& a, int step_idx, hipStream_t s)
{
    int64_t nb = stage_blocks(a.n);
    if (a.energies)
        k_stage<STAGE, PREC, true><<<dim3((unsigned)nb), dim3(kBlock), 0, s>>>(a, step_idx);
    else
        k_stage<STAGE, PREC, false><<<dim3((unsigned)nb), dim3(kBlock), 0, s>>>(a, step_idx);
    return hipGetLastError();
}

template <int PREC>
hipError_t launch_stage_p(int stage, const StageArgs& a, int step_idx, hipStream_t s)
{
    switch (stage) {
    case 0: return launch_stage_t<0, PREC>(a, step_idx, s);
    case 1: return launch_stage_t<1, PREC>(a, step_idx, s);
    case 2: return launch_stage_t<2, PREC>(a, step_idx, s);
    default: return launch_stage_t<3, PREC>(a, step_idx, s);
    }
}

}  // namespace

int64_t stage_blocks(int64_t n)
{
    int64_t nb = (n + kWaves - 1) / kWaves;
    return (nb + 7) / 8 * 8;  // multiple of 8 for the XCD remap
}

hipError_t launch_stage(int stage, int precision, const StageArgs& a, int step_idx, hipStream_t s)
{
    if (precision == PDM_FP64) return launch_stage_p<PDM_FP64>(stage, a, step_idx, s);
    return launch_stage_p<PDM_FP32C>(stage, a, step_idx, s);
}

hipError_t launch_energy_final(const StageArgs& a, int64_t nblocks, double* out2, hipStream_t s)
{
    k_energy_final<<<dim3(1), dim3(256), 0, s>>>(a.part, nblocks, out2);
    return hipGetLastError();
}

hipError_t launch_forces(int precision, const ForceArgs& a, hipStream_t s)
{
    unsigned nb = (unsigned)((a.n + kWaves - 1) / kWaves);
    if (nb == 0) return hipSuccess;
    if (precision == PDM_FP64) k_forces<PDM_FP64><<<dim3(nb), dim3(kBlock), 0, s>>>(a);
    else k_forces<PDM_FP32C><<<dim3(nb), dim3(kBlock), 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_tangent(int precision, const ForceArgs& a, const double* u3, double* y3, hipStream_t s)
{
    unsigned nb = (unsigned)((a.n + kWaves - 1) / kWaves);
    if (nb == 0) return hipSuccess;
    if (precision == PDM_FP64) k_tangent<PDM_FP64><<<dim3(nb), dim3(kBlock), 0, s>>>(a, u3, y3);
    else k_tangent<PDM_FP32C><<<dim3(nb), dim3(kBlock), 0, s>>>(a, u3, y3);
    return hipGetLastError();
}

}  // namespace pdm
