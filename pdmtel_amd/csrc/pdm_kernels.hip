// pdm_kernels.hip -- LDS-staged bond-force gathers fused with the RK4 stage update.
//
// One workgroup per tile.  Phase 1 stages the tile's halo (positions and this
// stage's displacements, FP64, SoA) and its families' row starts and own slots
// into LDS with coalesced loads.  Phase 2: each 64-lane wavefront takes one
// family at a time; every lane keeps kUnroll entries (a uint16 LDS slot and a
// coefficient each, streamed from HBM) in flight, and the first chunk of the
// wavefront's next family is loaded while the current one is computed.  The
// neighbour's data comes from LDS; the family sum is reduced across the
// wavefront with xor shuffles into an LDS result slot.  Phase 3: one thread per
// tile node applies the stage update (velocity buffer, boundary conditions, k,
// next stage's displacement) on contiguous state arrays.  One launch is one
// stage of RK4 (pdDataManager.cpp:1994-2134); a step is four.
//
// Per-bond arithmetic:
//   PDM_FP64:  CalcBondForceCCI's FP64 expression (pdDataManager2.cpp:1293-1351):
//              r = |xi|, p = |xi+eta|, s = (p - r)/r, f = ((coef*s)*(xi+eta))/p, with
//              coef = bdryEff_ij * dx^3 * beta * c_ij precomputed in the same association
//              order, so each bond force equals the reference's when pow(x, 2.0) is the
//              correctly rounded square.
//   PDM_FP32C: xi and eta formed in FP64 and rounded to FP32; the stretch from the
//              cancellation-free form s = (2 xi.eta + eta.eta) / ((p + r) r); FP32
//              coefficients; family sums kept as two-float (hi, lo) pairs, also across
//              lanes; the stage update in FP64.
#include <cmath>

#include "pdm_internal.h"

namespace pdm {
namespace {

constexpr int kUnroll = 4;              // entries per lane in flight
constexpr int kChunk = 64 * kUnroll;    // entries per wavefront chunk

__device__ __forceinline__ int64_t launch_index(int remap)
{
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous run of tiles so its L2 serves the neighbouring tiles' halos.
    int64_t b = blockIdx.x;
    if (!remap) return b;
    int64_t nb = gridDim.x;  // a multiple of 8
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

// LDS layout of a tile: SoA halo arrays of H doubles, then per-node results
// (4 doubles), row starts (nt + 1 int64) and own slots (nt ints).
struct Lds {
    double *x, *y, *z, *ux, *uy, *uz, *ec;
    double* res;
    int64_t* row;
    int* self;
};

__device__ __forceinline__ Lds lds_view(double* base, int H, int tile_max, bool with_ecrit)
{
    Lds l;
    l.x = base;
    l.y = base + H;
    l.z = base + 2 * H;
    l.ux = base + 3 * H;
    l.uy = base + 4 * H;
    l.uz = base + 5 * H;
    l.ec = with_ecrit ? base + 6 * H : nullptr;
    l.res = base + (with_ecrit ? 7 : 6) * H;
    l.row = reinterpret_cast<int64_t*>(l.res + 4 * tile_max);
    l.self = reinterpret_cast<int*>(l.row + tile_max + 1);
    return l;
}

__device__ __forceinline__ void stage_tile(const TileView& tv, const d4* __restrict__ u, int64_t h0, int H,
                                           int64_t a0, int nt, const Lds& l, bool with_ecrit)
{
    for (int q = threadIdx.x; q < H; q += blockDim.x) {
        int32_t id = tv.halo[h0 + q];
        d4 x = tv.X[id];
        d4 v = u[id];
        l.x[q] = x.x;
        l.y[q] = x.y;
        l.z[q] = x.z;
        l.ux[q] = v.x;
        l.uy[q] = v.y;
        l.uz[q] = v.z;
        if (with_ecrit) l.ec[q] = x.w;
    }
    for (int q = threadIdx.x; q <= nt; q += blockDim.x) l.row[q] = tv.rpi[a0 + q];
    for (int q = threadIdx.x; q < nt; q += blockDim.x) l.self[q] = tv.self_slot[a0 + q];
}

template <int PREC>
struct Chunk;
template <>
struct Chunk<PDM_FP64> {
    int s[kUnroll];
    double c[kUnroll];
};
template <>
struct Chunk<PDM_FP32C> {
    int s[kUnroll];
    float c[kUnroll];
};

template <int PREC>
__device__ __forceinline__ void load_chunk(const TileView& tv, int64_t base, int64_t e1, int lane, Chunk<PREC>& ch)
{
#pragma unroll
    for (int k = 0; k < kUnroll; ++k) {
        const int64_t e = base + lane + 64 * k;
        const bool v = e < e1;
        ch.s[k] = v ? (int)__builtin_nontemporal_load(&tv.slot[e]) : -1;
        if constexpr (PREC == PDM_FP64) ch.c[k] = v ? __builtin_nontemporal_load(&tv.coef64[e]) : 0.0;
        else ch.c[k] = v ? __builtin_nontemporal_load(&tv.coef32[e]) : 0.0f;
    }
}

struct Self {
    double x, y, z, u, v, w, ecrit;
};

// Per-lane accumulators: FP64 sums, or two-float (hi, lo) pairs.
template <int PREC>
struct Acc;
template <>
struct Acc<PDM_FP64> {
    double f0, f1, f2, el;
    __device__ void zero() { f0 = f1 = f2 = el = 0.0; }
};
template <>
struct Acc<PDM_FP32C> {
    float h0, h1, h2, l0, l1, l2, el;
    __device__ void zero() { h0 = h1 = h2 = l0 = l1 = l2 = el = 0.0f; }
};

// One bond, reference FP64 arithmetic.
template <bool ENERGY, bool UE>
__device__ __forceinline__ void bond(const Lds& l, int s, double c, const Self& me, double ivc, Acc<PDM_FP64>& a)
{
    double xi1 = l.x[s] - me.x, xi2 = l.y[s] - me.y, xi3 = l.z[s] - me.z;
    double et1 = l.ux[s] - me.u, et2 = l.uy[s] - me.v, et3 = l.uz[s] - me.w;
    double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
    double a1 = xi1 + et1, a2 = xi2 + et2, a3 = xi3 + et3;
    double p = sqrt(a1 * a1 + a2 * a2 + a3 * a3);
    double u = p - r;
    double st = u / r;
    double ecrit = me.ecrit;
    if (!UE) {
        double ej = l.ec[s];
        ecrit = (ej < me.ecrit) ? ej : me.ecrit;
    }
    if (u <= 0.0 || (u > 0.0 && st <= ecrit)) {
        double cs = c * st;
        a.f0 += (cs * a1) / p;
        a.f1 += (cs * a2) / p;
        a.f2 += (cs * a3) / p;
        if (ENERGY) a.el += 0.25 * (c * ivc) * st * u;
    }
}

// One bond, FP32 with the cancellation-free stretch and two-float sums.
template <bool ENERGY, bool UE>
__device__ __forceinline__ void bond(const Lds& l, int s, float c, const Self& me, double ivc, Acc<PDM_FP32C>& a)
{
    float x1 = (float)(l.x[s] - me.x), x2 = (float)(l.y[s] - me.y), x3 = (float)(l.z[s] - me.z);
    float e1 = (float)(l.ux[s] - me.u), e2 = (float)(l.uy[s] - me.v), e3 = (float)(l.uz[s] - me.w);
    float r2 = x1 * x1 + x2 * x2 + x3 * x3;
    float q = 2.0f * (x1 * e1 + x2 * e2 + x3 * e3) + (e1 * e1 + e2 * e2 + e3 * e3);
    float rinv = __builtin_amdgcn_rsqf(r2);
    float r = r2 * rinv;
    float p = __builtin_amdgcn_sqrtf(r2 + q);
    float u = q * __builtin_amdgcn_rcpf(p + r);
    float st = u * rinv;
    float ecrit = (float)me.ecrit;
    if (!UE) {
        float ej = (float)l.ec[s];
        ecrit = (ej < ecrit) ? ej : ecrit;
    }
    if (u <= 0.0f || st <= ecrit) {
        float g = c * st * __builtin_amdgcn_rcpf(p);
        float t, err;
        two_sum(a.h0, g * (x1 + e1), t, err); a.h0 = t; a.l0 += err;
        two_sum(a.h1, g * (x2 + e2), t, err); a.h1 = t; a.l1 += err;
        two_sum(a.h2, g * (x3 + e3), t, err); a.h2 = t; a.l2 += err;
        if (ENERGY) a.el += 0.25f * (c * (float)ivc) * st * u;
    }
}

__device__ __forceinline__ void reduce_store(Acc<PDM_FP64>& a, int lane, double* out)
{
    double f0 = wave_sum(a.f0), f1 = wave_sum(a.f1), f2 = wave_sum(a.f2), el = wave_sum(a.el);
    if (lane == 0) {
        out[0] = f0;
        out[1] = f1;
        out[2] = f2;
        out[3] = el;
    }
}

__device__ __forceinline__ void reduce_store(Acc<PDM_FP32C>& a, int lane, double* out)
{
    for (int off = 32; off > 0; off >>= 1) {
        float o0 = __shfl_xor(a.h0, off, 64), o1 = __shfl_xor(a.h1, off, 64), o2 = __shfl_xor(a.h2, off, 64);
        float m0 = __shfl_xor(a.l0, off, 64), m1 = __shfl_xor(a.l1, off, 64), m2 = __shfl_xor(a.l2, off, 64);
        float t, err;
        two_sum(a.h0, o0, t, err); a.h0 = t; a.l0 = a.l0 + m0 + err;
        two_sum(a.h1, o1, t, err); a.h1 = t; a.l1 = a.l1 + m1 + err;
        two_sum(a.h2, o2, t, err); a.h2 = t; a.l2 = a.l2 + m2 + err;
        a.el += __shfl_xor(a.el, off, 64);
    }
    if (lane == 0) {
        out[0] = (double)a.h0 + (double)a.l0;
        out[1] = (double)a.h1 + (double)a.l1;
        out[2] = (double)a.h2 + (double)a.l2;
        out[3] = (double)a.el;
    }
}

// Family sums of the tile's nodes into l.res; wavefront w takes nodes w, w + NW, ...
template <int PREC, bool ENERGY, bool UE, int NW>
__device__ __forceinline__ void tile_families(const TileView& tv, const Lds& l, int nt)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int loc = wave;
    if (loc >= nt) return;
    int64_t e0 = l.row[loc], e1 = l.row[loc + 1];
    Chunk<PREC> cur;
    load_chunk<PREC>(tv, e0, e1, lane, cur);
    for (;;) {
        const int nloc = loc + NW;
        int64_t n0 = 0, n1 = 0;
        Chunk<PREC> nxt;
        if (nloc < nt) {
            n0 = l.row[nloc];
            n1 = l.row[nloc + 1];
            load_chunk<PREC>(tv, n0, n1, lane, nxt);
        }
        const int ss = l.self[loc];
        Self me{l.x[ss], l.y[ss], l.z[ss], l.ux[ss], l.uy[ss], l.uz[ss], UE ? tv.ecrit0 : l.ec[ss]};
        Acc<PREC> acc;
        acc.zero();
#pragma unroll
        for (int k = 0; k < kUnroll; ++k)
            if (cur.s[k] >= 0) bond<ENERGY, UE>(l, cur.s[k], cur.c[k], me, tv.inv_vol_cube, acc);
        for (int64_t base = e0 + kChunk; base < e1; base += kChunk) {
            Chunk<PREC> more;
            load_chunk<PREC>(tv, base, e1, lane, more);
#pragma unroll
            for (int k = 0; k < kUnroll; ++k)
                if (more.s[k] >= 0) bond<ENERGY, UE>(l, more.s[k], more.c[k], me, tv.inv_vol_cube, acc);
        }
        reduce_store(acc, lane, l.res + 4 * loc);
        if (nloc >= nt) break;
        loc = nloc;
        e0 = n0;
        e1 = n1;
        cur = nxt;
    }
}

__device__ __forceinline__ double time_factor_dev(const BCDev& b, double t)
{
    if (b.type == 2) return (t <= b.t_end) ? 1.0 : 0.0;
    if (b.load == 1.0) return (t <= b.t_end) ? t / b.t_end : 1.0;
    return 1.0;
}

// One RK4 stage (STAGE 0..3) for one tile.  Reads a.uin, writes a.uout.
template <int STAGE, int PREC, bool ENERGY, bool UE, int NW>
__global__ __launch_bounds__(64 * NW) void k_stage(StageArgs a, int step_idx)
{
    extern __shared__ double L[];
    __shared__ double sh_e[2][NW];
    const TileView& tv = a.tv;
    const int64_t k = launch_index(tv.xcd_remap);
    if (k >= tv.ntl) return;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int H = (int)(tv.halo_ptr[k + 1] - h0);
    const Lds l = lds_view(L, H, tv.tile_max, !UE);
    stage_tile(tv, a.uin, h0, H, a0, nt, l, !UE);
    __syncthreads();
    tile_families<PREC, ENERGY, UE, NW>(tv, l, nt);
    __syncthreads();

    const StepParam sp = a.steps[step_idx];
    const double dt = sp.dt;
    const double dts = (STAGE < 2) ? 0.5 * dt : dt;
    double ke_i = 0.0, ee_i = 0.0;
    for (int loc = threadIdx.x; loc < nt; loc += 64 * NW) {
        const int64_t i = a0 + loc;
        const double* res = l.res + 4 * loc;
        const int m = a.mat[i];
        const int b = a.bc[i];
        const double rho = a.mats[m].rho;
        const double nbi = (double)a.nb[i];
        double Ft[3];
        for (int c = 0; c < 3; ++c) Ft[c] = res[c] + nbi * (a.mats[m].g[c] * rho);
        const double eli = res[3];
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
            const d4 X = tv.X[i];
            for (int c = 0; c < 3; ++c) {
                if (bd.f[c] != 1) continue;
                if (bd.type == 1) { Up[c] = bd.v[c] * fac; vbp[c] = 0.0; }
                else if (bd.type == 2) { vbp[c] = bd.v[c] * fac; }
                else if (bd.type == 3) {
                    Up[c] = (bd.v[3 * c] * X.x + bd.v[3 * c + 1] * X.y + bd.v[3 * c + 2] * X.z) * fac;
                    vbp[c] = 0.0;
                }
            }
        }
        double kk[3];
        for (int c = 0; c < 3; ++c) kk[c] = vbp[c] * dts;
        d4 nu;
        double* nup = &nu.x;
        nu.w = 0.0;
        if (STAGE == 0) {
            a.k1[i] = d4{kk[0], kk[1], kk[2], 0.0};
            for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + 0.5 * kk[c];
            a.Vb[i] = vb;
            if (ENERGY) a.elast[i] = eli;
        } else if (STAGE == 1) {
            a.k23[i] = d4{kk[0], kk[1], kk[2], 0.0};
            for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + 0.5 * kk[c];
            a.Vb[i] = vb;
            if (ENERGY) a.elast[i] += eli;
        } else if (STAGE == 2) {
            d4 k2 = a.k23[i];
            a.k23[i] = d4{kk[0] + k2.x, kk[1] + k2.y, kk[2] + k2.z, 0.0};  // k3 += k24 (pdDataManager.cpp:2052)
            for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + kk[c];
            a.Vb[i] = vb;
            if (ENERGY) a.elast[i] += eli;
        } else {
            const double dt6 = 1.0 / 6.0;
            const d4 k1 = a.k1[i], k23 = a.k23[i], V = a.V[i];
            const double* k1p = &k1.x;
            const double* k23p = &k23.x;
            const double* Vp = &V.x;
            for (int c = 0; c < 3; ++c) {
                if (bcnode) Up[c] = Up[c] + dt6 * (k1p[c] + kk[c] + 2.0 * k23p[c]);
                else Up[c] = Up[c] + dt6 * (k1p[c] + kk[c] + 2.0 * k23p[c]) - dt * a.damping * Vp[c] * dt / rho;
            }
            for (int c = 0; c < 3; ++c) vbp[c] += -a.damping * Vp[c] * dt / rho;
            vb.w = 0.0;
            a.V[i] = vb;
            for (int c = 0; c < 3; ++c) nup[c] = Up[c];
            if (ENERGY) {
                const double elt = a.elast[i] + eli;
                const double velsq = vbp[0] * vbp[0] + vbp[1] * vbp[1] + vbp[2] * vbp[2];
                ee_i += a.vol * elt;
                ke_i += 0.5 * rho * velsq * a.vol;
            }
        }
        if (b >= 0 || STAGE == 3) {
            U.w = 0.0;
            a.U[i] = U;
        }
        a.uout[i] = nu;
    }
    if (STAGE == 3 && ENERGY) {
        const int wave = threadIdx.x >> 6;
        ke_i = wave_sum(ke_i);
        ee_i = wave_sum(ee_i);
        if ((threadIdx.x & 63) == 0) {
            sh_e[0][wave] = ke_i;
            sh_e[1][wave] = ee_i;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double ks = 0.0, es = 0.0;
            for (int w = 0; w < NW; ++w) { ks += sh_e[0][w]; es += sh_e[1][w]; }
            a.part[2 * k] = ks;
            a.part[2 * k + 1] = es;
        }
    }
}

// Sum pairs (ke, ee): block b reduces a contiguous chunk of `in` into out[2b], out[2b+1].
__global__ __launch_bounds__(256) void k_reduce_pairs(const double* __restrict__ in, int64_t n, int64_t chunk,
                                                      double* __restrict__ out)
{
    __shared__ double sk[256], se[256];
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t b1 = (b0 + chunk < n) ? b0 + chunk : n;
    double ks = 0.0, es = 0.0;
    for (int64_t q = b0 + threadIdx.x; q < b1; q += 256) {
        ks += in[2 * q];
        es += in[2 * q + 1];
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
        out[2 * blockIdx.x] = sk[0];
        out[2 * blockIdx.x + 1] = se[0];
    }
}

template <int PREC, bool UE, int NW>
__global__ __launch_bounds__(64 * NW) void k_forces(TileView tv, const d4* __restrict__ u, d4* __restrict__ F,
                                                    double* __restrict__ elast)
{
    extern __shared__ double L[];
    const int64_t k = blockIdx.x;
    if (k >= tv.ntl) return;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int H = (int)(tv.halo_ptr[k + 1] - h0);
    const Lds l = lds_view(L, H, tv.tile_max, !UE);
    stage_tile(tv, u, h0, H, a0, nt, l, !UE);
    __syncthreads();
    tile_families<PREC, true, UE, NW>(tv, l, nt);
    __syncthreads();
    for (int loc = threadIdx.x; loc < nt; loc += 64 * NW) {
        const double* res = l.res + 4 * loc;
        F[a0 + loc] = d4{res[0], res[1], res[2], 0.0};
        if (elast) elast[a0 + loc] = res[3];
    }
}

// Tangent of the bond force at eta = 0: y_i = sum_j coef_ij xi (xi . (u_i - u_j)) / r^3.
template <int PREC, int NW>
__global__ __launch_bounds__(64 * NW) void k_tangent(TileView tv, const d4* __restrict__ u, d4* __restrict__ y)
{
    extern __shared__ double L[];
    const int64_t k = launch_index(tv.xcd_remap);
    if (k >= tv.ntl) return;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int H = (int)(tv.halo_ptr[k + 1] - h0);
    const Lds l = lds_view(L, H, tv.tile_max, false);
    stage_tile(tv, u, h0, H, a0, nt, l, false);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int loc = wave; loc < nt; loc += NW) {
        const int ss = l.self[loc];
        const double xi = l.x[ss], yi = l.y[ss], zi = l.z[ss];
        const double ui = l.ux[ss], vi = l.uy[ss], wi = l.uz[ss];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        const int64_t e1 = l.row[loc + 1];
        for (int64_t base = l.row[loc]; base < e1; base += kChunk) {
            Chunk<PREC> ch;
            load_chunk<PREC>(tv, base, e1, lane, ch);
#pragma unroll
            for (int q = 0; q < kUnroll; ++q) {
                const int s = ch.s[q];
                if (s < 0) continue;
                if constexpr (PREC == PDM_FP64) {
                    const double c = ch.c[q];
                    double x1 = l.x[s] - xi, x2 = l.y[s] - yi, x3 = l.z[s] - zi;
                    double d0 = ui - l.ux[s], d1 = vi - l.uy[s], d2 = wi - l.uz[s];
                    double r2 = x1 * x1 + x2 * x2 + x3 * x3;
                    double proj = (x1 * d0 + x2 * d1 + x3 * d2) * c / (r2 * sqrt(r2));
                    s0 += proj * x1;
                    s1 += proj * x2;
                    s2 += proj * x3;
                } else {
                    const float c = ch.c[q];
                    float x1 = (float)(l.x[s] - xi), x2 = (float)(l.y[s] - yi), x3 = (float)(l.z[s] - zi);
                    float d0 = (float)(ui - l.ux[s]), d1 = (float)(vi - l.uy[s]), d2 = (float)(wi - l.uz[s]);
                    float r2 = x1 * x1 + x2 * x2 + x3 * x3;
                    float ri = __builtin_amdgcn_rsqf(r2);
                    float proj = (x1 * d0 + x2 * d1 + x3 * d2) * c * (ri * ri * ri);
                    s0 += (double)(proj * x1);
                    s1 += (double)(proj * x2);
                    s2 += (double)(proj * x3);
                }
            }
        }
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        s2 = wave_sum(s2);
        if (lane == 0) {
            l.res[4 * loc] = s0;
            l.res[4 * loc + 1] = s1;
            l.res[4 * loc + 2] = s2;
        }
    }
    __syncthreads();
    for (int loc = threadIdx.x; loc < nt; loc += 64 * NW)
        y[a0 + loc] = d4{l.res[4 * loc], l.res[4 * loc + 1], l.res[4 * loc + 2], 0.0};
}

int64_t padded_grid(int64_t ntl) { return (ntl + 7) / 8 * 8; }

// Raise a kernel's dynamic-LDS limit once per instantiation (gfx950 has 160 KB per CU).
template <auto Kernel>
hipError_t set_lds(size_t bytes)
{
    static size_t configured = 0;  // one per kernel instantiation
    if (bytes <= configured) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(Kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) configured = bytes;
    return e;
}

constexpr int kWavesDefault = 8;

template <int STAGE, int PREC, bool ENERGY, bool UE>
hipError_t launch_stage_t(const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    constexpr int NW = kWavesDefault;
    hipError_t e = set_lds<k_stage<STAGE, PREC, ENERGY, UE, NW>>(lds);
    if (e != hipSuccess) return e;
    k_stage<STAGE, PREC, ENERGY, UE, NW>
        <<<dim3((unsigned)padded_grid(a.tv.ntl)), dim3(64 * NW), lds, s>>>(a, step_idx);
    return hipGetLastError();
}

template <int STAGE, int PREC>
hipError_t launch_stage_e(const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    if (a.energies) {
        if (a.tv.uniform_ecrit) return launch_stage_t<STAGE, PREC, true, true>(a, step_idx, lds, s);
        return launch_stage_t<STAGE, PREC, true, false>(a, step_idx, lds, s);
    }
    if (a.tv.uniform_ecrit) return launch_stage_t<STAGE, PREC, false, true>(a, step_idx, lds, s);
    return launch_stage_t<STAGE, PREC, false, false>(a, step_idx, lds, s);
}

template <int PREC>
hipError_t launch_stage_p(int stage, const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    switch (stage) {
    case 0: return launch_stage_e<0, PREC>(a, step_idx, lds, s);
    case 1: return launch_stage_e<1, PREC>(a, step_idx, lds, s);
    case 2: return launch_stage_e<2, PREC>(a, step_idx, lds, s);
    default: return launch_stage_e<3, PREC>(a, step_idx, lds, s);
    }
}

template <int PREC, bool UE>
hipError_t launch_forces_t(const TileView& tv, const d4* u, d4* F, double* elast, size_t lds, hipStream_t s)
{
    constexpr int NW = kWavesDefault;
    hipError_t e = set_lds<k_forces<PREC, UE, NW>>(lds);
    if (e != hipSuccess) return e;
    k_forces<PREC, UE, NW><<<dim3((unsigned)tv.ntl), dim3(64 * NW), lds, s>>>(tv, u, F, elast);
    return hipGetLastError();
}

template <int PREC>
hipError_t launch_tangent_t(const TileView& tv, const d4* u, d4* y, size_t lds, hipStream_t s)
{
    constexpr int NW = kWavesDefault;
    hipError_t e = set_lds<k_tangent<PREC, NW>>(lds);
    if (e != hipSuccess) return e;
    k_tangent<PREC, NW><<<dim3((unsigned)padded_grid(tv.ntl)), dim3(64 * NW), lds, s>>>(tv, u, y);
    return hipGetLastError();
}

}  // namespace

size_t stage_lds_bytes(const TileView& tv, bool with_ecrit)
{
    return sizeof(double) * ((size_t)(with_ecrit ? 7 : 6) * tv.halo_max + 4 * (size_t)tv.tile_max) +
           sizeof(int64_t) * ((size_t)tv.tile_max + 1) + sizeof(int) * (size_t)tv.tile_max;
}

int64_t energy_part2_size(int64_t nparts) { return nparts / 1024 + 2; }

hipError_t launch_stage(int stage, int precision, const StageArgs& a, int step_idx, hipStream_t s)
{
    if (a.tv.ntl == 0) return hipSuccess;
    size_t lds = stage_lds_bytes(a.tv, !a.tv.uniform_ecrit);
    if (precision == PDM_FP64) return launch_stage_p<PDM_FP64>(stage, a, step_idx, lds, s);
    return launch_stage_p<PDM_FP32C>(stage, a, step_idx, lds, s);
}

hipError_t launch_energy_final(const double* part, int64_t nparts, double* part2, double* out2, hipStream_t s)
{
    const int64_t chunk = 1024;
    int64_t nb = (nparts + chunk - 1) / chunk;
    if (nb < 1) nb = 1;
    k_reduce_pairs<<<dim3((unsigned)nb), dim3(256), 0, s>>>(part, nparts, chunk, part2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_reduce_pairs<<<dim3(1), dim3(256), 0, s>>>(part2, nb, nb, out2);
    return hipGetLastError();
}

hipError_t launch_forces(int precision, const TileView& tv, const d4* u, d4* F, double* elast, hipStream_t s)
{
    if (tv.ntl == 0) return hipSuccess;
    size_t lds = stage_lds_bytes(tv, !tv.uniform_ecrit);
    if (precision == PDM_FP64) {
        if (tv.uniform_ecrit) return launch_forces_t<PDM_FP64, true>(tv, u, F, elast, lds, s);
        return launch_forces_t<PDM_FP64, false>(tv, u, F, elast, lds, s);
    }
    if (tv.uniform_ecrit) return launch_forces_t<PDM_FP32C, true>(tv, u, F, elast, lds, s);
    return launch_forces_t<PDM_FP32C, false>(tv, u, F, elast, lds, s);
}

hipError_t launch_tangent(int precision, const TileView& tv, const d4* u, d4* y, hipStream_t s)
{
    if (tv.ntl == 0) return hipSuccess;
    size_t lds = stage_lds_bytes(tv, false);
    if (precision == PDM_FP64) return launch_tangent_t<PDM_FP64>(tv, u, y, lds, s);
    return launch_tangent_t<PDM_FP32C>(tv, u, y, lds, s);
}

}  // namespace pdm
