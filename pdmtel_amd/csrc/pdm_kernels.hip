// pdm_kernels.hip -- LDS-staged bond-force gathers fused with the RK4 stage update.
//
// One workgroup per tile.  Phase 1 stages the tile's halo as 48-byte records
// (position and this stage's displacement; FP64, or fixed point and FP64 u S with the
// FP32 coefficient inputs) plus the families' row starts and own slots into LDS with
// coalesced loads; one extra record far outside the body, with zero weight, backs the
// row padding (CCI rows hold only entries with a nonzero coefficient).  Phase 2: each
// 64-lane wavefront works on NPW families at a time, 64/NPW lanes per family; every lane
// loads aligned blocks of entries (uint16 LDS slots, and FP64 coefficients or FP32
// correction bytes, streamed from HBM as vector loads): a first block of 12 (FP32
// operations: 8 + 4) or 4 (FP64) loaded while the previous round is computed, then
// blocks of 4 each loaded while the one before is computed.  The neighbour's record is
// three ds_read_b128.  Family sums
// are reduced inside the lane group with DPP moves (quad_perm, row_half_mirror,
// row_mirror) and the gfx950 permlane16/32 swaps, never through LDS.  Phase 3:
// one thread per tile node applies the stage update (velocity buffer, boundary
// conditions, k, next stage's displacement) on contiguous state arrays.  One
// launch is one stage of RK4 (pdDataManager.cpp:1994-2134); a step is four.
//
// Per-bond arithmetic:
//   PDM_FP64:   CalcBondForceCCI's FP64 expression (pdDataManager2.cpp:1293-1351):
//               r = |xi|, p = |xi+eta|, s = (p - r)/r, f = ((coef*s)*(xi+eta))/p, with
//               coef = bdryEff_ij * dx^3 * beta * c_ij precomputed in the same association
//               order, so each bond force equals the reference's when pow(x, 2.0) is the
//               correctly rounded square.
//   PDM_FP32C:  xi and eta formed exactly and rounded to FP32; the stretch from the
//               cancellation-free form s = (2 xi.eta + eta.eta) / ((p + r) r); the
//               coefficient formed from the two records and the entry's correction byte;
//               each lane's bond forces summed in FP32 and the lane sums of a family
//               reduced in FP32, then widened to FP64 for the stage update.
//   PDM_FP32C2: the same bond force with two-float (hi, lo) FP32 family sums: FP32 partials
//               of 4 bonds added to each lane's pair with TwoSum, the lanes' hi + lo summed
//               in FP64.  The stage update is FP64 in every mode.
#include <cmath>
#include <cstddef>
#include <type_traits>
#include <utility>

#include "pdm_geom.h"
#include "pdm_internal.h"

namespace pdm {
namespace {

constexpr double kFar = 1.0e3;          // coordinate of the padding record (m)

// Specimen of launch index k (one specimen when the map is absent).
__device__ __forceinline__ int tile_specimen(const TileView& tv, int64_t k) { return tv.tile_spec ? tv.tile_spec[k] : 0; }

__device__ __forceinline__ int64_t launch_index(int remap)
{
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous run of tiles so its L2 serves the neighbouring tiles' halos.
    int64_t b = blockIdx.x;
    if (!remap) return b;
    int64_t nb = gridDim.x;  // a multiple of 8
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// ------------------------------------------------------------ cross-lane sums
// DPP controls (CDNA): quad_perm [1,0,3,2] = xor 1, quad_perm [2,3,0,1] = xor 2,
// row_half_mirror pairs the two quads of each 8-lane half, row_mirror the two halves
// of each 16-lane row.  Applied in that order they leave the row sum in every lane.
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;

// (update_dpp with old 0 and bound_ctrl: every lane of these permutations is written, and the
// form lets the compiler fold the move into the add that consumes it, v_add_f32_dpp)
template <int CTRL>
__device__ __forceinline__ float dppf(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ double dppd(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// v from lane i ^ 16 (SWAP16) or i ^ 32, via the gfx950 permlane swaps.
template <bool SWAP16>
__device__ __forceinline__ int perm_xor(int v)
{
    auto p = SWAP16 ? __builtin_amdgcn_permlane16_swap(v, v, false, false)
                    : __builtin_amdgcn_permlane32_swap(v, v, false, false);
    // with both operands v: p[0] holds the even (lower) half's values, p[1] the odd (upper)'s
    const int lane = threadIdx.x & 63;
    const bool upper = SWAP16 ? ((lane >> 4) & 1) : (lane >> 5);
    return upper ? p[0] : p[1];
}

template <bool SWAP16>
__device__ __forceinline__ double perm_xord(double v)
{
    return __hiloint2double(perm_xor<SWAP16>(__double2hiint(v)), perm_xor<SWAP16>(__double2loint(v)));
}

template <bool SWAP16>
__device__ __forceinline__ float perm_xorf(float v)
{
    return __int_as_float(perm_xor<SWAP16>(__float_as_int(v)));
}

// v + v(lane ^ 16) (SWAP16) or v + v(lane ^ 32): the swap of v with itself leaves one
// lane's own value in one result and its partner's in the other, so their sum is the
// pair sum in every lane, with no select.
template <bool SWAP16>
__device__ __forceinline__ double pair_sumd(double v)
{
    const int hi = __double2hiint(v), lo = __double2loint(v);
    auto ph = SWAP16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                     : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    auto pl = SWAP16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                     : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    return __hiloint2double(ph[0], pl[0]) + __hiloint2double(ph[1], pl[1]);
}

template <bool SWAP16>
__device__ __forceinline__ float pair_sumf(float v)
{
    const int b = __float_as_int(v);
    auto p = SWAP16 ? __builtin_amdgcn_permlane16_swap(b, b, false, false)
                    : __builtin_amdgcn_permlane32_swap(b, b, false, false);
    return __int_as_float(p[0]) + __int_as_float(p[1]);
}

// Sum over aligned groups of LPN lanes (16, 32 or 64); every lane gets its group's sum.
template <int LPN>
__device__ __forceinline__ double group_sum(double v)
{
    v += dppd<kDppXor1>(v);
    v += dppd<kDppXor2>(v);
    v += dppd<kDppHalfMirror>(v);
    v += dppd<kDppMirror>(v);
    if (LPN >= 32) v = pair_sumd<true>(v);
    if (LPN >= 64) v = pair_sumd<false>(v);
    return v;
}

template <int LPN>
__device__ __forceinline__ float group_sumf(float v)
{
    v += dppf<kDppXor1>(v);
    v += dppf<kDppXor2>(v);
    v += dppf<kDppHalfMirror>(v);
    v += dppf<kDppMirror>(v);
    if (LPN >= 32) v = pair_sumf<true>(v);
    if (LPN >= 64) v = pair_sumf<false>(v);
    return v;
}

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum<64>(v); }

// ------------------------------------------------------------ LDS tile layout
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Halo record of the FP64 operations: x, y, z, ux, uy, uz.
struct alignas(16) Rec {
    double x, y, z, u, v, w;
};

// Halo record of the FP32 operations, scaled by S = SpecDev::fscale (2^30 over the
// largest reach of a halo): the position as int32 fixed point, rint(X S) of the node
// less the tile reference (an integer in the same units), so xi = X_j - X_i is an exact
// integer difference rounded once to FP32.  Each node is rounded once, whichever tile
// stages it, so a pair's xi is the same integer in both directions and f_ij = -f_ji
// holds to the bit (rounding (X - tile centre) S per tile instead makes the two
// directions differ by a unit in some pairs: the bond forces then no longer cancel in
// pairs, and lightly damped oscillations drift in phase by ~1e-6 over 1,000 steps).  The
// displacement u S in FP64, so eta = u_j - u_i is formed exactly before its rounding
// to FP32 (an FP32 displacement, even relative to a tile reference, rounds eta at the
// scale of the displacement field across the halo instead of eta's own).  One 48-byte
// record per node (x, y, z as int32, ecrit as FP32 bits | ux, uy | uz, bdryEff and
// dx^3 c / 2 as FP32), read as three 16-byte LDS loads from one address.  The bond force
// is scale-free in these units; the energy carries one factor 1/S.
struct alignas(16) RecF {
    u32x4 a;     // fixed-point x, y, z; ecrit (FP32 bits)
    f64x2 u;     // u S (x, y)
    f64x2 w;     // u S (z); bdryEff and dx^3 c / 2 as two FP32 in the second double
};

__device__ __forceinline__ float lo_f(double d) { return __int_as_float((int)__double2loint(d)); }
__device__ __forceinline__ float hi_f(double d) { return __int_as_float((int)__double2hiint(d)); }

// Stage-update inputs of one tile node, loaded while the halo is staged.
struct EpiRec {
    double v[5][3];  // velocity source (V or Vb), U, k1, k23, V: as the stage needs them
    double el;       // elastic energy density of the earlier stages
    int m, b, nb, pad;
};

// H + 1 records (the last one is the padding record), optional per-record ecrit,
// per-node results (4 doubles), row starts (nt + 1 int64), own slots (nt ints);
// for the stage kernels also the stage-update inputs and the material / BC tables.
struct Lds {
    Rec* rec;        // FP64 records
    RecF* recf;      // FP32 records
    double* ec;      // per-record ecrit (FP64 records)
    double* res;
    int64_t* row;
    int* self;
    EpiRec* epi;
    MatDev* mats;
    BCDev* bcs;
    int pad;  // slot of the padding record
};

__host__ __device__ inline size_t round16(size_t b) { return (b + 15) / 16 * 16; }

__host__ __device__ inline size_t lds_rec_bytes(int H, bool with_ecrit, bool f32)
{
    const size_t n = (size_t)H + 1;
    if (f32) return n * sizeof(RecF);
    return n * sizeof(Rec) + (with_ecrit ? round16(n * 8) : 0);
}

__host__ __device__ inline size_t lds_bytes(int H, int tile_max, bool with_ecrit, bool with_epi, bool f32)
{
    size_t b = lds_rec_bytes(H, with_ecrit, f32);
    b += round16(sizeof(double) * 4 * (size_t)tile_max + sizeof(int64_t) * ((size_t)tile_max + 1) +
                 sizeof(int) * (size_t)tile_max);
    if (with_epi) b += sizeof(EpiRec) * (size_t)tile_max + sizeof(MatDev) * kMaxMat + sizeof(BCDev) * kMaxBC;
    return b;
}

__device__ __forceinline__ Lds lds_view(double* base, int H, int tile_max, bool with_ecrit, bool with_epi, bool f32)
{
    Lds l{};
    char* p = reinterpret_cast<char*>(base);
    const size_t n = (size_t)H + 1;
    if (f32) {
        l.recf = reinterpret_cast<RecF*>(p);
        p += n * sizeof(RecF);
    } else {
        l.rec = reinterpret_cast<Rec*>(p);
        p += n * sizeof(Rec);
        if (with_ecrit) {
            l.ec = reinterpret_cast<double*>(p);
            p += round16(n * 8);
        }
    }
    l.res = reinterpret_cast<double*>(p);
    l.row = reinterpret_cast<int64_t*>(l.res + 4 * tile_max);
    l.self = reinterpret_cast<int*>(l.row + tile_max + 1);
    p = reinterpret_cast<char*>(base) + lds_bytes(H, tile_max, with_ecrit, false, f32);
    if (with_epi) {
        l.epi = reinterpret_cast<EpiRec*>(p);
        l.mats = reinterpret_cast<MatDev*>(p + sizeof(EpiRec) * (size_t)tile_max);
        l.bcs = reinterpret_cast<BCDev*>(reinterpret_cast<char*>(l.mats) + sizeof(MatDev) * kMaxMat);
    }
    l.pad = H;
    return l;
}

constexpr unsigned kFarFix = 1u << 30;  // fixed-point coordinate of the padding record

__device__ __forceinline__ void frec_store(const Lds& l, int q, const d4& x, const d4& v, const BK& b, const d4& xr,
                                           double S)
{
    // rint(X S) and the reference are integers below 2^53: their difference is exact
    RecF r;
    r.a = u32x4{(unsigned)__double2int_rn(rint(x.x * S) - xr.x), (unsigned)__double2int_rn(rint(x.y * S) - xr.y),
                (unsigned)__double2int_rn(rint(x.z * S) - xr.z), __float_as_uint((float)x.w)};
    r.u = f64x2{v.x * S, v.y * S};
    r.w = f64x2{v.z * S, __hiloint2double(__float_as_int(b.kc), __float_as_int(b.be))};
    l.recf[q] = r;
}

// Stage the tile's halo records (record q = internal node halo[h0 + q]), the padding
// record, the row starts and own slots.  Two records per thread per pass, both id
// loads issued before the record loads.
template <bool F32>
__device__ __forceinline__ void stage_tile(const TileView& tv, const SpecDev& sp, const d4* __restrict__ u,
                                           int64_t k, int64_t h0, int H, int64_t a0, int nt, const Lds& l,
                                           bool with_ecrit)
{
    d4 xr{};
    const double S = sp.fscale;
    if (F32) xr = tv.tref[k];
    for (int q0 = threadIdx.x; q0 < H; q0 += 2 * blockDim.x) {
        const int q1 = q0 + blockDim.x;
        const int32_t id0 = tv.halo[h0 + q0];
        const int32_t id1 = (q1 < H) ? tv.halo[h0 + q1] : id0;
        const d4 x0 = tv.X[id0], v0 = u[id0];
        const d4 x1 = tv.X[id1], v1 = u[id1];
        if (F32) {
            const BK b0 = tv.bk[id0], b1 = tv.bk[id1];
            frec_store(l, q0, x0, v0, b0, xr, S);
            if (q1 < H) frec_store(l, q1, x1, v1, b1, xr, S);
        } else {
            l.rec[q0] = Rec{x0.x, x0.y, x0.z, v0.x, v0.y, v0.z};
            if (with_ecrit) l.ec[q0] = x0.w;
            if (q1 < H) {
                l.rec[q1] = Rec{x1.x, x1.y, x1.z, v1.x, v1.y, v1.z};
                if (with_ecrit) l.ec[q1] = x1.w;
            }
        }
    }
    if (threadIdx.x == 0) {
        if (F32) {
            RecF r;
            r.a = u32x4{kFarFix, kFarFix, kFarFix, __float_as_uint(1.0f)};
            r.u = f64x2{0.0, 0.0};
            r.w = f64x2{0.0, 0.0};  // bdryEff 0, c 0: coefficient 0
            l.recf[H] = r;
        } else {
            l.rec[H] = Rec{kFar, kFar, kFar, 0.0, 0.0, 0.0};
            if (with_ecrit) l.ec[H] = 1.0;
        }
    }
    for (int q = threadIdx.x; q <= nt; q += blockDim.x) l.row[q] = tv.rpi[a0 + q];
    for (int q = threadIdx.x; q < nt; q += blockDim.x) l.self[q] = tv.self_slot[a0 + q];
}

// The family owner's own record.
struct Self {
    double x, y, z, u, v, w, ecrit;
    __device__ static Self make(const Lds& l, const SpecDev& sp, int ss, bool ue)
    {
        const Rec r = l.rec[ss];
        return Self{r.x, r.y, r.z, r.u, r.v, r.w, ue ? sp.ecrit0 : l.ec[ss]};
    }
};

struct SelfF {
    int x, y, z;
    double u, v, w;
    float ecrit;
    float be, kc;         // bdryEff, dx^3 c / 2
    float beta0, beta1;   // beta(r) constants of the specimen
    __device__ static SelfF make(const Lds& l, const SpecDev& sp, int ss, bool ue)
    {
        const RecF& r = l.recf[ss];
        const u32x4 a = r.a;
        const f64x2 b = r.u;
        const f64x2 w = r.w;
        return SelfF{(int)a.x, (int)a.y, (int)a.z, b.x, b.y, w.x, ue ? (float)sp.ecrit0 : __uint_as_float(a.w),
                     lo_f(w.y), hi_f(w.y), sp.beta0, sp.beta1};
    }
};

// ------------------------------------------------------------ per-bond operations
// Each operation: the coefficient type it streams, its record type (Self / SelfF),
// a per-lane accumulator, the bond update, the lane-group reduction and the store
// of (fx, fy, fz, energy density).  Energy: acc.el sums coef * s * u; the bond energy
// bdryEff * beta * c * u^2 / (4 r) is that times 1 / (4 dx^3), applied once per family.

// Correctly rounded FP64 square root for the bond lengths: v_rsq_f64 refined by the same
// Newton-Raphson sequence the compiler emits for sqrt(), without the input scaling and the
// zero / infinity selects that sequence spends on operands outside [2^-767, 2^1000] (one
// rarely taken branch falls back to sqrt() there).  r and p must equal the reference's
// libm square roots to the bit: u = p - r cancels all but a few of their digits.
__device__ __forceinline__ double sqrt_rn(double x)
{
    if (!(x >= 0x1p-767 && x <= 0x1p+1000)) return sqrt(x);
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

struct OpF64 {  // PDM_FP64, reference arithmetic
    using Coef = double;
    using SelfT = Self;
    struct Acc {
        double f0, f1, f2, el;
    };
    template <bool ENERGY, bool UE>
    __device__ static void bond(const Lds& l, unsigned s, double c, const Self& me, Acc& a)
    {
        const Rec r = l.rec[s];
        double xi1 = r.x - me.x, xi2 = r.y - me.y, xi3 = r.z - me.z;
        double et1 = r.u - me.u, et2 = r.v - me.v, et3 = r.w - me.w;
        double rr = sqrt_rn(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
        double a1 = xi1 + et1, a2 = xi2 + et2, a3 = xi3 + et3;
        double p = sqrt_rn(a1 * a1 + a2 * a2 + a3 * a3);
        double u = p - rr;
        // r and p, hence u, are the reference's to the bit; its two divisions (u / r and
        // f / p) share one reciprocal here, 1/r = p / (r p) and 1/p = r / (r p), which moves
        // each bond force by a few ulp of itself
        const double inv = 1.0 / (rr * p);
        double st = u * (p * inv);
        double ecrit = me.ecrit;
        if (!UE) {
            double ej = l.ec[s];
            ecrit = (ej < me.ecrit) ? ej : me.ecrit;
        }
        // u <= 0 || (u > 0 && st <= ecrit), evaluated without branches (st has u's sign; a NaN
        // stretch fails both tests, as in the reference)
        const bool ok = (u <= 0.0) | (st <= ecrit);
        const double cs = ok ? c * st : 0.0;
        const double g = cs * (rr * inv);
        a.f0 = __builtin_fma(g, a1, a.f0);
        a.f1 = __builtin_fma(g, a2, a.f1);
        a.f2 = __builtin_fma(g, a3, a.f2);
        if (ENERGY) a.el = __builtin_fma(cs, u, a.el);
    }
    template <int LPN>
    __device__ static void reduce(Acc& a)
    {
        a.f0 = group_sum<LPN>(a.f0);
        a.f1 = group_sum<LPN>(a.f1);
        a.f2 = group_sum<LPN>(a.f2);
        a.el = group_sum<LPN>(a.el);
    }
    __device__ static void store(const Acc& a, double ivc4, double* out)
    {
        out[0] = a.f0;
        out[1] = a.f1;
        out[2] = a.f2;
        out[3] = a.el * ivc4;
    }
};

struct Bond32 {
    float g, a1, a2, a3, e, cs;  // force factor, xi + eta, u, coef * s
};

// ---- coefficients of the FP32 force operations
// coef = bdryEff_ij dx^3 beta(r) c_ij is formed from the two records as
// (be_i + be_j) min(kc_i, kc_j) (beta_32(r) + (corr - 128) 2^-26), kc = dx^3 c / 2, where
// beta_32 is beta evaluated in FP32 from the record geometry and corr (uint8 per entry)
// takes beta_32 + (corr - 128) 2^-26 to coef64 / ((be_i + be_j) min(kc_i, kc_j)) within
// 2^-27 (k_fly_corr); the FP32 rounding of that sum and of the product then leave the
// formed coefficient within about one FP32 ulp of the FP64 one.  One byte per entry
// instead of an FP32 coefficient, and the correction also absorbs the FP32 rounding of
// the constants and of be, kc.  The -128 2^-26 offset is folded into beta_32's constant and
// clamp bounds.  CCI rows hold no entry with coefficient 0 (cut bonds and beta 0 are
// left out when the rows are built); row padding reads the padding record (be = kc = 0).
constexpr float kCorrStep = 0x1p-26f;

// r^2, 1/r and r of a record pair, in record units (S m).
__device__ __forceinline__ void rec_r(const u32x4& ra, const SelfF& me, float& r2, float& rinv, float& rr)
{
    const float x1 = (float)((int)ra.x - me.x), x2 = (float)((int)ra.y - me.y), x3 = (float)((int)ra.z - me.z);
    r2 = __builtin_fmaf(x3, x3, __builtin_fmaf(x2, x2, x1 * x1));
    rinv = __builtin_amdgcn_rsqf(r2);
    rr = r2 * rinv;
}

// beta (pdDataManager3.cpp:1801-1831) in FP32, (delta + dx/2 - r) / dx clamped to [0, 1], minus
// 128 2^-26.
__device__ __forceinline__ float fly_beta(float rr, const SelfF& me)
{
    return __builtin_amdgcn_fmed3f(__builtin_fmaf(-rr, me.beta1, me.beta0), -128.0f * kCorrStep,
                                   1.0f - 128.0f * kCorrStep);
}

// (be_i + be_j) min(kc_i, kc_j) from the neighbour record's second double; with one
// material per specimen (UE) only be_i + be_j, and the family sums are scaled by kc0 in
// FP64 when they are stored.
template <bool UE>
__device__ __forceinline__ float fly_base(const SelfF& me, double wy)
{
    if (UE) return me.be + lo_f(wy);
    return (me.be + lo_f(wy)) * __builtin_fminf(me.kc, hi_f(wy));
}

// FP32 bond force from the scaled FP32 records, cancellation-free stretch:
// q = |xi+eta|^2 - |xi|^2 = eta . (2 xi + eta), s = (p - r) / r = q / ((p + r) r).
// cu: the entry's coefficient correction byte, as a float.
template <bool ENERGY, bool UE>
__device__ __forceinline__ Bond32 bond32(const Lds& l, unsigned s, float cu, const SelfF& me)
{
    const RecF& rec = l.recf[s];
    const u32x4 ra = rec.a;
    const f64x2 rb = rec.u;
    const f64x2 rw = rec.w;
    const double rz = rw.x;
    const float x1 = (float)((int)ra.x - me.x), x2 = (float)((int)ra.y - me.y), x3 = (float)((int)ra.z - me.z);
    const float e1 = (float)(rb.x - me.u), e2 = (float)(rb.y - me.v), e3 = (float)(rz - me.w);
    const float a1 = x1 + e1, a2 = x2 + e2, a3 = x3 + e3;
    float r2, rinv, rr;
    rec_r(ra, me, r2, rinv, rr);
    const float q = __builtin_fmaf(e3, x3 + a3, __builtin_fmaf(e2, x2 + a2, e1 * (x1 + a1)));
    const float p2 = r2 + q;
    const float pinv = __builtin_amdgcn_rsqf(p2);
    const float p = p2 * pinv;
    const float st = q * __builtin_amdgcn_rcpf((p + rr) * rr);
    const float c = fly_base<UE>(me, rw.y) * __builtin_fmaf(cu, kCorrStep, fly_beta(rr, me));
    // u <= 0 || s <= ecrit, with u = s r and r > 0
    float ecrit = me.ecrit;
    if (!UE) {
        const float ej = __uint_as_float(ra.w);
        ecrit = (ej < ecrit) ? ej : ecrit;
    }
    const float cs = (st <= __builtin_fmaxf(ecrit, 0.0f)) ? c * st : 0.0f;
    Bond32 b;
    b.g = cs * pinv;
    b.a1 = a1;
    b.a2 = a2;
    b.a3 = a3;
    b.e = ENERGY ? st * rr : 0.0f;  // u; the energy term is cs u
    b.cs = cs;
    return b;
}

struct OpF32 {  // PDM_FP32C: FP32 bond force, FP32 sums per lane and across the lane group
    using Coef = float;
    static constexpr bool kCoef = false;  // coefficients formed from the records and correction bytes
    using SelfT = SelfF;
    struct Acc {
        float p0, p1, p2;
        float el;
    };
    template <bool ENERGY, bool UE>
    __device__ static void bond(const Lds& l, unsigned s, float c, const SelfF& me, Acc& a)
    {
        const Bond32 b = bond32<ENERGY, UE>(l, s, c, me);
        a.p0 = __builtin_fmaf(b.g, b.a1, a.p0);
        a.p1 = __builtin_fmaf(b.g, b.a2, a.p1);
        a.p2 = __builtin_fmaf(b.g, b.a3, a.p2);
        if (ENERGY) a.el = __builtin_fmaf(b.cs, b.e, a.el);
    }
    template <int LPN>
    __device__ static void reduce(Acc& a)
    {
        a.p0 = group_sumf<LPN>(a.p0);
        a.p1 = group_sumf<LPN>(a.p1);
        a.p2 = group_sumf<LPN>(a.p2);
        a.el = group_sumf<LPN>(a.el);
    }
    __device__ static void store(const Acc& a, double ivc4, double* out)
    {
        out[0] = (double)a.p0;
        out[1] = (double)a.p1;
        out[2] = (double)a.p2;
        out[3] = (double)a.el * ivc4;
    }
};

// PDM_FP32C2: FP32 bond forces, two-float (hi, lo) family sums.  Each lane sums blocks of
// kBlockN bond forces in FP32 (the block partial p) and adds each block partial to its
// (hi, lo) pair with TwoSum, so the only roundings besides those of the bond forces
// themselves are the (kBlockN - 1) additions inside a block, each on a partial of kBlockN
// terms.  Each lane's hi + lo is exact in FP64, and the lane group sums those in FP64 (16
// to 64 terms, DPP and permlane steps on doubles).  A TwoSum per bond and TwoSum at every
// butterfly step (the first version) rounded nothing less that matters -- every term is
// already an FP32 bond force -- and cost 35 % of the stage time against 17 % for this.
struct OpF32TF {
    using Coef = float;
    static constexpr bool kCoef = false;
    using SelfT = SelfF;
    struct Acc {
        float p0, p1, p2;  // FP32 partials of the current block
        float h0, h1, h2, l0, l1, l2, el;
        double s0, s1, s2;  // the family sums, after reduce
    };
    template <bool ENERGY, bool UE>
    __device__ static void bond(const Lds& l, unsigned s, float c, const SelfF& me, Acc& a)
    {
        const Bond32 b = bond32<ENERGY, UE>(l, s, c, me);
        a.p0 = __builtin_fmaf(b.g, b.a1, a.p0);
        a.p1 = __builtin_fmaf(b.g, b.a2, a.p1);
        a.p2 = __builtin_fmaf(b.g, b.a3, a.p2);
        if (ENERGY) a.el = __builtin_fmaf(b.cs, b.e, a.el);
    }
    __device__ static void flush(Acc& a)
    {
        float t, err;
        two_sum(a.h0, a.p0, t, err); a.h0 = t; a.l0 += err;
        two_sum(a.h1, a.p1, t, err); a.h1 = t; a.l1 += err;
        two_sum(a.h2, a.p2, t, err); a.h2 = t; a.l2 += err;
        a.p0 = a.p1 = a.p2 = 0.0f;
    }
    template <int LPN>
    __device__ static void reduce(Acc& a)
    {
        a.s0 = group_sum<LPN>((double)a.h0 + (double)a.l0);
        a.s1 = group_sum<LPN>((double)a.h1 + (double)a.l1);
        a.s2 = group_sum<LPN>((double)a.h2 + (double)a.l2);
        a.el = group_sumf<LPN>(a.el);
    }
    __device__ static void store(const Acc& a, double ivc4, double* out)
    {
        out[0] = a.s0;
        out[1] = a.s1;
        out[2] = a.s2;
        out[3] = (double)a.el * ivc4;
    }
};

// Tangent of the bond force at eta = 0: y_i = sum_j w_ij xi (xi . (u_i - u_j)), with the
// per-entry tangent weight w = coef / r^3 streamed in place of the force coefficient
// (k_tangent_coef builds it once: the operator is fixed, so no root or division per bond).
// W = double: the FP64 operator; W = float: the weights stored in FP32 (a matvec streams 6
// instead of 10 bytes per entry), the arithmetic in FP64.  FP32 rounding of the weights
// leaves the operator symmetric and fixed, so CG keeps its FP64 convergence up to the small
// operator difference; FP32 arithmetic instead adds rounding noise to every product, which
// slows CG's convergence 2-3 fold.
template <class W>
struct OpTangent {
    using Coef = W;
    using SelfT = Self;
    struct Acc {
        double f0, f1, f2, el;
    };
    template <bool ENERGY, bool UE>
    __device__ static void bond(const Lds& l, unsigned s, W c, const Self& me, Acc& a)
    {
        const Rec r = l.rec[s];
        const double x1 = r.x - me.x, x2 = r.y - me.y, x3 = r.z - me.z;
        const double d0 = me.u - r.u, d1 = me.v - r.v, d2 = me.w - r.w;
        const double proj = __builtin_fma(x3, d2, __builtin_fma(x2, d1, x1 * d0)) * (double)c;
        a.f0 = __builtin_fma(proj, x1, a.f0);
        a.f1 = __builtin_fma(proj, x2, a.f1);
        a.f2 = __builtin_fma(proj, x3, a.f2);
    }
    template <int LPN>
    __device__ static void reduce(Acc& a)
    {
        a.f0 = group_sum<LPN>(a.f0);
        a.f1 = group_sum<LPN>(a.f1);
        a.f2 = group_sum<LPN>(a.f2);
    }
    __device__ static void store(const Acc& a, double, double* out)
    {
        out[0] = a.f0;
        out[1] = a.f1;
        out[2] = a.f2;
        out[3] = 0.0;
    }
};
using OpT64 = OpTangent<double>;
using OpT32 = OpTangent<float>;
// PDM_FP32C2's operator: each weight rounded to nearest at 39 significant bits (relative
// error at most 2^-40) and streamed as the FP32 truncation of that plus its next 16
// significand bits (tc16), rebuilt exactly when used by placing those bits below the
// converted FP32 significand: 4 + 2 bytes per entry besides the slot.  The FP32 rounding of
// the weights alone perturbs the smooth, slowly converging modes of the tangent by ~6e-8
// (L/dx)^2 relative (their bond terms cancel), which costs CG a third more iterations; near
// a 1e-10 residual even a 2^-38 truncation bias costs 15 %.
struct OpT32C : OpTangent<double> {
    static constexpr bool kWeightLo = true;
};

// Diagonal of the tangent operator (Jacobi preconditioner): d_c = sum_j w_ij xi_c^2, w the
// tangent weight coef / r^3.
struct OpD64 {
    using Coef = double;
    using SelfT = Self;
    struct Acc {
        double f0, f1, f2, el;
    };
    template <bool ENERGY, bool UE>
    __device__ static void bond(const Lds& l, unsigned s, double c, const Self& me, Acc& a)
    {
        const Rec r = l.rec[s];
        const double x1 = r.x - me.x, x2 = r.y - me.y, x3 = r.z - me.z;
        a.f0 = __builtin_fma(c * x1, x1, a.f0);
        a.f1 = __builtin_fma(c * x2, x2, a.f1);
        a.f2 = __builtin_fma(c * x3, x3, a.f2);
    }
    template <int LPN>
    __device__ static void reduce(Acc& a)
    {
        OpT64::template reduce<LPN>(reinterpret_cast<OpT64::Acc&>(a));
    }
    __device__ static void store(const Acc& a, double, double* out)
    {
        out[0] = a.f0;
        out[1] = a.f1;
        out[2] = a.f2;
        out[3] = 0.0;
    }
};

// The off-diagonal entries of each node's 3x3 tangent block: sum_j w xi_1 xi_2, xi_1 xi_3,
// xi_2 xi_3 (OpD64 gives the diagonal), for the block-Jacobi preconditioner.
struct OpD64X : OpD64 {
    template <bool ENERGY, bool UE>
    __device__ static void bond(const Lds& l, unsigned s, double c, const Self& me, Acc& a)
    {
        const Rec r = l.rec[s];
        const double x1 = r.x - me.x, x2 = r.y - me.y, x3 = r.z - me.z;
        a.f0 = __builtin_fma(c * x1, x2, a.f0);
        a.f1 = __builtin_fma(c * x1, x3, a.f1);
        a.f2 = __builtin_fma(c * x2, x3, a.f2);
    }
};

// Whether an operation reads FP32 records.
template <class OP>
constexpr bool kF32Rec = std::is_same<typename OP::SelfT, SelfF>::value;

// Whether an operation streams a per-entry coefficient (the FP32 force operations form
// theirs from the records and stream one correction byte per entry instead).
template <class OP, class = void>
struct CoefStream {
    static constexpr bool value = true;
};
template <class OP>
struct CoefStream<OP, std::enable_if_t<!OP::kCoef>> {
    static constexpr bool value = false;
};
template <class OP>
constexpr bool kCoefStream = CoefStream<OP>::value;

// Whether an operation streams FP32 weights with a 16-bit correction each (OpT32C).
template <class OP, class = void>
struct WeightLo {
    static constexpr bool value = false;
};
template <class OP>
struct WeightLo<OP, std::enable_if_t<OP::kWeightLo>> {
    static constexpr bool value = true;
};
template <class OP>
constexpr bool kWeightLo = WeightLo<OP>::value;

// What a block holds per entry besides its slot: 0 the streamed coefficient, 1 the
// correction byte of the FP32 force operations, 2 an FP32 weight and its 16-bit correction.
template <class OP>
constexpr int kChunkKind = kWeightLo<OP> ? 2 : kCoefStream<OP> ? 0 : 1;

// Operations without FP32 partial sums have nothing to flush.
template <class OP, class = void>
struct HasFlush {
    static constexpr bool value = false;
};
template <class OP>
struct HasFlush<OP, decltype(OP::flush(std::declval<typename OP::Acc&>()))> {
    static constexpr bool value = true;
};
template <class OP>
__device__ __forceinline__ void flush(typename OP::Acc& a)
{
    if constexpr (HasFlush<OP>::value) OP::flush(a);
}

// One lane's block: B consecutive family entries, slots packed two per word, and the
// coefficients or, for the FP32 force operations, the correction bytes packed four per word.
template <class OP, int B, int KIND = kChunkKind<OP>>
struct Chunk {
    unsigned sp[B / 2];
    typename OP::Coef c[B];
    __device__ __forceinline__ unsigned slot(int k) const { return (sp[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu; }
    __device__ __forceinline__ typename OP::Coef coef(int k) const { return c[k]; }
};
template <class OP, int B>
struct Chunk<OP, B, 2> {
    unsigned sp[B / 2];
    float w[B];
    unsigned lo[B / 2];  // two 16-bit corrections per word
    __device__ __forceinline__ unsigned slot(int k) const { return (sp[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu; }
    __device__ __forceinline__ double coef(int k) const
    {
        // FP64 significand bits 28..13 (an FP32 significand fills 51..29 after conversion)
        const unsigned x = (k & 1) ? (lo[k >> 1] >> 3) : (lo[k >> 1] << 13);
        const double d = (double)w[k];
        return __hiloint2double(__double2hiint(d), __double2loint(d) | (int)(x & 0x1FFFE000u));
    }
};
template <class OP, int B>
struct Chunk<OP, B, 1> {
    unsigned sp[B / 2];
    unsigned cb[B / 4];
    __device__ __forceinline__ unsigned slot(int k) const { return (sp[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu; }
    __device__ __forceinline__ float coef(int k) const { return (float)((cb[k >> 2] >> ((k & 3) * 8)) & 0xFFu); }
};

// Entries [e, e + 8) of a row ending at e1 (rows are whole blocks of kRowAlign = 8
// entries), or, at or past e1, 8 references to the padding record with coefficient 0.
template <class OP, int B>
__device__ __forceinline__ void load8(const TileView& tv, int64_t e, int64_t e1, unsigned pad2, Chunk<OP, B>& ch,
                                      int o)
{
    static_assert(kRowAlign == 8, "sub-blocks of 8 entries");
    if (e < e1) {
        const u32x4 sv = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(tv.slot + e));
#pragma unroll
        for (int k = 0; k < 4; ++k) ch.sp[o / 2 + k] = sv[k];
        if constexpr (kWeightLo<OP>) {
            const f32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tv.coef32 + e));
            const f32x4 c1 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tv.coef32 + e + 4));
            const u32x4 lv = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(tv.wlo + e));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ch.w[o + k] = c0[k];
                ch.w[o + 4 + k] = c1[k];
                ch.lo[o / 2 + k] = lv[k];
            }
        } else if constexpr (!kCoefStream<OP>) {
            const u32x2 cv = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(tv.corr + e));
            ch.cb[o / 4] = cv[0];
            ch.cb[o / 4 + 1] = cv[1];
        } else if constexpr (sizeof(typename OP::Coef) == 4) {
            const f32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tv.coef32 + e));
            const f32x4 c1 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tv.coef32 + e + 4));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ch.c[o + k] = c0[k];
                ch.c[o + 4 + k] = c1[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const f64x2 cv = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(tv.coef64 + e + 2 * k));
                ch.c[o + 2 * k] = cv[0];
                ch.c[o + 2 * k + 1] = cv[1];
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) ch.sp[o / 2 + k] = pad2;
        if constexpr (kWeightLo<OP>) {
#pragma unroll
            for (int k = 0; k < 8; ++k) ch.w[o + k] = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) ch.lo[o / 2 + k] = 0;
        } else if constexpr (kCoefStream<OP>) {
#pragma unroll
            for (int k = 0; k < 8; ++k) ch.c[o + k] = 0;
        } else {
            ch.cb[o / 4] = ch.cb[o / 4 + 1] = 0x80808080u;  // zero corrections
        }
    }
}

// Entries [e, e + 4) of a row ending at e1 into a chunk at entry offset o (e a multiple of
// 4, rows whole blocks of 8), or, at or past e1, 4 references to the padding record with
// coefficient 0: an 8-byte slot load and a 4-byte correction load or one or two 16-byte
// coefficient loads.
template <class OP, int B>
__device__ __forceinline__ void load4(const TileView& tv, int64_t e, int64_t e1, unsigned pad2, Chunk<OP, B>& ch,
                                      int o)
{
    if (e < e1) {
        const u32x2 sv = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(tv.slot + e));
        ch.sp[o / 2] = sv[0];
        ch.sp[o / 2 + 1] = sv[1];
        if constexpr (kWeightLo<OP>) {
            const f32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tv.coef32 + e));
            const u32x2 lv = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(tv.wlo + e));
#pragma unroll
            for (int k = 0; k < 4; ++k) ch.w[o + k] = c0[k];
            ch.lo[o / 2] = lv[0];
            ch.lo[o / 2 + 1] = lv[1];
        } else if constexpr (!kCoefStream<OP>) {
            ch.cb[o / 4] = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(tv.corr + e));
        } else if constexpr (sizeof(typename OP::Coef) == 4) {
            const f32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(tv.coef32 + e));
#pragma unroll
            for (int k = 0; k < 4; ++k) ch.c[o + k] = c0[k];
        } else {
            const f64x2 c0 = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(tv.coef64 + e));
            const f64x2 c1 = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(tv.coef64 + e + 2));
            ch.c[o] = c0[0];
            ch.c[o + 1] = c0[1];
            ch.c[o + 2] = c1[0];
            ch.c[o + 3] = c1[1];
        }
    } else {
        ch.sp[o / 2] = ch.sp[o / 2 + 1] = pad2;
        if constexpr (kWeightLo<OP>) {
#pragma unroll
            for (int k = 0; k < 4; ++k) ch.w[o + k] = 0.0f;
            ch.lo[o / 2] = ch.lo[o / 2 + 1] = 0;
        } else if constexpr (kCoefStream<OP>) {
#pragma unroll
            for (int k = 0; k < 4; ++k) ch.c[o + k] = 0;
        } else {
            ch.cb[o / 4] = 0x80808080u;
        }
    }
}

// A family is read as a first block of kBlock0 entries per lane, loaded during the
// previous round, then blocks of 4 entries per lane until the row ends, each loaded while
// the one before is computed.  FP32 operations: the first block is 8 entries at e0 + 8 l
// plus 4 at e0 + 8 L + 4 l (lane l of L), so with 16 lanes it covers 192 entries, an
// interior CCI family at delta = 3.015 dx (178 entries with beta > 0) whole; the FP64
// operations keep 4, and the two-float sums 8 at e0 + 8 l in the 80-VGPR kernel (their
// accumulators take six more registers; 12 would spill there).  The later blocks of 4 lie
// at e0 + kBlock0 L + 4 l, then 4 L further each time.  Each load instruction of a lane
// group covers contiguous memory.
template <class OP, bool OCC6 = false>
constexpr int kBlock0 = (sizeof(typename OP::Coef) == 8) ? 4 : (HasFlush<OP>::value && OCC6) ? 8 : 12;
constexpr int kBlockN = 4;

template <class OP, int LPN, int B0>
__device__ __forceinline__ void load_first(const TileView& tv, int64_t e0, int sl, int64_t e1, int pad,
                                           Chunk<OP, B0>& ch)
{
    const unsigned pad2 = (unsigned)pad * 0x10001u;
    if constexpr (B0 == 12) {
        load8<OP, 12>(tv, e0 + 8 * sl, e1, pad2, ch, 0);
        load4<OP, 12>(tv, e0 + 8 * LPN + 4 * sl, e1, pad2, ch, 8);
    } else if constexpr (B0 == 8) {
        load8<OP, 8>(tv, e0 + 8 * sl, e1, pad2, ch, 0);
    } else {
        load4<OP, 4>(tv, e0 + 4 * sl, e1, pad2, ch, 0);
    }
}

// One round of a lane group: the family of tile node loc (first block from `cur`, which the
// caller loaded), with the next round's first block loaded into the caller's other buffer
// beforehand.
template <class OP, bool ENERGY, bool UE, int LPN, int B0>
__device__ __forceinline__ void family_round(const TileView& tv, const SpecDev& sp, const Lds& l, int nt, int loc,
                                             int64_t e0, int64_t e1, int sl, const Chunk<OP, B0>& cur, double ivc4)
{
    const int ss = (loc < nt) ? l.self[loc] : l.pad;
    const typename OP::SelfT me = OP::SelfT::make(l, sp, ss, UE);
    typename OP::Acc acc{};
    int64_t base = e0 + LPN * B0;
    const unsigned pad2 = (unsigned)l.pad * 0x10001u;
    Chunk<OP, kBlockN> nx;
    if (base < e1) load4<OP, kBlockN>(tv, base + kBlockN * sl, e1, pad2, nx, 0);
#pragma unroll
    for (int k = 0; k < B0; ++k) {
        OP::template bond<ENERGY, UE>(l, cur.slot(k), cur.coef(k), me, acc);
        if (k % kBlockN == kBlockN - 1) flush<OP>(acc);  // the first block in blocks of kBlockN
    }
    while (base < e1) {
        const Chunk<OP, kBlockN> c = nx;
        base += LPN * kBlockN;
        if (base < e1) load4<OP, kBlockN>(tv, base + kBlockN * sl, e1, pad2, nx, 0);
#pragma unroll
        for (int k = 0; k < kBlockN; ++k) OP::template bond<ENERGY, UE>(l, c.slot(k), c.coef(k), me, acc);
        flush<OP>(acc);
    }
    OP::template reduce<LPN>(acc);
    if (sl == 0 && loc < nt) {
        double* out = l.res + 4 * loc;
        OP::store(acc, ivc4, out);
        if constexpr (!kCoefStream<OP> && UE)
            for (int c = 0; c < 4; ++c) out[c] *= sp.kc0;  // the coefficient's common factor
    }
}

// Family sums of the tile's nodes into l.res.  Round r of wavefront w covers nodes
// (r NW + w) NPW + g for lane groups g = 0..NPW-1; the first block of round r + 1
// is loaded before round r is computed.  Rounds alternate between two block
// buffers, so no registers are copied between rounds.
template <class OP, bool ENERGY, bool UE, int NW, int NPW, bool OCC6 = false>
__device__ __forceinline__ void tile_families(const TileView& tv, const SpecDev& sp, const Lds& l, int nt)
{
    constexpr int LPN = 64 / NPW;
    constexpr int B = kBlock0<OP, OCC6>;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane / LPN, sl = lane % LPN;
    // FP32 records are scaled by S: the energy sum coef s (s r S) carries one factor S
    // (the host forms both scales once, so they stay in scalar registers)
    const double ivc4 = kF32Rec<OP> ? sp.ivc4_rec : sp.ivc4;
    int first = wave * NPW;
    if (first >= nt) return;
    int loc = first + g;
    int64_t e0 = (loc < nt) ? l.row[loc] : 0, e1 = (loc < nt) ? l.row[loc + 1] : 0;
    Chunk<OP, B> c0, c1;
    load_first<OP, LPN, B>(tv, e0, sl, e1, l.pad, c0);
    if constexpr (!kF32Rec<OP>) {
        // FP64 operations: one block buffer and a copy per round (fewer registers)
        for (;;) {
            const int nfirst = first + NW * NPW;
            const int nloc = nfirst + g;
            int64_t n0 = 0, n1 = 0;
            if (nfirst < nt) {
                if (nloc < nt) {
                    n0 = l.row[nloc];
                    n1 = l.row[nloc + 1];
                }
                load_first<OP, LPN, B>(tv, n0, sl, n1, l.pad, c1);
            }
            family_round<OP, ENERGY, UE, LPN, B>(tv, sp, l, nt, loc, e0, e1, sl, c0, ivc4);
            if (nfirst >= nt) return;
            first = nfirst;
            loc = nloc;
            e0 = n0;
            e1 = n1;
            c0 = c1;
        }
    }
    for (;;) {
        // round with c0, prefetch into c1
        int nfirst = first + NW * NPW;
        int nloc = nfirst + g;
        int64_t n0 = 0, n1 = 0;
        if (nfirst < nt) {
            if (nloc < nt) {
                n0 = l.row[nloc];
                n1 = l.row[nloc + 1];
            }
            load_first<OP, LPN, B>(tv, n0, sl, n1, l.pad, c1);
        }
        family_round<OP, ENERGY, UE, LPN, B>(tv, sp, l, nt, loc, e0, e1, sl, c0, ivc4);
        if (nfirst >= nt) break;
        first = nfirst;
        loc = nloc;
        e0 = n0;
        e1 = n1;
        // round with c1, prefetch into c0
        nfirst = first + NW * NPW;
        nloc = nfirst + g;
        n0 = 0;
        n1 = 0;
        if (nfirst < nt) {
            if (nloc < nt) {
                n0 = l.row[nloc];
                n1 = l.row[nloc + 1];
            }
            load_first<OP, LPN, B>(tv, n0, sl, n1, l.pad, c0);
        }
        family_round<OP, ENERGY, UE, LPN, B>(tv, sp, l, nt, loc, e0, e1, sl, c1, ivc4);
        if (nfirst >= nt) break;
        first = nfirst;
        loc = nloc;
        e0 = n0;
        e1 = n1;
    }
}

__device__ __forceinline__ double time_factor_dev(const BCDev& b, double t)
{
    if (b.type == 2) return (t <= b.t_end) ? 1.0 : 0.0;
    if (b.load == 1.0) return (t <= b.t_end) ? t / b.t_end : 1.0;
    return 1.0;
}

template <int PREC>
struct OpFor;
template <>
struct OpFor<PDM_FP64> {
    using Force = OpF64;
    using Tangent = OpT64;
};
template <>
struct OpFor<PDM_FP32C> {
    using Force = OpF32;
    using Tangent = OpT32;
};
template <>
struct OpFor<PDM_FP32C2> {
    using Force = OpF32TF;
    using Tangent = OpT32;
};

// Stage-update inputs of internal node i as stage STAGE needs them.
template <int STAGE, bool ENERGY>
struct EpiLoad {
    d4 v0, v1, v2, v3, v4;
    double el;
    int m, b, nb;
    __device__ __forceinline__ void load(const StageArgs& a, int64_t i)
    {
        v0 = (STAGE == 0) ? a.V[i] : a.Vb[i];
        v1 = a.U[i];
        if (STAGE == 3) v2 = a.k1[i];
        if (STAGE >= 2) v3 = a.k23[i];
        if (STAGE == 3) v4 = a.V[i];
        if (ENERGY && STAGE > 0) el = a.elast[i];
        m = a.mat[i];
        b = a.bc[i];
        nb = a.nb[i];
    }
    __device__ __forceinline__ void store(EpiRec& e) const
    {
        const d4* v[5] = {&v0, &v1, &v2, &v3, &v4};
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            if ((q == 2 && STAGE != 3) || (q == 3 && STAGE < 2) || (q == 4 && STAGE != 3)) continue;
            e.v[q][0] = v[q]->x;
            e.v[q][1] = v[q]->y;
            e.v[q][2] = v[q]->z;
        }
        if (ENERGY && STAGE > 0) e.el = el;
        e.m = m;
        e.b = b;
        e.nb = nb;
    }
};

// The specimen's material and BC tables into LDS.
__device__ __forceinline__ void load_tables(const StageArgs& a, int spec, const Lds& l)
{
    static_assert(sizeof(MatDev) % 8 == 0 && sizeof(BCDev) % 8 == 0, "tables are double words");
    constexpr int nm = (int)(sizeof(MatDev) * kMaxMat / sizeof(double));
    constexpr int nb = (int)(sizeof(BCDev) * kMaxBC / sizeof(double));
    const double* ms = reinterpret_cast<const double*>(a.mats + (int64_t)spec * kMaxMat);
    const double* bs = reinterpret_cast<const double*>(a.bcs + (int64_t)spec * kMaxBC);
    double* md = reinterpret_cast<double*>(l.mats);
    double* bd = reinterpret_cast<double*>(l.bcs);
    for (int q = threadIdx.x; q < nm + nb; q += blockDim.x) {
        if (q < nm) md[q] = ms[q];
        else bd[q - nm] = bs[q - nm];
    }
}

// The stage update of a tile's nodes from their family sums in l.res (fx, fy, fz, energy
// density): velocity buffer, BC overwrite, k, the next stage's displacements and, at stage 4,
// the RK4 update and the tile's KE / EE partials (CalcBondBasedMotionRK4's tail and RK4,
// pdDataManager.cpp:1958-1990, 2058-2110).
template <int STAGE, bool ENERGY, int NW>
__device__ __forceinline__ void tile_update(const StageArgs& a, const SpecDev& sp, const Lds& l, int64_t k, int64_t a0,
                                            int nt, int step_idx, int spec, double (*sh_e)[NW])
{
    const TileView& tv = a.tv;
    const StepParam stp = a.steps[(int64_t)step_idx * a.nspec + spec];
    const double dt = stp.dt;
    const double dts = (STAGE < 2) ? 0.5 * dt : dt;
    double ke_i = 0.0, ee_i = 0.0;
    for (int loc = threadIdx.x; loc < nt; loc += 64 * NW) {
        const int64_t i = a0 + loc;
        const double* res = l.res + 4 * loc;
        const EpiRec& ep = l.epi[loc];
        const int m = ep.m;
        const int b = ep.b;
        const MatDev& md = l.mats[m];
        const double rho = md.rho;
        const double nbi = (double)ep.nb;
        double Ft[3];
        for (int c = 0; c < 3; ++c) Ft[c] = res[c] + nbi * (md.g[c] * rho);
        const double eli = res[3];
        bool bcnode = false;
        double fac = 1.0;
        if (b >= 0) {
            const BCDev& bd = l.bcs[b];
            bcnode = bd.bcnode != 0;
            fac = time_factor_dev(bd, stp.t);
            if (bd.type == 4)
                for (int c = 0; c < 3; ++c) Ft[c] += bd.trac[c] * fac;
        }
        double vbp[3], Up[3];
        for (int c = 0; c < 3; ++c) {
            vbp[c] = ep.v[0][c];
            Up[c] = ep.v[1][c];
        }
        for (int c = 0; c < 3; ++c) vbp[c] += Ft[c] * dts / rho;
        // UpdateNodeStatusWithBC (pdDataManager.cpp:2497-2602)
        if (b >= 0) {
            const BCDev& bd = l.bcs[b];
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
        double nup[3];
        if (STAGE == 0) {
            a.k1[i] = d4{kk[0], kk[1], kk[2], 0.0};
            for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + 0.5 * kk[c];
            a.Vb[i] = d4{vbp[0], vbp[1], vbp[2], 0.0};
            if (ENERGY) a.elast[i] = eli;
        } else if (STAGE == 1) {
            a.k23[i] = d4{kk[0], kk[1], kk[2], 0.0};
            for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + 0.5 * kk[c];
            a.Vb[i] = d4{vbp[0], vbp[1], vbp[2], 0.0};
            if (ENERGY) a.elast[i] = ep.el + eli;
        } else if (STAGE == 2) {
            const double* k2 = ep.v[3];
            a.k23[i] = d4{kk[0] + k2[0], kk[1] + k2[1], kk[2] + k2[2], 0.0};  // k3 += k24 (pdDataManager.cpp:2052)
            for (int c = 0; c < 3; ++c) nup[c] = bcnode ? Up[c] : Up[c] + kk[c];
            a.Vb[i] = d4{vbp[0], vbp[1], vbp[2], 0.0};
            if (ENERGY) a.elast[i] = ep.el + eli;
        } else {
            const double dt6 = 1.0 / 6.0;
            const double* k1p = ep.v[2];
            const double* k23p = ep.v[3];
            const double* Vp = ep.v[4];
            for (int c = 0; c < 3; ++c) {
                if (bcnode) Up[c] = Up[c] + dt6 * (k1p[c] + kk[c] + 2.0 * k23p[c]);
                else Up[c] = Up[c] + dt6 * (k1p[c] + kk[c] + 2.0 * k23p[c]) - dt * sp.damping * Vp[c] * dt / rho;
            }
            for (int c = 0; c < 3; ++c) vbp[c] += -sp.damping * Vp[c] * dt / rho;
            a.V[i] = d4{vbp[0], vbp[1], vbp[2], 0.0};
            for (int c = 0; c < 3; ++c) nup[c] = Up[c];
            if (ENERGY) {
                const double elt = ep.el + eli;
                const double velsq = vbp[0] * vbp[0] + vbp[1] * vbp[1] + vbp[2] * vbp[2];
                ee_i += sp.vol * elt;
                ke_i += 0.5 * rho * velsq * sp.vol;
            }
        }
        if (b >= 0 || STAGE == 3) a.U[i] = d4{Up[0], Up[1], Up[2], 0.0};
        a.uout[i] = d4{nup[0], nup[1], nup[2], 0.0};
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
            double* part = a.part + 2 * (tv.ntl * (int64_t)step_idx + k);  // [step][tile]
            part[0] = ks;
            part[1] = es;
        }
    }
}

// One RK4 stage (STAGE 0..3) for one tile.  Reads a.uin, writes a.uout.
// At most 128 VGPRs (4 waves per SIMD, two 8-wave tiles per CU), or with OCC6 at most 80
// (6 waves per SIMD): the launch picks OCC6 for the FP32 operations on uniform-material
// tiles when the LDS holds more tiles per CU than the 128-VGPR kernel can (three 8-wave
// tiles of edge 4 dx at delta = 3.015 dx, about 51 KB each), so that the staging and
// epilogue of one tile overlap the bond loops of the others.  With room for fewer tiles
// the tighter register budget would only spill.
template <int STAGE, int PREC, bool ENERGY, bool UE, int NW, int NPW, bool OCC6 = false>
__global__ __launch_bounds__(64 * NW, (OCC6 ? 6 : 4)) void k_stage(StageArgs a, int step_idx)
{
    extern __shared__ double L[];
    __shared__ double sh_e[2][NW];
    const TileView& tv = a.tv;
    int64_t k = launch_index(tv.xcd_remap);
    if (k >= (tv.lmap ? tv.nmap : tv.ntl)) return;
    if (tv.lmap) k = tv.lmap[k];
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int H = (int)(tv.halo_ptr[k + 1] - h0);
    const int spec = tile_specimen(tv, k);
    const SpecDev sp = tv.specs[spec];
    constexpr bool F32 = kF32Rec<typename OpFor<PREC>::Force>;
    const Lds l = lds_view(L, H, tv.tile_max, !UE, true, F32);
    // the stage-update inputs (first node of each thread) are loaded together with the halo
    EpiLoad<STAGE, ENERGY> pre;
    const bool has_pre = (int)threadIdx.x < nt;
    if (has_pre) pre.load(a, a0 + threadIdx.x);
    stage_tile<F32>(tv, sp, a.uin, k, h0, H, a0, nt, l, !UE);
    load_tables(a, spec, l);
    if (has_pre) pre.store(l.epi[threadIdx.x]);
    for (int loc = threadIdx.x + 64 * NW; loc < nt; loc += 64 * NW) {
        EpiLoad<STAGE, ENERGY> more;
        more.load(a, a0 + loc);
        more.store(l.epi[loc]);
    }
    __syncthreads();
    tile_families<typename OpFor<PREC>::Force, ENERGY, UE, NW, NPW, OCC6>(tv, sp, l, nt);
    __syncthreads();
    tile_update<STAGE, ENERGY, NW>(a, sp, l, k, a0, nt, step_idx, spec, sh_e);
}

// Per-step, per-specimen energy totals: block b = step * nspec + spec sums the tile
// partials (ke, ee) of that step over the specimen's tiles into out[2b], out[2b+1], in a
// fixed order.
__global__ __launch_bounds__(256) void k_reduce_steps(const double* __restrict__ part, int64_t ntl,
                                                      const int64_t* __restrict__ spec_tile_ptr, int nspec,
                                                      double* __restrict__ out)
{
    __shared__ double sk[256], se[256];
    const int64_t step = blockIdx.x / nspec, spec = blockIdx.x % nspec;
    const double* in = part + 2 * ntl * step;
    const int64_t t0 = spec_tile_ptr[spec], t1 = spec_tile_ptr[spec + 1];
    double ks = 0.0, es = 0.0;
    for (int64_t q = t0 + threadIdx.x; q < t1; q += 256) {
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

// Family sums (force or tangent) for every node, no state update: y = (fx, fy, fz), energy.
// With dot set, also the tile's share of u . y (CG's p . Kp) into dot[2 k] (dot[2 k + 1] = 0),
// k the launch index.
template <class OP, bool UE, int NW, int NPW>
__global__ __launch_bounds__(64 * NW) void k_gather(TileView tv, const d4* __restrict__ u, d4* __restrict__ y,
                                                    double* __restrict__ elast, double* __restrict__ dot)
{
    __shared__ double sh_dot[NW];
    extern __shared__ double L[];
    const int64_t k = launch_index(tv.xcd_remap);
    if (k >= tv.ntl) return;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int H = (int)(tv.halo_ptr[k + 1] - h0);
    constexpr bool F32 = kF32Rec<OP>;
    const Lds l = lds_view(L, H, tv.tile_max, !UE, false, F32);
    const SpecDev sp = tv.specs[tile_specimen(tv, k)];
    stage_tile<F32>(tv, sp, u, k, h0, H, a0, nt, l, !UE);
    __syncthreads();
    tile_families<OP, true, UE, NW, NPW>(tv, sp, l, nt);
    __syncthreads();
    double pq = 0.0;
    for (int loc = threadIdx.x; loc < nt; loc += 64 * NW) {
        const double* res = l.res + 4 * loc;
        y[a0 + loc] = d4{res[0], res[1], res[2], 0.0};
        if (elast) elast[a0 + loc] = res[3];
        if (dot) {
            const d4 v = u[a0 + loc];
            pq += v.x * res[0] + v.y * res[1] + v.z * res[2];
        }
    }
    if (dot) {
        pq = wave_sum(pq);
        if ((threadIdx.x & 63) == 0) sh_dot[threadIdx.x >> 6] = pq;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < NW; ++w) t += sh_dot[w];
            dot[2 * k] = t;
            dot[2 * k + 1] = 0.0;
        }
    }
}

// Tangent weights w = coef / r^3 of every internal entry, r from the FP64 positions, one
// workgroup per tile: tc64 in FP64, tc32 (may be null) rounded to FP32.  Entries with
// coefficient 0 (row padding) get 0.
__global__ __launch_bounds__(256) void k_tangent_coef(TileView tv, double* __restrict__ tc64,
                                                      float* __restrict__ tc32, int16_t* __restrict__ tc16)
{
    const int64_t k = blockIdx.x;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int64_t e0 = tv.rpi[a0], e1 = tv.rpi[a0 + nt];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const double c = tv.coef64[e];
        double w = 0.0;
        if (c != 0.0) {
            // the row of e: the last tile node whose row starts at or before e
            int lo = 0, hi = nt - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (tv.rpi[a0 + mid] <= e) lo = mid;
                else hi = mid - 1;
            }
            const d4 xi = tv.X[a0 + lo];
            const d4 xj = tv.X[tv.halo[h0 + tv.slot[e]]];
            const double x1 = xj.x - xi.x, x2 = xj.y - xi.y, x3 = xj.z - xi.z;
            const double r2 = x1 * x1 + x2 * x2 + x3 * x3;
            w = c / (r2 * sqrt(r2));
        }
        tc64[e] = w;
        if (tc16) {
            // w rounded to nearest at 39 significant bits (an integer add on the magnitude
            // bits), then split into its FP32 truncation and the next 16 significand bits
            // (none for FP32 subnormals, which no weight reaches)
            const double w39 = __longlong_as_double((__double_as_longlong(w) + 0x1000LL) & ~0x1FFFLL);
            const float wt = __double2float_rz(w39);
            const bool normal = fabsf(wt) >= 0x1p-126f;
            tc32[e] = wt;
            tc16[e] = normal ? (int16_t)(uint16_t)((unsigned)__double2loint(w39) >> 13) : (int16_t)0;
        } else if (tc32) {
            tc32[e] = (float)w;
        }
    }
}

// Correction bytes of the FP32 force operations: for each entry, q = rint((coef64 / base -
// fly_beta) / 2^-26) clamped to 0..255, with base and fly_beta evaluated by the same
// functions, on the same FP32 records, as the force kernels.  One workgroup per
// tile; *clamped counts entries whose correction fell outside the byte.
__global__ __launch_bounds__(256) void k_fly_corr(TileView tv, const d4* __restrict__ u, uint8_t* __restrict__ corr,
                                                  int* __restrict__ clamped)
{
    extern __shared__ double L[];
    const int64_t k = blockIdx.x;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int H = (int)(tv.halo_ptr[k + 1] - h0);
    const SpecDev sp = tv.specs[tile_specimen(tv, k)];
    const Lds l = lds_view(L, H, tv.tile_max, false, false, true);
    stage_tile<true>(tv, sp, u, k, h0, H, a0, nt, l, false);
    __syncthreads();
    int bad = 0;
    const bool ue = tv.uniform_mat != 0;
    for (int loc = 0; loc < nt; ++loc) {
        const SelfF me = SelfF::make(l, sp, l.self[loc], true);
        for (int64_t e = l.row[loc] + threadIdx.x; e < l.row[loc + 1]; e += blockDim.x) {
            const RecF& rec = l.recf[tv.slot[e]];
            float r2, rinv, rr;
            rec_r(rec.a, me, r2, rinv, rr);
            const float base = ue ? fly_base<true>(me, rec.w.y) : fly_base<false>(me, rec.w.y);
            const double scale = ue ? sp.kc0 : 1.0;
            const float b32 = fly_beta(rr, me);
            int q = 128;
            if (base > 0.0f) {
                const double d = rint((tv.coef64[e] / ((double)base * scale) - (double)b32) / (double)kCorrStep);
                bad += (d < 0.0 || d > 255.0) ? 1 : 0;
                q = (int)fmin(fmax(d, 0.0), 255.0);
            }
            corr[e] = (uint8_t)q;
        }
    }
    if (bad) atomicAdd(clamped, bad);
}

// ------------------------------------------------------------ fixed Gaussian integration
// The fixed-Gaussian branch of CalcBondBasedMotionRK4's family loop (pdDataManager.cpp:
// 1918-1957): for each unbroken entry (i, j), the CCI bond force (pdDataManager2.cpp:
// 1293-1351, volume outside) at each Gauss point g of node j, with xi = X_g - X_i, eta =
// u_g - u_i, beta over the cell size (w1 w2 w3)^(1/3) dx / 2, summed with the weights
// dx^3/8 w1 w2 w3.  coef64 holds bdryEff_ij (0: cut entry or row padding).  One wavefront per
// node; its lanes step through the row's live (entry, point) pairs (f.pairs: uncut entries,
// beta > 0), so that consecutive lanes read consecutive Gauss points; a pair left out adds
// exactly 0 to every sum.  Returns (fx, fy, fz, energy density) in every lane.
// One live pair's inputs: the entry's coefficient, node j's position and the Gauss point's
// displacement, and the pair's ecrit and spring constant.
struct FgiPairIn {
    double bij, ecrit, spring;
    d4 xj, ug;
    int g;
};

__device__ __forceinline__ FgiPairIn fgi_pair_in(const TileView& tv, const FgiArgs& f, const MatDev* mats,
                                                 const int32_t* __restrict__ mat, const MatDev& mi, int64_t eb,
                                                 int64_t idx)
{
    FgiPairIn q;
    const uint2 pr = f.pairs[idx];
    const int32_t j = (int32_t)pr.y;
    q.g = (int)(pr.x & 0xFFFFu);
    q.bij = tv.coef64[eb + (int64_t)(pr.x >> 16)];
    q.xj = tv.X[j];
    q.ug = f.gin[(int64_t)j * f.ng + q.g];
    if (tv.uniform_mat) {
        q.ecrit = mi.ecrit;
        q.spring = mi.spring;
    } else {
        const MatDev& mj = mats[mat[j]];
        q.ecrit = (mj.ecrit < mi.ecrit) ? mj.ecrit : mi.ecrit;
        q.spring = (mj.spring < mi.spring) ? mj.spring : mi.spring;
    }
    return q;
}

// The FP64 bond force of one pair (CalcBondForceCCI's expression, volume outside) times the
// point's weight, added to the lane's sums.
__device__ __forceinline__ void fgi_pair_add(const FgiArgs& f, const double* __restrict__ tab, const d4& xi,
                                             const d4& ui, const FgiPairIn& q, double& f0, double& f1, double& f2,
                                             double& el)
{
    const double* t = tab + kFgiTab * q.g;
    const double xi1 = (t[0] + q.xj.x) - xi.x, xi2 = (t[1] + q.xj.y) - xi.y, xi3 = (t[2] + q.xj.z) - xi.z;
    const double a1 = xi1 + (q.ug.x - ui.x), a2 = xi2 + (q.ug.y - ui.y), a3 = xi3 + (q.ug.z - ui.z);
    const double r = sqrt_rn(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
    const double p = sqrt_rn(a1 * a1 + a2 * a2 + a3 * a3);
    const double beta = g_beta(r, f.horizon, t[4], false);
    const double u = p - r;
    // one reciprocal for the divisions by r and p: 1/r = p / (r p), 1/p = r / (r p)
    const double inv = 1.0 / (r * p);
    const double rinv = p * inv;
    const double st = u * rinv;
    // u <= 0 || (u > 0 && st <= ecrit), without branches
    const bool ok = (u <= 0.0) | (st <= q.ecrit);
    const double c = ok ? q.bij * beta * q.spring * st : 0.0;
    const double w = t[3] * (c * (r * inv));
    f0 = __builtin_fma(w, a1, f0);
    f1 = __builtin_fma(w, a2, f1);
    f2 = __builtin_fma(w, a3, f2);
    el += ok ? q.bij * 0.5 * (0.5 * beta * (q.spring * rinv) * u * u) * t[5] : 0.0;
}

__device__ __forceinline__ void fgi_family(const TileView& tv, const FgiArgs& f, const double* __restrict__ tab,
                                           const MatDev* mats, const int32_t* __restrict__ mat, int64_t h0, int64_t i,
                                           const d4& ui, double acc[4])
{
    (void)h0;
    const int lane = threadIdx.x & 63;
    const d4 xi = tv.X[i];
    const MatDev& mi = mats[mat[i]];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, el = 0.0;
    const int64_t eb = tv.rpi[i];
    const int64_t p1 = f.pair_ptr[i + 1];
    // two pairs per lane per pass (idx and idx + 64: the lane's order of summation is unchanged),
    // so that each lane has two chains of dependent loads in flight
    int64_t idx = f.pair_ptr[i] + lane;
    for (; idx + 64 < p1; idx += 128) {
        const FgiPairIn qa = fgi_pair_in(tv, f, mats, mat, mi, eb, idx);
        const FgiPairIn qb = fgi_pair_in(tv, f, mats, mat, mi, eb, idx + 64);
        fgi_pair_add(f, tab, xi, ui, qa, f0, f1, f2, el);
        fgi_pair_add(f, tab, xi, ui, qb, f0, f1, f2, el);
    }
    if (idx < p1) fgi_pair_add(f, tab, xi, ui, fgi_pair_in(tv, f, mats, mat, mi, eb, idx), f0, f1, f2, el);
    acc[0] = wave_sum(f0);
    acc[1] = wave_sum(f1);
    acc[2] = wave_sum(f2);
    acc[3] = wave_sum(el);
}

// fgi_family in FP32 bond arithmetic (PDM_FP32C, PDM_FP32C2): each pair's xi = (X_j + offset)
// - X_i and eta = u_g - u_i are formed in FP64 and rounded once to FP32; the stretch is formed
// without cancellation, s = eta . (2 xi + eta) / ((p + r) r), with r, p from rsq; beta over the
// point's cell size and the ecrit test in FP32.  Sums per lane in FP32 (TF: two-float TwoSum
// pairs), then across the wavefront in FP64.  The Gauss-point displacements and the RK4 state
// stay FP64.
__device__ __forceinline__ float g_beta32(float dist, float delta, float dx)
{
    const float inner = delta - 0.5f * dx;
    const float outer = delta + 0.5f * dx;
    if (dist <= inner) return 1.0f;
    if (dist <= outer) return (outer - dist) / dx;
    return 0.0f;
}

template <bool TF>
__device__ __forceinline__ void fgi_pair_add32(const double* __restrict__ tab, float delta, const d4& xi,
                                               const d4& ui, const FgiPairIn& q, float& f0, float& f1, float& f2,
                                               float& l0, float& l1, float& l2, float& el)
{
    const double* t = tab + kFgiTab * q.g;
    const float bij = (float)q.bij, ecrit = (float)q.ecrit, spring = (float)q.spring;
    const float x1 = (float)((t[0] + q.xj.x) - xi.x), x2 = (float)((t[1] + q.xj.y) - xi.y),
                x3 = (float)((t[2] + q.xj.z) - xi.z);
    const float e1 = (float)(q.ug.x - ui.x), e2 = (float)(q.ug.y - ui.y), e3 = (float)(q.ug.z - ui.z);
    const float a1 = x1 + e1, a2 = x2 + e2, a3 = x3 + e3;
    const float r2 = x1 * x1 + x2 * x2 + x3 * x3;
    const float p2 = a1 * a1 + a2 * a2 + a3 * a3;
    const float rinv = __builtin_amdgcn_rsqf(r2), pinv = __builtin_amdgcn_rsqf(p2);
    const float r = r2 * rinv, p = p2 * pinv;
    const float num = e1 * (2.0f * x1 + e1) + e2 * (2.0f * x2 + e2) + e3 * (2.0f * x3 + e3);
    const float st = num * __builtin_amdgcn_rcpf((p + r) * r);
    const float beta = g_beta32(r, delta, (float)t[4]);
    const bool ok = (num <= 0.0f) | (st <= ecrit);
    const float c = ok ? (float)t[3] * (bij * beta * spring * st) * pinv : 0.0f;
    if (TF) {
        float sum, err;
        two_sum(f0, c * a1, sum, err); f0 = sum; l0 += err;
        two_sum(f1, c * a2, sum, err); f1 = sum; l1 += err;
        two_sum(f2, c * a3, sum, err); f2 = sum; l2 += err;
    } else {
        f0 += c * a1;
        f1 += c * a2;
        f2 += c * a3;
    }
    const float u = st * r;
    el += ok ? bij * 0.5f * (0.5f * beta * (spring * rinv) * u * u) * (float)t[5] : 0.0f;
}

template <bool TF>
__device__ __forceinline__ void fgi_family32(const TileView& tv, const FgiArgs& f, const double* __restrict__ tab,
                                             const MatDev* mats, const int32_t* __restrict__ mat, int64_t h0, int64_t i,
                                             const d4& ui, double acc[4])
{
    (void)h0;
    const int lane = threadIdx.x & 63;
    const d4 xi = tv.X[i];
    const MatDev& mi = mats[mat[i]];
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, l0 = 0.0f, l1 = 0.0f, l2 = 0.0f, el = 0.0f;
    const float delta = (float)f.horizon;
    const int64_t eb = tv.rpi[i];
    const int64_t p1 = f.pair_ptr[i + 1];
    int64_t idx = f.pair_ptr[i] + lane;
    for (; idx + 64 < p1; idx += 128) {  // two pairs per lane per pass, as fgi_family
        const FgiPairIn qa = fgi_pair_in(tv, f, mats, mat, mi, eb, idx);
        const FgiPairIn qb = fgi_pair_in(tv, f, mats, mat, mi, eb, idx + 64);
        fgi_pair_add32<TF>(tab, delta, xi, ui, qa, f0, f1, f2, l0, l1, l2, el);
        fgi_pair_add32<TF>(tab, delta, xi, ui, qb, f0, f1, f2, l0, l1, l2, el);
    }
    if (idx < p1)
        fgi_pair_add32<TF>(tab, delta, xi, ui, fgi_pair_in(tv, f, mats, mat, mi, eb, idx), f0, f1, f2, l0, l1, l2,
                           el);
    acc[0] = wave_sum((double)f0 + (double)l0);
    acc[1] = wave_sum((double)f1 + (double)l1);
    acc[2] = wave_sum((double)f2 + (double)l2);
    acc[3] = wave_sum((double)el);
}

// The adaptive-integration branch of CalcBondBasedMotionRK4's family loop (pdDataManager.cpp:
// 1888-1916) for one node i per workgroup: for each uncut entry (i, j), CalcBondBasedForce
// (pdIntgManager2.cpp:92-310) at every trapezoidal point of node j's cell clipped to node i's
// horizon, beta = 1, summed with the points' weights.  The point's relative displacement:
// - f.ai == 2 (no MLS): G_j xi from node j's DISPLGRAD condition;
// - a BC node j (mode 0): its DISPL / DISPLGRAD value at the step's time in the flagged
//   components, 0 in the others (pdIntgManager2.cpp:115-172), less u_i;
// - otherwise the MLS interpolation sum_k phi_k src[cb_k] over node j's contributing nodes
//   (pdIntgManager2.cpp:173-186), src holding the stage's trial displacements, less u_i.
// A point's elastic energy density is scaled by its share of a cell, w / dx^3.  The threads step
// through the node's points; the points of an entry are consecutive threads sharing node j, and
// their coefficients are stored [contributing node][point], so a wavefront reads them in
// contiguous runs.  out[i] = (fx, fy, fz, energy density), sums in a fixed order.
constexpr int kAiThreads = 256;

__global__ __launch_bounds__(kAiThreads) void k_ai_family(TileView tv, FgiArgs f, const int32_t* __restrict__ mat,
                                                          const int32_t* __restrict__ bc, const MatDev* __restrict__ mats,
                                                          const BCDev* __restrict__ bcs, const d4* __restrict__ uin,
                                                          const d4* __restrict__ src, int mode,
                                                          const StepParam* __restrict__ steps, int step_idx,
                                                          d4* __restrict__ out)
{
    __shared__ double sh[4][kAiThreads / 64];
    const int64_t i = blockIdx.x;
    const double t = (mode == 0) ? steps[step_idx].t : 0.0;
    const d4 ui = uin[i];
    const d4 xi = tv.X[i];
    const MatDev mi = mats[mat[i]];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, el = 0.0;
    const int64_t eb = tv.rpi[i];
    const int64_t q1 = f.ai_ptr[i + 1];
    for (int64_t q = f.ai_ptr[i] + threadIdx.x; q < q1; q += kAiThreads) {
        const d4 w = f.ai_xw[q];
        const int64_t e = eb + (int64_t)f.ai_eo[q];
        const double bij = tv.coef64[e];
        const int32_t j = f.ai_j[q];
        const MatDev& mj = mats[mat[j]];
        const double ecrit = (mj.ecrit < mi.ecrit) ? mj.ecrit : mi.ecrit;
        const double spring = (mj.spring < mi.spring) ? mj.spring : mi.spring;
        const int b = bc[j];
        double e1, e2, e3;
        if (f.ai == 2) {
            const double* G = bcs[b].v;
            e1 = G[0] * w.x + G[1] * w.y + G[2] * w.z;
            e2 = G[3] * w.x + G[4] * w.y + G[5] * w.z;
            e3 = G[6] * w.x + G[7] * w.y + G[8] * w.z;
        } else {
            double v[3] = {0.0, 0.0, 0.0};
            if (mode == 0 && b >= 0 && bcs[b].bcnode) {
                const BCDev& bd = bcs[b];
                const double fac = time_factor_dev(bd, t);
                const double xg[3] = {w.x + xi.x, w.y + xi.y, w.z + xi.z};
                for (int c = 0; c < 3; ++c) {
                    if (bd.f[c] != 1) continue;
                    if (bd.type == 1) v[c] = bd.v[c] * fac;
                    else if (bd.type == 3)
                        v[c] = (bd.v[3 * c] * xg[0] + bd.v[3 * c + 1] * xg[1] + bd.v[3 * c + 2] * xg[2]) * fac;
                }
            } else {
                const int64_t k0 = f.cb_ptr[j], k1 = f.cb_ptr[j + 1];
                const int64_t st = f.ai_ps[q];
                const double* ph = f.ai_phi + f.ai_po[q];
#pragma unroll 4
                for (int64_t k = k0; k < k1; ++k) {
                    const double c = *ph;
                    ph += st;
                    const d4 x = src[f.cb[k]];
                    v[0] += x.x * c;
                    v[1] += x.y * c;
                    v[2] += x.z * c;
                }
            }
            e1 = v[0] - ui.x;
            e2 = v[1] - ui.y;
            e3 = v[2] - ui.z;
        }
        const double a1 = w.x + e1, a2 = w.y + e2, a3 = w.z + e3;
        const double r = sqrt_rn(w.x * w.x + w.y * w.y + w.z * w.z);
        const double p = sqrt_rn(a1 * a1 + a2 * a2 + a3 * a3);
        const double u = p - r;
        // one reciprocal for the divisions by r and p: 1/r = p / (r p), 1/p = r / (r p)
        const double inv = 1.0 / (r * p);
        const double rinv = p * inv;
        const double sr = u * rinv;
        const bool ok = (u <= 0.0) | (sr <= ecrit);
        const double c = ok ? bij * spring * sr : 0.0;
        const double wf = w.w * (c * (r * inv));
        f0 = __builtin_fma(wf, a1, f0);
        f1 = __builtin_fma(wf, a2, f1);
        f2 = __builtin_fma(wf, a3, f2);
        el += ok ? bij * 0.5 * (0.5 * (spring * rinv) * u * u) * (w.w / f.ai_vol) : 0.0;
    }
    const double r0 = wave_sum(f0), r1 = wave_sum(f1), r2 = wave_sum(f2), r3 = wave_sum(el);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][wave] = r0;
        sh[1][wave] = r1;
        sh[2][wave] = r2;
        sh[3][wave] = r3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < 4; ++c)
            for (int q = 0; q < kAiThreads / 64; ++q) s4[c] += sh[c][q];
        out[i] = d4{s4[0], s4[1], s4[2], s4[3]};
    }
}

constexpr int kFgiWaves = 4;

// Gauss-point displacements, one thread per (node, point): CalcIntgPointDisplRK4
// (pdDataManager.cpp:2402-2495).  A BC node's points take the prescribed DISPL / DISPLGRAD
// components at the step's time (UpdateIntgPointStatusWithBC, :2604-2703) and keep the others;
// the other nodes' points interpolate u with the MLS coefficients (n > 1) or copy it (n = 1).
// f.mode 1 updates the non-BC nodes only, mode 2 interpolates every node.
__global__ __launch_bounds__(256) void k_fgi_gauss(FgiArgs f, const d4* __restrict__ u, int step_idx)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= f.n * f.ng) return;
    const int64_t a = q / f.ng;
    const int g = (int)(q - a * f.ng);
    const int b = f.bc[a];
    if (f.mode != 2 && b >= 0 && f.bcs[b].bcnode) {
        if (f.mode == 1) return;
        const BCDev& bd = f.bcs[b];
        const double fac = time_factor_dev(bd, f.steps[step_idx].t);
        const double* t = f.tab + kFgiTab * g;
        const d4 x = f.X[a];
        const double xg[3] = {t[0] + x.x, t[1] + x.y, t[2] + x.z};
        d4 G = f.gout[q];
        double v[3] = {G.x, G.y, G.z};
        for (int c = 0; c < 3; ++c) {
            if (bd.f[c] != 1) continue;
            if (bd.type == 1) v[c] = bd.v[c] * fac;
            else if (bd.type == 3) v[c] = (bd.v[3 * c] * xg[0] + bd.v[3 * c + 1] * xg[1] + bd.v[3 * c + 2] * xg[2]) * fac;
        }
        f.gout[q] = d4{v[0], v[1], v[2], 0.0};
        return;
    }
    if (!f.phi) {
        const d4 x = u[a];
        f.gout[q] = d4{x.x, x.y, x.z, 0.0};
        return;
    }
    const int64_t k0 = f.cb_ptr[a], k1 = f.cb_ptr[a + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    // unrolled so that several iterations' loads are in flight; the sums keep their order
#pragma unroll 4
    for (int64_t k = k0; k < k1; ++k) {
        const double c = f.phi[k * f.ng + g];
        const d4 x = u[f.cb[k]];
        s0 += x.x * c;
        s1 += x.y * c;
        s2 += x.z * c;
    }
    f.gout[q] = d4{s0, s1, s2, 0.0};
}

// LDS of the FGI kernels: the stage layout without halo records, then the point table.
__host__ __device__ inline size_t fgi_lds_bytes(int tile_max, int ng)
{
    return lds_bytes(0, tile_max, false, true, false) + sizeof(double) * kFgiTab * (size_t)ng;
}

// One RK4 stage of a fixed-Gaussian specimen for one tile: family sums over the neighbours'
// Gauss points (a.uin: this stage's displacement of the family owner, f.gin: the Gauss
// points'), then the same stage update as k_stage.
// FAM: 0 fixed Gaussian FP64, 1 adaptive integration (sums from k_ai_family), 2 fixed Gaussian
// in FP32 bond arithmetic, 3 the same with two-float sums.
template <int STAGE, bool ENERGY, int FAM>
__global__ __launch_bounds__(64 * kFgiWaves) void k_fgi_stage(StageArgs a, FgiArgs f, int step_idx)
{
    extern __shared__ double L[];
    __shared__ double sh_e[2][kFgiWaves];
    const TileView& tv = a.tv;
    const int64_t k = blockIdx.x;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const SpecDev sp = tv.specs[0];
    const Lds l = lds_view(L, 0, tv.tile_max, false, true, false);
    double* tab = reinterpret_cast<double*>(reinterpret_cast<char*>(L) + lds_bytes(0, tv.tile_max, false, true, false));
    for (int q = threadIdx.x; q < kFgiTab * f.ng; q += blockDim.x) tab[q] = f.tab[q];
    load_tables(a, 0, l);
    for (int loc = threadIdx.x; loc < nt; loc += blockDim.x) {
        EpiLoad<STAGE, ENERGY> in;
        in.load(a, a0 + loc);
        in.store(l.epi[loc]);
    }
    __syncthreads();
    for (int loc = threadIdx.x >> 6; loc < nt; loc += kFgiWaves) {
        const int64_t i = a0 + loc;
        double acc[4];
        if (FAM == 1) {  // the family sums of k_ai_family
            const d4 r = f.ai_res[i];
            acc[0] = r.x;
            acc[1] = r.y;
            acc[2] = r.z;
            acc[3] = r.w;
        } else if (FAM == 2) {
            fgi_family32<false>(tv, f, tab, l.mats, a.mat, h0, i, a.uin[i], acc);
        } else if (FAM == 3) {
            fgi_family32<true>(tv, f, tab, l.mats, a.mat, h0, i, a.uin[i], acc);
        } else {
            fgi_family(tv, f, tab, l.mats, a.mat, h0, i, a.uin[i], acc);
        }
        if ((threadIdx.x & 63) == 0)
            for (int c = 0; c < 4; ++c) l.res[4 * loc + c] = acc[c];
    }
    __syncthreads();
    tile_update<STAGE, ENERGY, kFgiWaves>(a, sp, l, k, a0, nt, step_idx, 0, sh_e);
    if (STAGE == 3) return;  // the Gauss points then follow the new U (a.uout)
    // The Gauss points interpolate the RK4 trial displacement displStep = u1 + c k of every
    // node (pdDataManager.cpp:2040-2056), where the force loop reads a BC node's own U
    // (a.uout): a BC node's trial is its step-start U plus c k.
    __syncthreads();
    const double dts = (STAGE < 2) ? 0.5 * a.steps[step_idx].dt : a.steps[step_idx].dt;
    const double c = (STAGE < 2) ? 0.5 : 1.0;
    // two trial buffers: the next stage reads this one while its own output goes to the other
    d4* trial = f.trial + (STAGE & 1) * f.n;
    for (int loc = threadIdx.x; loc < nt; loc += blockDim.x) {
        const int64_t i = a0 + loc;
        const EpiRec& ep = l.epi[loc];
        if (ep.b < 0 || !l.bcs[ep.b].bcnode) {
            trial[i] = a.uout[i];
            continue;
        }
        d4 u1;
        if (STAGE == 0) {
            u1 = d4{ep.v[1][0], ep.v[1][1], ep.v[1][2], 0.0};
            f.u0[i] = u1;
        } else {
            u1 = f.u0[i];
        }
        const d4 vb = a.Vb[i];
        trial[i] = d4{u1.x + c * (vb.x * dts), u1.y + c * (vb.y * dts), u1.z + c * (vb.z * dts), 0.0};
    }
}

// Fixed-Gaussian family sums of every node for the displacement u (f.gin interpolated from u);
// FAM as for k_fgi_stage (0, 2, 3).
template <int FAM>
__global__ __launch_bounds__(64 * kFgiWaves) void k_fgi_forces(StageArgs a, FgiArgs f, const d4* __restrict__ u,
                                                               d4* __restrict__ F, double* __restrict__ elast)
{
    extern __shared__ double L[];
    const TileView& tv = a.tv;
    const int64_t k = blockIdx.x;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const Lds l = lds_view(L, 0, tv.tile_max, false, true, false);
    double* tab = reinterpret_cast<double*>(reinterpret_cast<char*>(L) + lds_bytes(0, tv.tile_max, false, true, false));
    for (int q = threadIdx.x; q < kFgiTab * f.ng; q += blockDim.x) tab[q] = f.tab[q];
    load_tables(a, 0, l);
    __syncthreads();
    for (int loc = threadIdx.x >> 6; loc < nt; loc += kFgiWaves) {
        const int64_t i = a0 + loc;
        double acc[4];
        if (FAM == 2) fgi_family32<false>(tv, f, tab, l.mats, a.mat, h0, i, u[i], acc);
        else if (FAM == 3) fgi_family32<true>(tv, f, tab, l.mats, a.mat, h0, i, u[i], acc);
        else fgi_family(tv, f, tab, l.mats, a.mat, h0, i, u[i], acc);
        if ((threadIdx.x & 63) == 0) {
            F[i] = d4{acc[0], acc[1], acc[2], 0.0};
            if (elast) elast[i] = acc[3];
        }
    }
}

// Live (entry, Gauss point) pairs of each node's row, for fgi_family: an uncut entry
// (coefficient != 0) and a point whose beta is nonzero, with r formed exactly as fgi_family
// forms it.  One workgroup per tile, one wavefront per node, one lane per entry; the pairs
// are written in (entry, point) order, each with node j's internal id, so that the gathers
// reach node j's position and Gauss points in one dependent load instead of three (slot,
// halo, node).
__device__ __forceinline__ bool fgi_point_live(const FgiArgs& f, int g, const d4& xi, const d4& xj)
{
    const double* t = f.tab + kFgiTab * g;
    const double xi1 = (t[0] + xj.x) - xi.x, xi2 = (t[1] + xj.y) - xi.y, xi3 = (t[2] + xj.z) - xi.z;
    const double r = sqrt(xi1 * xi1 + xi2 * xi2 + xi3 * xi3);
    return g_beta(r, f.horizon, t[4], false) != 0.0;
}

template <bool FILL>
__global__ __launch_bounds__(64 * kFgiWaves) void k_fgi_pairs(TileView tv, FgiArgs f, int64_t* __restrict__ cnt,
                                                              int64_t* __restrict__ ent, uint2* __restrict__ pairs)
{
    const int64_t k = blockIdx.x;
    const int32_t tid = tv.tiles[k];
    const int64_t a0 = tv.tile_start[tid];
    const int nt = (int)(tv.tile_start[tid + 1] - a0);
    const int64_t h0 = tv.halo_ptr[k];
    const int lane = threadIdx.x & 63;
    const int ng = f.ng;
    for (int loc = threadIdx.x >> 6; loc < nt; loc += kFgiWaves) {
        const int64_t i = a0 + loc;
        const d4 xi = tv.X[i];
        const int64_t eb = tv.rpi[i], ee = tv.rpi[i + 1];
        int64_t base = FILL ? f.pair_ptr[i] : 0;
        int64_t rc = 0, re = 0;
        for (int64_t e0 = eb; e0 < ee; e0 += 64) {
            const int64_t e = e0 + lane;
            const bool uncut = e < ee && tv.coef64[e] != 0.0;
            d4 xj = xi;
            int32_t j = 0;
            if (uncut) {  // cut and padding entries hold the padding slot
                j = tv.halo[h0 + tv.slot[e]];
                xj = tv.X[j];
            }
            int c = 0;
            if (uncut)
                for (int g = 0; g < ng; ++g) c += fgi_point_live(f, g, xi, xj) ? 1 : 0;
            // exclusive prefix of the lanes' counts
            int incl = c;
            for (int off = 1; off < 64; off <<= 1) {
                const int v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            const int total = __shfl(incl, 63, 64);
            if (FILL && c > 0) {
                int64_t o = base + (incl - c);
                const uint32_t eo = (uint32_t)(e - eb) << 16;
                for (int g = 0; g < ng; ++g)
                    if (fgi_point_live(f, g, xi, xj)) pairs[o++] = make_uint2(eo | (uint32_t)g, (uint32_t)j);
            }
            base += total;
            rc += total;
            re += __popcll(__ballot(c > 0));
        }
        if (!FILL && lane == 0) {
            cnt[i] = rc;
            ent[i] = re;
        }
    }
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

template <int STAGE, int PREC, bool ENERGY, bool UE, int NW, int NPW, bool OCC6 = false>
hipError_t launch_stage_t(const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    hipError_t e = set_lds<k_stage<STAGE, PREC, ENERGY, UE, NW, NPW, OCC6>>(lds);
    if (e != hipSuccess) return e;
    k_stage<STAGE, PREC, ENERGY, UE, NW, NPW, OCC6>
        <<<dim3((unsigned)padded_grid(a.tv.lmap ? a.tv.nmap : a.tv.ntl)), dim3(64 * NW), lds, s>>>(a, step_idx);
    return hipGetLastError();
}

template <int STAGE, int PREC, int NW, int NPW>
hipError_t launch_stage_e(const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    const bool occ6 = stage_vgpr_cap(a.tv, PREC, lds) == 80;
    if (a.energies) {
        if (a.tv.uniform_mat) {
            if constexpr (PREC != PDM_FP64)
                if (occ6) return launch_stage_t<STAGE, PREC, true, true, NW, NPW, true>(a, step_idx, lds, s);
            return launch_stage_t<STAGE, PREC, true, true, NW, NPW>(a, step_idx, lds, s);
        }
        return launch_stage_t<STAGE, PREC, true, false, NW, NPW>(a, step_idx, lds, s);
    }
    if (a.tv.uniform_mat) {
        if constexpr (PREC != PDM_FP64)
            if (occ6) return launch_stage_t<STAGE, PREC, false, true, NW, NPW, true>(a, step_idx, lds, s);
        return launch_stage_t<STAGE, PREC, false, true, NW, NPW>(a, step_idx, lds, s);
    }
    return launch_stage_t<STAGE, PREC, false, false, NW, NPW>(a, step_idx, lds, s);
}

template <int PREC, int NW, int NPW>
hipError_t launch_stage_s(int stage, const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    switch (stage) {
    case 0: return launch_stage_e<0, PREC, NW, NPW>(a, step_idx, lds, s);
    case 1: return launch_stage_e<1, PREC, NW, NPW>(a, step_idx, lds, s);
    case 2: return launch_stage_e<2, PREC, NW, NPW>(a, step_idx, lds, s);
    default: return launch_stage_e<3, PREC, NW, NPW>(a, step_idx, lds, s);
    }
}

template <int PREC>
hipError_t launch_stage_p(int stage, const StageArgs& a, int step_idx, size_t lds, hipStream_t s)
{
    const int w = a.tv.waves, f = a.tv.fam_per_wave;
    if (w == 4) {
        if (f == 1) return launch_stage_s<PREC, 4, 1>(stage, a, step_idx, lds, s);
        if (f == 2) return launch_stage_s<PREC, 4, 2>(stage, a, step_idx, lds, s);
        return launch_stage_s<PREC, 4, 4>(stage, a, step_idx, lds, s);
    }
    if (f == 1) return launch_stage_s<PREC, 8, 1>(stage, a, step_idx, lds, s);
    if (f == 2) return launch_stage_s<PREC, 8, 2>(stage, a, step_idx, lds, s);
    return launch_stage_s<PREC, 8, 4>(stage, a, step_idx, lds, s);
}

template <class OP, bool UE>
hipError_t launch_gather_t(const TileView& tv, const d4* u, d4* y, double* elast, size_t lds, hipStream_t s,
                           double* dot = nullptr)
{
    constexpr int NW = 8, NPW = 4;
    hipError_t e = set_lds<k_gather<OP, UE, NW, NPW>>(lds);
    if (e != hipSuccess) return e;
    k_gather<OP, UE, NW, NPW><<<dim3((unsigned)padded_grid(tv.ntl)), dim3(64 * NW), lds, s>>>(tv, u, y, elast, dot);
    return hipGetLastError();
}

template <class OP>
hipError_t launch_gather_u(const TileView& tv, const d4* u, d4* y, double* elast, size_t lds, hipStream_t s,
                           double* dot = nullptr)
{
    if (tv.uniform_mat) return launch_gather_t<OP, true>(tv, u, y, elast, lds, s, dot);
    return launch_gather_t<OP, false>(tv, u, y, elast, lds, s, dot);
}

}  // namespace

// Register budget of the stage kernel a launch takes.  Tiles per CU: 16 / NW waves at 4
// per SIMD, 24 / NW at 6, LDS (dynamic + the static energy partials) permitting; the
// 80-VGPR kernel (6 waves per SIMD) exists for the FP32 operations on uniform-material
// tiles and is taken when the LDS holds more tiles per CU than the 128-VGPR one can use.
int stage_vgpr_cap(const TileView& tv, int precision, size_t lds)
{
    const int nw = tv.waves == 4 ? 4 : 8;
    const size_t per_tile = lds + sizeof(double) * 2 * nw;
    const bool occ6 = precision != PDM_FP64 && tv.uniform_mat && kLdsBytes / per_tile > (size_t)(16 / nw);
    return occ6 ? 80 : 128;
}

size_t stage_lds_bytes(const TileView& tv, bool with_ecrit, bool with_epi, bool f32)
{
    return lds_bytes(tv.halo_max, tv.tile_max, with_ecrit, with_epi, f32);
}


hipError_t launch_stage(int stage, int precision, const StageArgs& a, int step_idx, hipStream_t s)
{
    if (a.tv.ntl == 0 || (a.tv.lmap && a.tv.nmap == 0)) return hipSuccess;
    size_t lds = stage_lds_bytes(a.tv, !a.tv.uniform_mat, true, precision != PDM_FP64);
    switch (precision) {
    case PDM_FP64: return launch_stage_p<PDM_FP64>(stage, a, step_idx, lds, s);
    case PDM_FP32C: return launch_stage_p<PDM_FP32C>(stage, a, step_idx, lds, s);
    default: return launch_stage_p<PDM_FP32C2>(stage, a, step_idx, lds, s);
    }
}

hipError_t launch_energy_steps(const double* part, int64_t ntl, const int64_t* spec_tile_ptr, int nspec, int nsteps,
                               double* out, hipStream_t s)
{
    if (nsteps <= 0 || nspec <= 0) return hipSuccess;
    k_reduce_steps<<<dim3((unsigned)(nsteps * nspec)), dim3(256), 0, s>>>(part, ntl, spec_tile_ptr, nspec, out);
    return hipGetLastError();
}

hipError_t launch_forces(int precision, const TileView& tv, const d4* u, d4* F, double* elast, hipStream_t s)
{
    if (tv.ntl == 0) return hipSuccess;
    size_t lds = stage_lds_bytes(tv, !tv.uniform_mat, false, precision != PDM_FP64);
    switch (precision) {
    case PDM_FP64: return launch_gather_u<OpF64>(tv, u, F, elast, lds, s);
    case PDM_FP32C: return launch_gather_u<OpF32>(tv, u, F, elast, lds, s);
    default: return launch_gather_u<OpF32TF>(tv, u, F, elast, lds, s);
    }
}

template <int STAGE, bool ENERGY, int FAM>
hipError_t launch_fgi_stage_e(const StageArgs& a, const FgiArgs& f, int step_idx, size_t lds, hipStream_t s)
{
    hipError_t e = set_lds<k_fgi_stage<STAGE, ENERGY, FAM>>(lds);
    if (e != hipSuccess) return e;
    k_fgi_stage<STAGE, ENERGY, FAM><<<dim3((unsigned)a.tv.ntl), dim3(64 * kFgiWaves), lds, s>>>(a, f, step_idx);
    return hipGetLastError();
}

// The family kind of an FgiArgs: adaptive integration, or fixed Gaussian in its precision.
inline int fgi_family_kind(const FgiArgs& f)
{
    if (f.ai) return 1;
    return f.prec == PDM_FP32C ? 2 : f.prec == PDM_FP32C2 ? 3 : 0;
}

template <int STAGE, bool ENERGY>
hipError_t launch_fgi_stage_k(const StageArgs& a, const FgiArgs& f, int step_idx, size_t lds, hipStream_t s)
{
    switch (fgi_family_kind(f)) {
    case 1: return launch_fgi_stage_e<STAGE, ENERGY, 1>(a, f, step_idx, lds, s);
    case 2: return launch_fgi_stage_e<STAGE, ENERGY, 2>(a, f, step_idx, lds, s);
    case 3: return launch_fgi_stage_e<STAGE, ENERGY, 3>(a, f, step_idx, lds, s);
    default: return launch_fgi_stage_e<STAGE, ENERGY, 0>(a, f, step_idx, lds, s);
    }
}

template <int STAGE>
hipError_t launch_fgi_stage_t(const StageArgs& a, const FgiArgs& f, int step_idx, size_t lds, hipStream_t s)
{
    return a.energies ? launch_fgi_stage_k<STAGE, true>(a, f, step_idx, lds, s)
                      : launch_fgi_stage_k<STAGE, false>(a, f, step_idx, lds, s);
}

hipError_t launch_fgi_stage(int stage, const StageArgs& a, const FgiArgs& f, int step_idx, hipStream_t s)
{
    if (a.tv.ntl == 0) return hipSuccess;
    const size_t lds = fgi_lds_bytes(a.tv.tile_max, f.ng);
    switch (stage) {
    case 0: return launch_fgi_stage_t<0>(a, f, step_idx, lds, s);
    case 1: return launch_fgi_stage_t<1>(a, f, step_idx, lds, s);
    case 2: return launch_fgi_stage_t<2>(a, f, step_idx, lds, s);
    default: return launch_fgi_stage_t<3>(a, f, step_idx, lds, s);
    }
}

hipError_t launch_ai_family(const StageArgs& a, const FgiArgs& f, const d4* uin, const d4* src, int mode, int step_idx,
                            d4* out, hipStream_t s)
{
    if (f.n == 0) return hipSuccess;
    k_ai_family<<<dim3((unsigned)f.n), dim3(kAiThreads), 0, s>>>(a.tv, f, a.mat, a.bc, a.mats, a.bcs, uin, src, mode,
                                                                a.steps, step_idx, out);
    return hipGetLastError();
}

hipError_t launch_fgi_gauss(const FgiArgs& f, const d4* u, int step_idx, hipStream_t s)
{
    const int64_t m = f.n * f.ng;
    if (m == 0) return hipSuccess;
    k_fgi_gauss<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s>>>(f, u, step_idx);
    return hipGetLastError();
}

hipError_t launch_fgi_forces(const StageArgs& a, const FgiArgs& f, const d4* u, d4* F, double* elast, hipStream_t s)
{
    if (a.tv.ntl == 0) return hipSuccess;
    const size_t lds = fgi_lds_bytes(a.tv.tile_max, f.ng);
    const int kind = fgi_family_kind(f);
    hipError_t e = kind == 2 ? set_lds<k_fgi_forces<2>>(lds) : kind == 3 ? set_lds<k_fgi_forces<3>>(lds)
                                                                         : set_lds<k_fgi_forces<0>>(lds);
    if (e != hipSuccess) return e;
    if (kind == 2) k_fgi_forces<2><<<dim3((unsigned)a.tv.ntl), dim3(64 * kFgiWaves), lds, s>>>(a, f, u, F, elast);
    else if (kind == 3) k_fgi_forces<3><<<dim3((unsigned)a.tv.ntl), dim3(64 * kFgiWaves), lds, s>>>(a, f, u, F, elast);
    else k_fgi_forces<0><<<dim3((unsigned)a.tv.ntl), dim3(64 * kFgiWaves), lds, s>>>(a, f, u, F, elast);
    return hipGetLastError();
}

hipError_t launch_fgi_pairs(const TileView& tv, const FgiArgs& f, bool fill, int64_t* cnt, int64_t* ent,
                            uint2* pairs, hipStream_t s)
{
    if (tv.ntl == 0) return hipSuccess;
    if (fill)
        k_fgi_pairs<true><<<dim3((unsigned)tv.ntl), dim3(64 * kFgiWaves), 0, s>>>(tv, f, cnt, ent, pairs);
    else
        k_fgi_pairs<false><<<dim3((unsigned)tv.ntl), dim3(64 * kFgiWaves), 0, s>>>(tv, f, cnt, ent, pairs);
    return hipGetLastError();
}

hipError_t launch_fly_corr(const TileView& tv, const d4* u, uint8_t* corr, int* clamped, hipStream_t s)
{
    if (tv.ntl == 0) return hipSuccess;
    const size_t lds = lds_bytes(tv.halo_max, tv.tile_max, false, false, true);
    hipError_t e = set_lds<k_fly_corr>(lds);
    if (e != hipSuccess) return e;
    k_fly_corr<<<dim3((unsigned)tv.ntl), dim3(256), lds, s>>>(tv, u, corr, clamped);
    return hipGetLastError();
}

hipError_t launch_tangent_coef(const TileView& tv, double* tc64, float* tc32, int16_t* tc16, hipStream_t s)
{
    if (tv.ntl == 0) return hipSuccess;
    k_tangent_coef<<<dim3((unsigned)tv.ntl), dim3(256), 0, s>>>(tv, tc64, tc32, tc16);
    return hipGetLastError();
}

hipError_t launch_tangent_diag(const TileView& tv, const d4* u, d4* y, hipStream_t s, d4* off)
{
    if (tv.ntl == 0) return hipSuccess;
    const size_t lds = stage_lds_bytes(tv, !tv.uniform_mat, false, false);
    hipError_t e = launch_gather_u<OpD64>(tv, u, y, nullptr, lds, s);
    if (e != hipSuccess || !off) return e;
    return launch_gather_u<OpD64X>(tv, u, off, nullptr, lds, s);
}

hipError_t launch_tangent(int precision, const TileView& tv, const d4* u, d4* y, hipStream_t s, double* dot)
{
    if (tv.ntl == 0) return hipSuccess;
    size_t lds = stage_lds_bytes(tv, !tv.uniform_mat, false, false);  // FP64 records in every precision
    if (precision == PDM_FP64) return launch_gather_u<OpT64>(tv, u, y, nullptr, lds, s, dot);
    if (precision == PDM_FP32C2 && tv.wlo) return launch_gather_u<OpT32C>(tv, u, y, nullptr, lds, s, dot);
    return launch_gather_u<OpT32>(tv, u, y, nullptr, lds, s, dot);
}

}  // namespace pdm
