// pdm_tiles.hip -- spatial tiling of the family CSR for LDS-staged gathers.
//
// Nodes are binned into tiles of tile_edge x dx along each axis and renumbered
// tile by tile (within a tile: reference id order).  Each row of the internal
// CSR is the reference family mapped to internal ids and sorted.  For every
// tile a halo list holds the internal ids of the tile's nodes and all their
// neighbours, sorted; a family entry stores the uint16 position of its
// neighbour in that list.  CCI rows keep only the entries that carry a force (no
// cut, beta > 0).  Rows are padded to a multiple of kRowAlign entries (padding
// entries point at the tile's padding record with a zero coefficient) so that a
// lane can load a whole aligned block of slots and coefficients at once.
// The halo of a tile lies in the (2R+1)^3 block of
// tiles around it, R = ceil(reach / edge) per axis, reach = delta + dx/2 plus
// a margin: an accepted pair has dmin < delta + 1e-4 dx, and dmin is at least
// the axis distance minus half a cell minus the face tolerance.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pdm_geom.h"
#include "pdm_internal.h"

namespace pdm {
namespace {

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            h->err = std::string("HIP error ") + hipGetErrorString(e_) + " at " #x; \
            return PDM_E_HIP;                                                      \
        }                                                                          \
    } while (0)

constexpr int kMaxNbr = 343;  // (2R+1)^3 for R <= 3

struct TileGrid {
    int64_t nt[3];
    int R[3];
};

// Internal node arrays from the reference-order ones.
__global__ void k_permute_nodes(const int32_t* __restrict__ perm, const d4* __restrict__ Xd,
                                const int32_t* __restrict__ mat, const int32_t* __restrict__ bc_ref,
                                const double* __restrict__ be_ref, const double* __restrict__ ecrit_of_mat, int64_t n,
                                d4* __restrict__ Xi, int32_t* __restrict__ mati, int32_t* __restrict__ bci,
                                double* __restrict__ bei)
{
    int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    int32_t i = perm[a];
    d4 x = Xd[i];
    int m = mat[i];
    x.w = ecrit_of_mat[m];
    Xi[a] = x;
    mati[a] = m;
    bci[a] = bc_ref[i];
    bei[a] = be_ref[i];
}

constexpr uint32_t kPadEntry = 0xFFFFFFFFu;  // row padding (sorts last)

// packed[e'] = internal col << 1 | cut, rows in internal order, padded with kPadEntry.
__global__ void k_rows_internal(const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                                const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                const uint8_t* __restrict__ cut, const int64_t* __restrict__ rpi, int64_t n,
                                uint32_t* __restrict__ packed)
{
    int64_t a = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x % 64;
    if (a >= n) return;
    int32_t i = perm[a];
    int64_t b = rp[i], len = rp[i + 1] - b, o = rpi[a], plen = rpi[a + 1] - o;
    for (int64_t k = lane; k < plen; k += 64)
        packed[o + k] = (k < len) ? (((uint32_t)iperm[col[b + k]] << 1) | (uint32_t)cut[b + k]) : kPadEntry;
}

__global__ void k_sort_rows_u32(const int64_t* __restrict__ rpi, uint32_t* __restrict__ v, int64_t n)
{
    __shared__ uint32_t sh[kMaxFamily];
    int64_t a = blockIdx.x;
    if (a >= n) return;
    int64_t b = rpi[a];
    int len = (int)(rpi[a + 1] - b);
    int P = 1;
    while (P < len) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) sh[t] = (t < len) ? v[b + t] : 0xFFFFFFFFu;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int t = threadIdx.x; t < P; t += blockDim.x) {
                int p = t ^ jj;
                if (p > t) {
                    bool up = ((t & k) == 0);
                    uint32_t va = sh[t], vb = sh[p];
                    if ((va > vb) == up) {
                        sh[t] = vb;
                        sh[p] = va;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < len; t += blockDim.x) v[b + t] = sh[t];
}

// beta of the entry (i, j) from the internal positions, as k_coef evaluates it.
__device__ __forceinline__ double entry_beta(const d4& xi, const d4& xj, double horizon, double dx, bool original)
{
    double x1 = xj.x - xi.x, x2 = xj.y - xi.y, x3 = xj.z - xi.z;
    double r = sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    return g_beta(r, horizon, dx, original);
}

// CCI rows keep only the entries that carry a force.  A cut bond and an entry whose beta
// is 0 (r >= delta + dx/2) contribute exactly 0 to the force, the elastic energy and the
// tangent in every state (CalcBondForceCCI, pdDataManager2.cpp:1293-1351, multiplies its
// force and energy by beta), so the gathers skip them; the body force still counts every
// unbroken entry, so nb[a] holds that count from the full row.  One wavefront per row;
// the kept entries move to the front of the row in order (every lane has read its entry
// before the wavefront writes, and a write never passes the position being read).
__global__ void k_live_rows(const int64_t* __restrict__ rpi, uint32_t* __restrict__ packed,
                            const d4* __restrict__ Xi, double horizon, double dx, bool original, int64_t n,
                            int32_t* __restrict__ nb, int64_t* __restrict__ live)
{
    const int64_t a = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (a >= n) return;
    const int64_t e0 = rpi[a], e1 = rpi[a + 1];
    const d4 xi = Xi[a];
    int cnt = 0;
    int64_t w = e0;
    for (int64_t b = e0; b < e1; b += 64) {
        const int64_t e = b + lane;
        const uint32_t pk = (e < e1) ? packed[e] : kPadEntry;
        const bool uncut = pk != kPadEntry && (pk & 1u) == 0;
        const bool keep = uncut && entry_beta(xi, Xi[pk >> 1], horizon, dx, original) != 0.0;
        cnt += uncut ? 1 : 0;
        const uint64_t m = __ballot(keep);
        if (keep) packed[w + __popcll(m & ((1ull << lane) - 1ull))] = pk;
        w += __popcll(m);
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (lane == 0) {
        nb[a] = cnt;
        live[a] = w - e0;
    }
}

// dst rows (row pointers rpo, padded) from the first live[a] entries of the src rows.
__global__ void k_copy_rows(const int64_t* __restrict__ rps, const uint32_t* __restrict__ src,
                            const int64_t* __restrict__ rpo, const int64_t* __restrict__ live, int64_t n,
                            uint32_t* __restrict__ dst)
{
    const int64_t a = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (a >= n) return;
    const int64_t s0 = rps[a], o0 = rpo[a], len = rpo[a + 1] - o0, lv = live[a];
    for (int64_t k = lane; k < len; k += 64) dst[o0 + k] = (k < lv) ? src[s0 + k] : kPadEntry;
}

struct HaloArgs {
    const int32_t* tiles;
    const int64_t* tile_start;
    const int32_t* tileid;   // internal node -> tile id
    const int64_t* rpi;
    const uint32_t* packed;
    int64_t ntl;
    TileGrid g;
    // pass 1
    int64_t* halo_len;
    // pass 2
    const int64_t* halo_ptr;
    int32_t* halo;
    uint16_t* slot;
    uint16_t* self_slot;
    int* error;
};

// One workgroup per tile.  WRITE=false: halo length; WRITE=true: halo list and slots.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_halo(HaloArgs a)
{
    __shared__ int32_t nb_tile[kMaxNbr];
    __shared__ int32_t nb_off[kMaxNbr + 1];
    __shared__ int16_t grid_to_nb[kMaxNbr];
    __shared__ uint8_t flag[kMaxFlags];
    __shared__ uint16_t pre[kMaxFlags];
    __shared__ int32_t tsum[256];
    __shared__ int bad;
    const int64_t k = blockIdx.x;
    const int32_t tid = a.tiles[k];
    const int64_t ntx = a.g.nt[0], nty = a.g.nt[1], ntz = a.g.nt[2];
    const int64_t tx = tid % ntx, ty = (tid / ntx) % nty, tz = tid / (ntx * nty);
    const int Rx = a.g.R[0], Ry = a.g.R[1], Rz = a.g.R[2];
    const int gx = 2 * Rx + 1, gy = 2 * Ry + 1, gz = 2 * Rz + 1;
    if (threadIdx.x == 0) {
        bad = 0;
        int cnt = 0;
        int32_t off = 0;
        for (int dz = -Rz; dz <= Rz; ++dz)
            for (int dy = -Ry; dy <= Ry; ++dy)
                for (int dx = -Rx; dx <= Rx; ++dx) {
                    int g = (dx + Rx) + gx * ((dy + Ry) + gy * (dz + Rz));
                    int64_t x = tx + dx, y = ty + dy, z = tz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= ntx || y >= nty || z >= ntz) {
                        grid_to_nb[g] = -1;
                        continue;
                    }
                    int32_t t2 = (int32_t)(x + ntx * (y + nty * z));
                    grid_to_nb[g] = (int16_t)cnt;
                    nb_tile[cnt] = t2;
                    nb_off[cnt] = off;
                    off += (int32_t)(a.tile_start[t2 + 1] - a.tile_start[t2]);
                    ++cnt;
                }
        nb_off[cnt] = off;
        if (off > kMaxFlags) bad = 1;
        tsum[0] = cnt;
    }
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) atomicExch(a.error, 1);
        return;
    }
    const int nnb = tsum[0];
    const int F = nb_off[nnb];
    __syncthreads();
    for (int f = threadIdx.x; f < F; f += blockDim.x) flag[f] = 0;
    __syncthreads();
    const int64_t a0 = a.tile_start[tid], a1 = a.tile_start[tid + 1];
    const int own = grid_to_nb[Rx + gx * (Ry + gy * Rz)];
    for (int64_t q = a0 + threadIdx.x; q < a1; q += blockDim.x) flag[nb_off[own] + (int)(q - a0)] = 1;
    const int64_t e0 = a.rpi[a0], e1 = a.rpi[a1];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const uint32_t pk = a.packed[e];
        if (pk == kPadEntry) continue;
        int32_t c = (int32_t)(pk >> 1);
        int32_t tj = a.tileid[c];
        int64_t jx = tj % ntx, jy = (tj / ntx) % nty, jz = tj / (ntx * nty);
        int64_t dx = jx - tx, dy = jy - ty, dz = jz - tz;
        if (dx < -Rx || dx > Rx || dy < -Ry || dy > Ry || dz < -Rz || dz > Rz) {
            atomicExch(a.error, 2);
            continue;
        }
        int g = (int)((dx + Rx) + gx * ((dy + Ry) + gy * (dz + Rz)));
        int nbi = grid_to_nb[g];
        flag[nb_off[nbi] + (int)(c - a.tile_start[tj])] = 1;
    }
    __syncthreads();
    // block exclusive scan of the flags
    const int per = (F + (int)blockDim.x - 1) / (int)blockDim.x;
    const int f0 = threadIdx.x * per;
    int loc = 0;
    for (int f = f0; f < f0 + per && f < F; ++f) loc += flag[f];
    tsum[threadIdx.x] = loc;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        int add = (threadIdx.x >= (unsigned)off) ? tsum[threadIdx.x - off] : 0;
        __syncthreads();
        tsum[threadIdx.x] += add;
        __syncthreads();
    }
    int run = tsum[threadIdx.x] - loc;
    for (int f = f0; f < f0 + per && f < F; ++f) {
        pre[f] = (uint16_t)run;
        run += flag[f];
    }
    const int H = tsum[blockDim.x - 1];
    __syncthreads();
    if (!WRITE) {
        if (threadIdx.x == 0) a.halo_len[k] = H;
        return;
    }
    const int64_t hp = a.halo_ptr[k];
    for (int nbi = 0; nbi < nnb; ++nbi) {
        int64_t s0 = a.tile_start[nb_tile[nbi]];
        int cnt = nb_off[nbi + 1] - nb_off[nbi];
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
            int f = nb_off[nbi] + q;
            if (flag[f]) a.halo[hp + pre[f]] = (int32_t)(s0 + q);
        }
    }
    for (int64_t q = a0 + threadIdx.x; q < a1; q += blockDim.x) a.self_slot[q] = pre[nb_off[own] + (int)(q - a0)];
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const uint32_t pk = a.packed[e];
        if (pk == kPadEntry) {
            a.slot[e] = (uint16_t)H;  // the padding record
            continue;
        }
        int32_t c = (int32_t)(pk >> 1);
        if (pk & 1u) {  // a cut bond (fixed Gaussian rows keep them) reads the padding record
            a.slot[e] = (uint16_t)H;
            continue;
        }
        int32_t tj = a.tileid[c];
        int64_t jx = tj % ntx, jy = (tj / ntx) % nty, jz = tj / (ntx * nty);
        int g = (int)((jx - tx + Rx) + gx * ((jy - ty + Ry) + gy * (jz - tz + Rz)));
        if (g < 0 || g >= gx * gy * gz) continue;
        int nbi = grid_to_nb[g];
        a.slot[e] = pre[nb_off[nbi] + (int)(c - a.tile_start[tj])];
    }
}

struct CoefArgs {
    const d4* Xi;
    const int64_t* rpi;
    const uint32_t* packed;
    const int32_t* mati;
    const double* bei;
    double* coef64;
    float* coef32;
    BK* bk;
    int32_t* nb;             // unbroken entries per row (null: counted elsewhere)
    int64_t n;
    double horizon, dx, vcube;
    bool original_cci;
    bool fgi;                // fixed Gaussian integration: coef64 = bdryEff_ij only
    double spring[kMaxMat];
};

// coef = bdryEff_ij * dx^3 * beta * c_ij: the leading factors of CalcBondForceCCI's force
// expression (pdDataManager2.cpp:1329-1331) in the same association order.  Under fixed
// Gaussian integration beta and the weights belong to each Gauss point, so coef64 holds
// bdryEff_ij alone (0 for a cut bond).
__global__ void k_coef(CoefArgs a)
{
    int64_t i = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    int lane = threadIdx.x % 64;
    if (i >= a.n) return;
    d4 xi = a.Xi[i];
    double bi = a.bei[i];
    double ci = a.spring[a.mati[i]];
    if (lane == 0) a.bk[i] = BK{(float)bi, (float)(0.5 * a.vcube * ci)};
    int cnt = 0;
    for (int64_t e = a.rpi[i] + lane; e < a.rpi[i + 1]; e += 64) {
        uint32_t pk = a.packed[e];
        if (pk == kPadEntry) {
            a.coef64[e] = 0.0;
            a.coef32[e] = 0.0f;
            continue;
        }
        int32_t j = (int32_t)(pk >> 1);
        bool cut = (pk & 1u) != 0;
        d4 xj = a.Xi[j];
        double cj = a.spring[a.mati[j]];
        double spring = (cj < ci) ? cj : ci;
        double bij = 0.5 * (bi + a.bei[j]);
        double beta = entry_beta(xi, xj, a.horizon, a.dx, a.original_cci);
        double c = cut ? 0.0 : bij * a.vcube * beta * spring;
        a.coef64[e] = cut ? 0.0 : (a.fgi ? bij : c);
        a.coef32[e] = (float)c;
        cnt += cut ? 0 : 1;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (lane == 0 && a.nb) a.nb[i] = cnt;
}

// LDS-bank-aware order of each row (one thread per row).  With 16 lanes per family the
// FP32 gathers read a row's entries 0..127 as one block of 8 per lane and the rest in
// blocks of 4 per lane (64 entries per block round).  In one LDS read, 8 consecutive lanes
// read the positions of one group: 64 b + 8 j + a (j = 0..7) below 128, 128 + 32 b + 4 j + a
// (j = 0..7) above; a 48-byte record s starts in bank group 3 s mod 16.  Entries sorted by
// slot mod 16 are dealt to the positions ordered by (j, group), so equal residues land in
// different groups and the 8 records one group reads fall in distinct banks.  The order
// within a row is otherwise free: every entry, padding included, carries its own weight.
__global__ void k_bank_order(const int64_t* __restrict__ rpi, uint16_t* __restrict__ slot,
                             uint32_t* __restrict__ packed, uint16_t* __restrict__ tslot,
                             uint32_t* __restrict__ tpacked, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t e0 = rpi[i];
    const int L = (int)(rpi[i + 1] - e0);
    int start[16] = {0};
    for (int p = 0; p < L; ++p) start[slot[e0 + p] & 15]++;
    for (int r = 0, run = 0; r < 16; ++r) {
        const int c = start[r];
        start[r] = run;
        run += c;
    }
    for (int p = 0; p < L; ++p) {
        const int q = start[slot[e0 + p] & 15]++;
        tslot[e0 + q] = slot[e0 + p];
        tpacked[e0 + q] = packed[e0 + p];
    }
    int q = 0;
    for (int j = 0; j < 8; ++j) {
        for (int b = 0; b < 2; ++b)
            for (int a = 0; a < 8; ++a) {
                const int p = 64 * b + 8 * j + a;
                if (p >= L) continue;
                slot[e0 + p] = tslot[e0 + q];
                packed[e0 + p] = tpacked[e0 + q];
                ++q;
            }
        for (int b = 0; 128 + 32 * b < L; ++b)
            for (int a = 0; a < 4; ++a) {
                const int p = 128 + 32 * b + 4 * j + a;
                if (p >= L) continue;
                slot[e0 + p] = tslot[e0 + q];
                packed[e0 + p] = tpacked[e0 + q];
                ++q;
            }
    }
}

template <class T>
__global__ void k_copy_add(T* __restrict__ dst, const T* __restrict__ src, int64_t n, T off)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + off;
}

__global__ void k_fill_i32(int32_t* __restrict__ dst, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

}  // namespace

hipError_t copy_add_i64(int64_t* dst, const int64_t* src, int64_t n, int64_t off, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_copy_add<int64_t><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(dst, src, n, off);
    return hipGetLastError();
}

hipError_t copy_add_i32(int32_t* dst, const int32_t* src, int64_t n, int32_t off, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_copy_add<int32_t><<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(dst, src, n, off);
    return hipGetLastError();
}

hipError_t fill_i32(int32_t* dst, int64_t n, int32_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_fill_i32<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(dst, n, v);
    return hipGetLastError();
}

TileView tile_view(const pdm_handle* h)
{
    TileView v{};
    v.tiles = h->tiles;
    v.tile_start = h->tile_start;
    v.halo_ptr = h->halo_ptr;
    v.halo = h->halo;
    v.self_slot = h->self_slot;
    v.rpi = h->rpi;
    v.slot = h->slot;
    v.coef64 = h->coef64;
    v.coef32 = h->coef32;
    v.bk = h->bk;
    v.corr = h->corr;
    v.X = h->Xi;
    v.tref = h->tref;
    v.n = h->n;
    v.ntl = h->ntl;
    v.halo_max = h->halo_max;
    v.tile_max = h->tile_max;
    v.xcd_remap = 1;
    v.waves = h->waves;
    v.fam_per_wave = h->fam_per_wave;
    v.uniform_mat = h->uniform_mat;
    v.specs = h->specs_d;
    v.tile_spec = h->tile_spec_d;
    return v;
}

int build_tiles(pdm_handle* h)
{
    const Deck& d = h->deck;
    hipStream_t s = h->stream;
    const int64_t n = h->n;
    // tile grid over the node bounding box; bins start half a spacing below the lowest node
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) { lo[c] = 1e300; hi[c] = -1e300; }
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], h->X[3 * i + c]);
            hi[c] = std::max(hi[c], h->X[3 * i + c]);
        }
    TileGrid g{};
    double edge[3];
    const double reach = d.horizon + 0.5 * d.dx + 0.01 * d.dx;
    for (int c = 0; c < 3; ++c) {
        if (n == 0) { lo[c] = hi[c] = 0.0; }
        lo[c] -= 0.5 * d.dx;
        edge[c] = h->tile_edge[c] * d.dx;
        g.nt[c] = (int64_t)std::floor((hi[c] - lo[c]) / edge[c]) + 1;
        g.R[c] = (int)std::ceil(reach / edge[c]);
        if (g.R[c] > 3) {
            h->err = "tile edge too small for the horizon";
            return PDM_E_LIMIT;
        }
    }
    const int64_t ntiles = g.nt[0] * g.nt[1] * g.nt[2];
    // every halo node lies within (R + 1/2) tile edges of its tile's centre along each axis;
    // the FP32 records hold positions as int32 fixed point with 2^30 covering that reach
    double reach_c = 0.0;
    for (int c = 0; c < 3; ++c) reach_c = std::max(reach_c, (g.R[c] + 0.5) * edge[c]);
    h->fscale = std::ldexp(1.0, 30) / (reach_c + d.dx);
    // the FP32 records round X fscale to an integer in FP64: it must stay below 2^52
    for (int c = 0; c < 3; ++c)
        if (std::max(std::fabs(lo[c]), std::fabs(hi[c])) * h->fscale > std::ldexp(1.0, 52)) {
            h->err = "node coordinates too far from the origin for the fixed-point records";
            return PDM_E_LIMIT;
        }
    if ((double)g.nt[0] * (double)g.nt[1] * (double)g.nt[2] > 2.0e9) {
        h->err = "tile grid too large";
        return PDM_E_LIMIT;
    }
    // counting sort of nodes by tile (stable: reference id order within a tile)
    std::vector<int32_t> tid_ref(n);
    std::vector<int64_t> tstart(ntiles + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        int64_t q[3];
        for (int c = 0; c < 3; ++c) {
            q[c] = (int64_t)std::floor((h->X[3 * i + c] - lo[c]) / edge[c]);
            q[c] = std::min(std::max<int64_t>(q[c], 0), g.nt[c] - 1);
        }
        tid_ref[i] = (int32_t)(q[0] + g.nt[0] * (q[1] + g.nt[1] * q[2]));
        tstart[tid_ref[i] + 1]++;
    }
    for (int64_t t = 0; t < ntiles; ++t) tstart[t + 1] += tstart[t];
    std::vector<int64_t> cursor(tstart.begin(), tstart.end() - 1);
    h->perm.assign(n, 0);
    std::vector<int32_t> iperm(n), tileid(n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t a = cursor[tid_ref[i]]++;
        h->perm[a] = (int32_t)i;
        iperm[i] = (int32_t)a;
        tileid[a] = tid_ref[i];
    }
    std::vector<int32_t> tiles;
    int tmax = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
        int64_t cnt = tstart[t + 1] - tstart[t];
        if (cnt > 0) tiles.push_back((int32_t)t);
        tmax = std::max<int>(tmax, (int)cnt);
    }
    if (tmax > kMaxTileNodes) {
        h->err = "a tile holds more than " + std::to_string(kMaxTileNodes) + " nodes";
        return PDM_E_LIMIT;
    }
    h->ntiles = ntiles;
    h->ntl = (int64_t)tiles.size();
    h->tile_max = tmax;
    // internal row pointers
    std::vector<int64_t> rph(n + 1), rpih(n + 1, 0);
    HIPCHK(hipMemcpy(rph.data(), h->rp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    for (int64_t a = 0; a < n; ++a) {
        const int64_t len = rph[h->perm[a] + 1] - rph[h->perm[a]];
        rpih[a + 1] = rpih[a] + (len + kRowAlign - 1) / kRowAlign * kRowAlign;
    }
    const int64_t Ei = rpih[n];
    h->Ei = Ei;
    // one ecrit and one spring constant across the materials?
    h->uniform_mat = 1;
    for (size_t m = 1; m < d.mats.size(); ++m)
        if (d.mats[m].ecrit != d.mats[0].ecrit || h->spring[m] != h->spring[0]) h->uniform_mat = 0;

    int32_t *ipermd = nullptr, *tileidd = nullptr, *bcref = nullptr;
    double *ecm = nullptr, *bei = nullptr;
    uint32_t* packed = nullptr;
    int64_t* halo_len = nullptr;
    int* err = nullptr;
    // temporaries are released on every return path (after the stream has drained)
    DevTemps temps(s);
    temps.add(&ipermd);
    temps.add(&tileidd);
    temps.add(&bcref);
    temps.add(&ecm);
    temps.add(&bei);
    temps.add(&packed);
    temps.add(&halo_len);
    temps.add(&err);
    const int64_t n1 = std::max<int64_t>(n, 1), E1 = std::max<int64_t>(Ei, 1);
    HIPCHK(hipMalloc(&h->permd, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&ipermd, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&tileidd, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&bcref, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&ecm, sizeof(double) * kMaxMat));
    HIPCHK(hipMalloc(&bei, sizeof(double) * n1));
    HIPCHK(hipMalloc(&h->Xi, sizeof(d4) * n1));
    HIPCHK(hipMalloc(&h->mati, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&h->bci, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&h->nb, sizeof(int32_t) * n1));
    HIPCHK(hipMalloc(&h->rpi, sizeof(int64_t) * (n + 1)));
    HIPCHK(hipMalloc(&h->tile_start, sizeof(int64_t) * (ntiles + 1)));
    HIPCHK(hipMalloc(&h->tiles, sizeof(int32_t) * std::max<int64_t>(h->ntl, 1)));
    HIPCHK(hipMalloc(&h->tref, sizeof(d4) * std::max<int64_t>(h->ntl, 1)));
    HIPCHK(hipMalloc(&h->halo_ptr, sizeof(int64_t) * (h->ntl + 1)));
    HIPCHK(hipMalloc(&halo_len, sizeof(int64_t) * std::max<int64_t>(h->ntl, 1)));
    HIPCHK(hipMalloc(&h->self_slot, sizeof(uint16_t) * n1));
    HIPCHK(hipMalloc(&packed, sizeof(uint32_t) * E1));
    HIPCHK(hipMalloc(&h->bk, sizeof(BK) * n1));
    HIPCHK(hipMalloc(&err, sizeof(int)));
    double ech[kMaxMat] = {0};
    for (size_t m = 0; m < d.mats.size(); ++m) ech[m] = d.mats[m].ecrit;
    HIPCHK(hipMemcpyAsync(h->permd, h->perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ipermd, iperm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(tileidd, tileid.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(bcref, h->bc.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ecm, ech, sizeof(ech), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(h->rpi, rpih.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(h->tile_start, tstart.data(), sizeof(int64_t) * (ntiles + 1), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(h->tiles, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice, s));
    // tile centres in fixed-point units, rint(c fscale): the reference the FP32 halo
    // records subtract from each node's rint(X fscale)
    std::vector<d4> trefh(tiles.size());
    for (size_t k = 0; k < tiles.size(); ++k) {
        const int64_t t = tiles[k];
        const int64_t q[3] = {t % g.nt[0], (t / g.nt[0]) % g.nt[1], t / (g.nt[0] * g.nt[1])};
        double c[3];
        for (int a = 0; a < 3; ++a) c[a] = std::nearbyint((lo[a] + ((double)q[a] + 0.5) * edge[a]) * h->fscale);
        trefh[k] = d4{c[0], c[1], c[2], 0.0};
    }
    HIPCHK(hipMemcpy(h->tref, trefh.data(), sizeof(d4) * trefh.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(err, 0, sizeof(int), s));
    if (n) {
        k_permute_nodes<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
            h->permd, h->Xd, h->matd, bcref, h->be, ecm, n, h->Xi, h->mati, h->bci, bei);
        HIPCHK(hipGetLastError());
        unsigned nbw = (unsigned)((n + 3) / 4);
        k_rows_internal<<<dim3(nbw), dim3(256), 0, s>>>(h->permd, ipermd, h->rp, h->col, h->cut, h->rpi, n, packed);
        HIPCHK(hipGetLastError());
        k_sort_rows_u32<<<dim3((unsigned)n), dim3(256), 0, s>>>(h->rpi, packed, n);
        HIPCHK(hipGetLastError());
    }
    // CCI: rows of the entries that carry a force (fixed Gaussian integration weights each
    // Gauss point's beta itself and keeps every entry)
    const bool compact = d.cci_on && n > 0;
    h->Elive = h->E;
    if (compact) {
        int64_t *live = nullptr, *rpo = nullptr;
        uint32_t* packed2 = nullptr;
        struct Local {
            hipStream_t s;
            int64_t** a;
            int64_t** b;
            uint32_t** c;
            ~Local()
            {
                (void)hipStreamSynchronize(s);
                if (*a) (void)hipFree(*a);
                if (*b) (void)hipFree(*b);
                if (*c) (void)hipFree(*c);
            }
        } local{s, &live, &rpo, &packed2};
        HIPCHK(hipMalloc(&live, sizeof(int64_t) * n));
        k_live_rows<<<dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s>>>(h->rpi, packed, h->Xi, d.horizon, d.dx,
                                                                         !d.cci_mdf, n, h->nb, live);
        HIPCHK(hipGetLastError());
        std::vector<int64_t> lv(n);
        HIPCHK(hipMemcpyAsync(lv.data(), live, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<int64_t> rpo_h(n + 1, 0);
        for (int64_t a = 0; a < n; ++a)
            rpo_h[a + 1] = rpo_h[a] + (lv[a] + kRowAlign - 1) / kRowAlign * kRowAlign;
        const int64_t E2 = rpo_h[n];
        HIPCHK(hipMalloc(&rpo, sizeof(int64_t) * (n + 1)));
        HIPCHK(hipMalloc(&packed2, sizeof(uint32_t) * std::max<int64_t>(E2, 1)));
        HIPCHK(hipMemcpyAsync(rpo, rpo_h.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
        k_copy_rows<<<dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s>>>(h->rpi, packed, rpo, live, n, packed2);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h->rpi, rpo_h.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        std::swap(packed, packed2);  // the full rows are released with the other temporaries
        h->Ei = E2;
        h->Elive = 0;
        for (int64_t a = 0; a < n; ++a) h->Elive += lv[a];
    }
    const int64_t Ec = std::max<int64_t>(h->Ei, 1);
    HIPCHK(hipMalloc(&h->slot, sizeof(uint16_t) * Ec));
    HIPCHK(hipMalloc(&h->coef64, sizeof(double) * Ec));
    HIPCHK(hipMalloc(&h->coef32, sizeof(float) * Ec));
    HIPCHK(hipMalloc(&h->corr, Ec));
    HaloArgs ha{};
    ha.tiles = h->tiles;
    ha.tile_start = h->tile_start;
    ha.tileid = tileidd;
    ha.rpi = h->rpi;
    ha.packed = packed;
    ha.ntl = h->ntl;
    ha.g = g;
    ha.halo_len = halo_len;
    ha.error = err;
    if (h->ntl) {
        k_halo<false><<<dim3((unsigned)h->ntl), dim3(256), 0, s>>>(ha);
        HIPCHK(hipGetLastError());
    }
    int rc = scan_exclusive(h, halo_len, h->halo_ptr, h->ntl);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        h->err = herr == 1 ? "a tile neighbourhood holds more than " + std::to_string(kMaxFlags) + " nodes"
                           : "a family entry lies outside its tile neighbourhood";
        return PDM_E_LIMIT;
    }
    std::vector<int64_t> hp(h->ntl + 1);
    HIPCHK(hipMemcpy(hp.data(), h->halo_ptr, sizeof(int64_t) * (h->ntl + 1), hipMemcpyDeviceToHost));
    h->halo_total = hp[h->ntl];
    int64_t hmax = 0;
    for (int64_t k = 0; k < h->ntl; ++k) hmax = std::max(hmax, hp[k + 1] - hp[k]);
    if (hmax > kMaxHalo) {
        h->err = "a tile halo holds " + std::to_string(hmax) + " nodes, more than the LDS limit " +
                 std::to_string(kMaxHalo);
        return PDM_E_LIMIT;
    }
    h->halo_max = (int)hmax;
    {
        TileView tv = tile_view(h);
        size_t need = stage_lds_bytes(tv, !h->uniform_mat, true, false);  // FP64 records: the larger
        if (need > kLdsBytes) {
            h->err = "tile LDS footprint " + std::to_string(need) + " B exceeds " + std::to_string(kLdsBytes) +
                     " B; use a smaller tile_edge";
            return PDM_E_LIMIT;
        }
    }
    HIPCHK(hipMalloc(&h->halo, sizeof(int32_t) * std::max<int64_t>(h->halo_total, 1)));
    ha.halo_ptr = h->halo_ptr;
    ha.halo = h->halo;
    ha.slot = h->slot;
    ha.self_slot = h->self_slot;
    if (h->ntl) {
        k_halo<true><<<dim3((unsigned)h->ntl), dim3(256), 0, s>>>(ha);
        HIPCHK(hipGetLastError());
    }
    if (n && d.cci_on) {  // the fixed-Gaussian kernels read neighbours from global memory, in id order
        uint16_t* tslot = nullptr;
        uint32_t* tpacked = nullptr;
        HIPCHK(hipMalloc(&tslot, sizeof(uint16_t) * Ec));
        HIPCHK(hipMalloc(&tpacked, sizeof(uint32_t) * Ec));
        k_bank_order<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(h->rpi, h->slot, packed, tslot, tpacked, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipFree(tslot));
        HIPCHK(hipFree(tpacked));
    }
    CoefArgs ca{};
    ca.Xi = h->Xi;
    ca.rpi = h->rpi;
    ca.packed = packed;
    ca.mati = h->mati;
    ca.bei = bei;
    ca.coef64 = h->coef64;
    ca.coef32 = h->coef32;
    ca.bk = h->bk;
    ca.nb = compact ? nullptr : h->nb;  // compacted rows: nb was counted from the full rows
    ca.n = n;
    ca.horizon = d.horizon;
    ca.dx = d.dx;
    ca.vcube = d.dx * d.dx * d.dx;
    ca.original_cci = d.cci_on && !d.cci_mdf;
    ca.fgi = !d.cci_on;  // fixed Gaussian and adaptive integration
    for (size_t m = 0; m < d.mats.size() && m < (size_t)kMaxMat; ++m) ca.spring[m] = h->spring[m];
    if (n) {
        k_coef<<<dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s>>>(ca);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    if (!compact && n) {  // fixed Gaussian integration evaluates the uncut entries
        std::vector<int32_t> nbh(n);
        HIPCHK(hipMemcpy(nbh.data(), h->nb, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        h->Elive = 0;
        for (int64_t a = 0; a < n; ++a) h->Elive += nbh[a];
    }
    HIPCHK(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        h->err = "a family entry lies outside its tile neighbourhood";
        return PDM_E_LIMIT;
    }
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMalloc(&h->u[b], sizeof(d4) * n1));
        HIPCHK(hipMemsetAsync(h->u[b], 0, sizeof(d4) * n1, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return PDM_OK;
}

}  // namespace pdm
